// dem_diff.inl -- differentiable runs: reverse-mode derivatives of scene
// observables with respect to the per-scene contact laws.  Included at the
// end of dem_hip.hip (it uses the solver's kernels and device buffers).
//
// A run is the reference's pure-DEM loop (IntegrateDem,
// cuda_simulation_integrator.cpp:245-286): the first force evaluation
// (dt = 0), the half-step initialisation, then `steps` x (advance, forces,
// finish).  Observables per scene: the mean centre height z of its particles
// after the last step, and the plate load, the z force of one wall (the same
// wall index in every scene) on the particles, averaged over the steps.
//
// Forward: the run stores checkpoints (particle state and contact history)
// before step 1, 1 + K, 1 + 2K, ... .  Reverse: for each K-step segment,
// last to first, it restores the checkpoint, recomputes the segment while
// recording a tape (the state before every step, the contact lists with
// their history slots, the slot values each law read and the occupancy after
// expiry), then runs the reverse kernels step by step: finish + acceleration,
// particle-particle and particle-wall laws (dem_adjoint.hpp), component
// velocities and advance.  The history adjoint follows the slots: a slot
// that a contact acquired afresh takes no derivative from the older values.
// Particles never move rows during a differentiable run: the rows are put in
// cell order once before it starts, and neither the full scan nor the
// candidate-list rebuilds reorder them afterwards, so a row on the tape names
// the same particle throughout.  Candidate lists (when the solver has them)
// replace the full scan on the forward run and on every recomputed segment;
// they yield the full scan's contacts in its order, so the tape is the same.

namespace {

// Snapshot of one particle [field][n]: x3 v3 w3 q4 vh3 wh3 hq4 a3 al3
enum { P_X = 0, P_V = 3, P_W = 6, P_Q = 9, P_VH = 13, P_WH = 16, P_HQ = 19, P_A = 23, P_AL = 26, NSNAP = 29 };
// Adjoint of the carried state [field][n]: x3 q4 vh3 wh3 hq4 a3 al3
enum { L_X = 0, L_Q = 3, L_VH = 7, L_WH = 10, L_HQ = 13, L_A = 17, L_AL = 20, NLAM = 23 };
// Per-step scratch [field][n]: advanced q'(4), w'(3); adjoints of force (3),
// torque (3), q'(4), w'(3), vh(3), wh(3); component adjoint pos vel vh w wh (15)
enum { Z_Q = 0, Z_W = 4, Z_F = 7, Z_T = 10, Z_LQ = 13, Z_LW = 17, Z_LVH = 20, Z_LWH = 23, Z_B = 26, NZ = 41 };

__global__ void __launch_bounds__(BLOCK) k_snap(Soa S, double* out)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    if (p >= n) return;
    const int src[NSNAP] = {F_X,  F_Y,  F_Z,  F_VX,  F_VY,  F_VZ,  F_WX,  F_WY,  F_WZ,  F_QW,
                            F_QX, F_QY, F_QZ, F_VHX, F_VHY, F_VHZ, F_WHX, F_WHY, F_WHZ, F_HQW,
                            F_HQX, F_HQY, F_HQZ, F_AX, F_AY, F_AZ, F_AWX, F_AWY, F_AWZ};
#pragma unroll
    for (int k = 0; k < NSNAP; ++k) out[(long)k * n + p] = S.f(src[k], p);
}

// Step bookkeeping of a run: contact counts and history high-water marks.
__global__ void k_diff_count(const Scalars* sc, unsigned* wallCount, long long* pp, long long* pw, int* high,
                             int step)
{
    if (threadIdx.x != 0) return;
    pp[step] = sc->total_pairs;
    pw[step] = *wallCount;
    *wallCount = 0u;
    high[0] = max(high[0], sc->pp_high);
    high[1] = max(high[1], sc->body_high);
}

__device__ __forceinline__ V3 ld3(const double* a, int f, long n, long p)
{
    return mk(a[(long)f * n + p], a[(long)(f + 1) * n + p], a[(long)(f + 2) * n + p]);
}
__device__ __forceinline__ void st3(double* a, int f, long n, long p, V3 v)
{
    a[(long)f * n + p] = v.x;
    a[(long)(f + 1) * n + p] = v.y;
    a[(long)(f + 2) * n + p] = v.z;
}
__device__ __forceinline__ void ld4(const double* a, int f, long n, long p, double* q)
{
    for (int k = 0; k < 4; ++k) q[k] = a[(long)(f + k) * n + p];
}
__device__ __forceinline__ void st4(double* a, int f, long n, long p, const double* q)
{
    for (int k = 0; k < 4; ++k) a[(long)(f + k) * n + p] = q[k];
}

// Reverse of finish (or, for the first evaluation, of the half-step
// initialisation) and of the acceleration; rebuilds the contact records of
// the step from the state before it.
__global__ void __launch_bounds__(BLOCK) k_adj_prep(Soa S, const double* pre, const double* post, double dt,
                                                    bool initial, Recs rec, double* lam, double* z)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    if (p >= n) return;
    const int status = S.k(I_STATUS, p);
    V3 x = ld3(pre, P_X, n, p), v = ld3(pre, P_V, n, p), w = ld3(pre, P_W, n, p);
    const V3 vh = ld3(pre, P_VH, n, p), wh = ld3(pre, P_WH, n, p), a = ld3(pre, P_A, n, p),
             al = ld3(pre, P_AL, n, p);
    double q[4], hq[4];
    ld4(pre, P_Q, n, p, q);
    ld4(pre, P_HQ, n, p, hq);
    // advance (advancePositionKernel)
    if (!initial && status != -1) {
        x = mk(x.x + vh.x * dt, x.y + vh.y * dt, x.z + vh.z * dt);
        for (int k = 0; k < 4; ++k) q[k] = q[k] + hq[k] * dt;
        quat_unit(q[0], q[1], q[2], q[3]);
        v = mk(vh.x + a.x * (0.5 * dt), vh.y + a.y * (0.5 * dt), vh.z + a.z * (0.5 * dt));
        w = mk(wh.x + al.x * (0.5 * dt), wh.y + al.y * (0.5 * dt), wh.z + al.z * (0.5 * dt));
    }
    {
        Rec r;
        r.px = x.x; r.py = x.y; r.pz = x.z;
        r.vx = v.x; r.vy = v.y; r.vz = v.z;
        r.hx = vh.x; r.hy = vh.y; r.hz = vh.z;
        const V3 ww = rot(q[0], q[1], q[2], q[3], w), wwh = rot(q[0], q[1], q[2], q[3], wh);
        r.wx = ww.x; r.wy = ww.y; r.wz = ww.z;
        r.ux = wwh.x; r.uy = wwh.y; r.uz = wwh.z;
        r.mass = S.f(F_MASS, p);
        rec.main[p] = r;
        RecAux x;
        x.r = S.f(F_CR, p);
        x.status = status;
        x.scene = S.k(I_SCENE, p);
        rec.aux[p] = x;
    }
    st4(z, Z_Q, n, p, q);
    st3(z, Z_W, n, p, w);
    // post-step values
    const V3 alp = ld3(post, P_AL, n, p);
    const V3 lvh = ld3(lam, L_VH, n, p), lwh = ld3(lam, L_WH, n, p), la = ld3(lam, L_A, n, p),
             lal = ld3(lam, L_AL, n, p);
    double lhq[4];
    ld4(lam, L_HQ, n, p, lhq);
    double lq[4] = {0.0, 0.0, 0.0, 0.0};
    V3 lw = zero3();
    V3 lwhTot = lwh;
    V3 laTot, lalTot;
    if (!initial) {
        // finish: vh' = vh + a' dt; wh' = wh + al' dt; hq' = g(N(q' + g(q', w') dt/2), wh')
        ld4(lam, L_Q, n, p, lq);
        const double h = 0.5 * dt;
        const V3 whp = mk(wh.x + alp.x * dt, wh.y + alp.y * dt, wh.z + alp.z * dt);
        double g1[4];
        quat_dot(q[0], q[1], q[2], q[3], w.x, w.y, w.z, g1[0], g1[1], g1[2], g1[3]);
        double qt[4] = {q[0] + g1[0] * h, q[1] + g1[1] * h, q[2] + g1[2] * h, q[3] + g1[3] * h};
        double qh[4] = {qt[0], qt[1], qt[2], qt[3]};
        quat_unit(qh[0], qh[1], qh[2], qh[3]);
        double qhb[4] = {0.0, 0.0, 0.0, 0.0}, qtb[4] = {0.0, 0.0, 0.0, 0.0};
        quat_dot_vjp(qh, whp, lhq, qhb, lwhTot);
        quat_unit_vjp(qt, qhb, qtb);
        double qtbh[4];
        for (int k = 0; k < 4; ++k) {
            lq[k] += qtb[k];
            qtbh[k] = qtb[k] * h;
        }
        quat_dot_vjp(q, w, qtbh, lq, lw);
        laTot = add(la, scl(lvh, dt));
        lalTot = add(lal, scl(lwhTot, dt));
        st3(z, Z_LVH, n, p, lvh);
    } else {
        // half-step initialisation with h = dt/2: vh0 = v0 + a0 h; wh0 = w0 + al0 h;
        // hq0 = g(N(q0 + g(q0, w0) h), wh0)   (q0, w0, v0: inputs, no derivative)
        const double h = 0.5 * dt;
        const V3 wh0 = ld3(post, P_WH, n, p);
        double g1[4];
        quat_dot(q[0], q[1], q[2], q[3], w.x, w.y, w.z, g1[0], g1[1], g1[2], g1[3]);
        double qh[4] = {q[0] + g1[0] * h, q[1] + g1[1] * h, q[2] + g1[2] * h, q[3] + g1[3] * h};
        quat_unit(qh[0], qh[1], qh[2], qh[3]);
        double qhb[4] = {0.0, 0.0, 0.0, 0.0};
        quat_dot_vjp(qh, wh0, lhq, qhb, lwhTot);
        laTot = add(la, scl(lvh, h));
        lalTot = add(lal, scl(lwhTot, h));
        st3(z, Z_LVH, n, p, zero3());
    }
    // acceleration: a' = F / m; al' = Euler(R(q')^T T, w', I).  The world
    // torque is recovered from al' by inverting the Euler equations.
    const double m = S.f(F_MASS, p);
    const V3 inertia = mk(S.f(F_IX, p), S.f(F_IY, p), S.f(F_IZ, p));
    const V3 tb = mk(alp.x * inertia.x + (inertia.z - inertia.y) * w.z * w.y,
                     alp.y * inertia.y + (inertia.x - inertia.z) * w.x * w.z,
                     alp.z * inertia.z + (inertia.y - inertia.x) * w.y * w.x);
    const V3 T = rot(q[0], q[1], q[2], q[3], tb);
    V3 lT = zero3();
    euler_vjp(q, T, w, inertia, lalTot, lT, lq, lw);
    st3(z, Z_F, n, p, scl(laTot, 1.0 / m));
    st3(z, Z_T, n, p, lT);
    st4(z, Z_LQ, n, p, lq);
    st3(z, Z_LW, n, p, lw);
    st3(z, Z_LWH, n, p, lwhTot);
    for (int k = 0; k < 15; ++k) z[(long)(Z_B + k) * n + p] = 0.0;
}

__device__ __forceinline__ void add_body_grad(double* z, long n, long p, const BodyGrad& g)
{
    const V3 v[5] = {g.pos, g.vel, g.vh, g.w, g.wh};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        unsafeAtomicAdd(&z[(long)(Z_B + 3 * k) * n + p], v[k].x);
        unsafeAtomicAdd(&z[(long)(Z_B + 3 * k + 1) * n + p], v[k].y);
        unsafeAtomicAdd(&z[(long)(Z_B + 3 * k + 2) * n + p], v[k].z);
    }
}

// Owner-side body adjoint of the contacts held by a wave: the contacts of
// one owner are contiguous in the tape, so each run of lanes with the same
// owner is summed with shuffles (suffix sums within the run) and the run's
// first lane adds the total; one set of atomics per owner instead of one
// per contact.
__device__ __forceinline__ void add_body_grad_owner(double* z, long n, long p, bool active, const BodyGrad& g)
{
    const int lane = threadIdx.x & 63;
    const int key = active ? (int)p : -1 - lane;  // idle lanes are runs of their own
    double v[15] = {g.pos.x, g.pos.y, g.pos.z, g.vel.x, g.vel.y, g.vel.z, g.vh.x, g.vh.y, g.vh.z,
                    g.w.x,   g.w.y,   g.w.z,   g.wh.x,  g.wh.y,  g.wh.z};
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int next = __shfl_down(key, o, 64);  // every lane takes part in the shuffle
        const bool same = lane + o < 64 && next == key;
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            const double t = __shfl_down(v[k], o, 64);
            if (same) v[k] += t;
        }
    }
    const int before = __shfl_up(key, 1, 64);
    if (active && (lane == 0 || before != key)) {
#pragma unroll
        for (int k = 0; k < 15; ++k) unsafeAtomicAdd(&z[(long)(Z_B + k) * n + p], v[k]);
    }
}

// Adds a material adjoint to the scene's accumulator: one atomic per wave
// when the wave's lanes share a scene.
__device__ __forceinline__ void scene_add(double* acc, int scene, bool active, const MatGrad& g)
{
    const int lo = wave_min(active ? scene : INT_MAX), hi = wave_max(active ? scene : INT_MIN);
    const double v[4] = {active ? g.mu : 0.0, active ? g.logN : 0.0, active ? g.logS : 0.0,
                         active ? g.kN : 0.0};
    if (lo == hi) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double s = wave_sum(v[k]);
            if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(&acc[lo * 8 + k], s);
        }
    } else if (active) {
        for (int k = 0; k < 4; ++k) unsafeAtomicAdd(&acc[scene * 8 + k], v[k]);
    }
}

// Reverse of the particle-particle law over the step's contacts.
__global__ void __launch_bounds__(BLOCK) k_adj_pairs(long count, const int2* pair, const double* prevTape,
                                                     const unsigned long long* occTape, Recs rec,
                                                     const Material* mats, bool first, long n, double* z,
                                                     double* lamH, double* grads)
{
    const long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    const bool active = e < count;
    MatGrad g{0.0, 0.0, 0.0, 0.0};
    BodyGrad Ib = zero_body_grad();
    long owner = 0;
    int scene = 0;
    if (active) {
        const int2 pr = pair[e];
        const long p = pr.x, q = pr.y >> SLOT_BITS;
        const int slot = pr.y & (PP_SLOTS - 1);
        const Body I = rec_body(rec, p), J = rec_body(rec, q);
        scene = rec.aux[p].scene;
        Material m = mats[scene];
        if (first) m.dt = 0.0;
        double* hb = lamH + (p * PP_SLOTS + slot) * 9;
        double hbv[9];
        for (int k = 0; k < 9; ++k) hbv[k] = hb[k];
        double prevb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        BodyGrad Jb = zero_body_grad();
        pp_law_vjp(I, J, m, prevTape + e * 9, ld3(z, Z_F, n, p), ld3(z, Z_T, n, p), hbv, Ib, Jb, prevb, g);
        const bool fresh = !((occTape[p] >> slot) & 1ull);
        for (int k = 0; k < 9; ++k) hb[k] = fresh ? 0.0 : prevb[k];
        add_body_grad(z, n, q, Jb);
        owner = p;
    }
    add_body_grad_owner(z, n, owner, active, Ib);
    scene_add(grads, scene, active, g);
}

// Reverse of the particle-wall law over the step's wall contacts; the plate
// load's adjoint enters as a seed on the plate wall's force.
__global__ void __launch_bounds__(BLOCK) k_adj_walls(long count, const int4* wall, const double* prevTape,
                                                     const unsigned* occTape, Recs rec, const Material* mats,
                                                     const SceneParams* sp, Walls W, bool first, int plate,
                                                     const double* plateSeed, long n, double* z, double* lamHW,
                                                     double* grads)
{
    const long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    const bool active = e < count;
    MatGrad g{0.0, 0.0, 0.0, 0.0};
    int scene = 0;
    if (active) {
        const int4 t = wall[e];
        const long p = t.x;
        const int w = t.y, slot = t.z;
        const Body I = rec_body(rec, p);
        scene = rec.aux[p].scene;
        Material m = mats[scene];
        if (first) m.dt = 0.0;
        const int gw = sp[scene].wall0 + w;
        const V3 outward = mk(W.nrm[3 * gw], W.nrm[3 * gw + 1], W.nrm[3 * gw + 2]);
        const V3 origin = mk(W.o[3 * gw], W.o[3 * gw + 1], W.o[3 * gw + 2]);
        V3 cfb = ld3(z, Z_F, n, p);
        if (w == plate && !first) cfb.z += plateSeed[scene];
        double* hb = lamHW + (p * PW_SLOTS + slot) * 9;
        double hbv[9];
        for (int k = 0; k < 9; ++k) hbv[k] = hb[k];
        double prevb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        BodyGrad Ib = zero_body_grad();
        pw_law_vjp(I, outward, origin, m, prevTape + e * 9, cfb, ld3(z, Z_T, n, p), hbv, Ib, prevb, g);
        const bool fresh = !((occTape[p] >> slot) & 1u);
        for (int k = 0; k < 9; ++k) hb[k] = fresh ? 0.0 : prevb[k];
        add_body_grad(z, n, p, Ib);
    }
    scene_add(grads + 4, scene, active, g);
}

// Component velocities (world angular velocities = R(q') w) and the advance,
// reversed: leaves the adjoint of the state before the step in lam.
__global__ void __launch_bounds__(BLOCK) k_adj_body(Soa S, const double* pre, double dt, double* lam,
                                                    const double* z)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    if (p >= n) return;
    double q[4], lq[4];
    ld4(z, Z_Q, n, p, q);
    ld4(z, Z_LQ, n, p, lq);
    const V3 w = ld3(z, Z_W, n, p), wh = ld3(pre, P_WH, n, p);
    V3 lw = ld3(z, Z_LW, n, p), lwh = ld3(z, Z_LWH, n, p), lvh = ld3(z, Z_LVH, n, p);
    const V3 bpos = ld3(z, Z_B, n, p), bvel = ld3(z, Z_B + 3, n, p), bvh = ld3(z, Z_B + 6, n, p),
             bw = ld3(z, Z_B + 9, n, p), bwh = ld3(z, Z_B + 12, n, p);
    const V3 lx = add(ld3(lam, L_X, n, p), bpos);
    const V3 lv = bvel;
    lvh = add(lvh, bvh);
    rot_vjp(q[0], q[1], q[2], q[3], w, bw, lq, lw);
    rot_vjp(q[0], q[1], q[2], q[3], wh, bwh, lq, lwh);
    double lqPre[4] = {lq[0], lq[1], lq[2], lq[3]}, lhqPre[4] = {0.0, 0.0, 0.0, 0.0};
    V3 laPre = zero3(), lalPre = zero3();
    if (S.k(I_STATUS, p) != -1) {
        // x' = x + vh dt; q' = N(q + hq dt); v' = vh + a dt/2; w' = wh + al dt/2
        lvh = add(lvh, scl(lx, dt));
        double q0[4], hq[4];
        ld4(pre, P_Q, n, p, q0);
        ld4(pre, P_HQ, n, p, hq);
        double qt[4] = {q0[0] + hq[0] * dt, q0[1] + hq[1] * dt, q0[2] + hq[2] * dt, q0[3] + hq[3] * dt};
        double qtb[4] = {0.0, 0.0, 0.0, 0.0};
        quat_unit_vjp(qt, lq, qtb);
        for (int k = 0; k < 4; ++k) {
            lqPre[k] = qtb[k];
            lhqPre[k] = qtb[k] * dt;
        }
        lvh = add(lvh, lv);
        laPre = scl(lv, 0.5 * dt);
        lwh = add(lwh, lw);
        lalPre = scl(lw, 0.5 * dt);
    }
    st3(lam, L_X, n, p, lx);
    st4(lam, L_Q, n, p, lqPre);
    st3(lam, L_VH, n, p, lvh);
    st3(lam, L_WH, n, p, lwh);
    st4(lam, L_HQ, n, p, lhqPre);
    st3(lam, L_A, n, p, laPre);
    st3(lam, L_AL, n, p, lalPre);
}

__global__ void __launch_bounds__(BLOCK) k_adj_seed(Soa S, const SceneParams* sp, const double* dHeight,
                                                    double* lam)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    if (p >= n) return;
    for (int k = 0; k < NLAM; ++k) lam[(long)k * n + p] = 0.0;
    const int s = S.k(I_SCENE, p);
    lam[(long)(L_X + 2) * n + p] = dHeight[s] / (double)(sp[s].row1 - sp[s].row0);
}

}  // namespace

// ---------------------------------------------------------------- host side
// Device copy of what a step depends on: particle state and contact history
// (history slots below the run's high-water marks only).
struct Checkpoint {
    double* d = nullptr;
    unsigned long long* ppOcc = nullptr;
    unsigned long long* ppSeen = nullptr;
    uint32_t* ppKey = nullptr;
    double* ppVal = nullptr;
    unsigned* pwOcc = nullptr;
    unsigned* pwSeen = nullptr;
    uint32_t* pwKey = nullptr;
    double* pwVal = nullptr;
    int ppW = 0, pwW = 0;  // slots stored per particle
    int step = 0;
    void release()
    {
        void* ptrs[] = {d, ppOcc, ppSeen, ppKey, ppVal, pwOcc, pwSeen, pwKey, pwVal};
        for (void* q : ptrs)
            if (q) hipFree(q);
        *this = Checkpoint{};
    }
};

struct DiffRun {
    int steps = 0, K = 0, plate = 0;
    std::vector<Checkpoint> cps;  // before step 1 + c K
    Checkpoint init, fin;
    std::vector<long long> ppCount, pwCount;
    long long* dPP = nullptr;
    long long* dPW = nullptr;
    int* dHigh = nullptr;
    unsigned* dWallCount = nullptr;
    double* plateAcc = nullptr;
    double* snaps = nullptr;  // [K + 1][NSNAP][n]
    unsigned long long* occT = nullptr;
    unsigned* pwOccT = nullptr;
    int2* pairT = nullptr;
    double* pairPrevT = nullptr;
    long long pairCap = 0, pairPrevCap = 0;
    int4* wallT = nullptr;
    double* wallPrevT = nullptr;
    long long wallCap = 0, wallPrevCap = 0;
    double* lam = nullptr;
    double* z = nullptr;
    double* lamH = nullptr;
    double* lamHW = nullptr;
    double* grads = nullptr;
    double* seeds = nullptr;  // [2][scenes]: height, plate (already / steps)
    bool ready = false;
    void release()
    {
        for (auto& c : cps) c.release();
        cps.clear();
        init.release();
        fin.release();
        void* ptrs[] = {dPP, dPW, dHigh, dWallCount, plateAcc, snaps, occT, pwOccT, pairT, pairPrevT, wallT,
                        wallPrevT, lam, z, lamH, lamHW, grads, seeds};
        for (void* q : ptrs)
            if (q) hipFree(q);
        *this = DiffRun{};
    }
};

void diff_release(DiffRun* r)
{
    if (!r) return;
    r->release();
    delete r;
}

namespace {

int save_checkpoint(dem_solver* s, Checkpoint& c, int ppW, int pwW)
{
    const long n = s->n;
    hipStream_t st = s->stream;
    if (!c.d) {
        TRY(dalloc(c.d, (size_t)NDF * n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.ppOcc, n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.ppSeen, 2 * n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.pwOcc, n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.pwSeen, 2 * n, "hipMalloc(checkpoint)"));
    }
    if (ppW > c.ppW) {
        if (c.ppKey) hipFree(c.ppKey);
        if (c.ppVal) hipFree(c.ppVal);
        c.ppKey = nullptr; c.ppVal = nullptr;
        TRY(dalloc(c.ppKey, (size_t)ppW * n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.ppVal, (size_t)ppW * 9 * n, "hipMalloc(checkpoint)"));
    }
    if (pwW > c.pwW) {
        if (c.pwKey) hipFree(c.pwKey);
        if (c.pwVal) hipFree(c.pwVal);
        c.pwKey = nullptr; c.pwVal = nullptr;
        TRY(dalloc(c.pwKey, (size_t)pwW * n, "hipMalloc(checkpoint)"));
        TRY(dalloc(c.pwVal, (size_t)pwW * 9 * n, "hipMalloc(checkpoint)"));
    }
    c.ppW = std::max(c.ppW, ppW);
    c.pwW = std::max(c.pwW, pwW);
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    };
    HIP_TRY(cp(c.d, s->dState[s->cur], sizeof(double) * NDF * n), "checkpoint");
    HIP_TRY(cp(c.ppOcc, s->H.ppOcc, sizeof(unsigned long long) * n), "checkpoint");
    HIP_TRY(cp(c.ppSeen, s->H.ppSeen, sizeof(unsigned long long) * 2 * n), "checkpoint");
    HIP_TRY(cp(c.pwOcc, s->H.pwOcc, sizeof(unsigned) * n), "checkpoint");
    HIP_TRY(cp(c.pwSeen, s->H.pwSeen, sizeof(unsigned) * 2 * n), "checkpoint");
    if (ppW > 0) {
        HIP_TRY(hipMemcpy2DAsync(c.ppKey, sizeof(uint32_t) * ppW, s->H.ppKey, sizeof(uint32_t) * PP_SLOTS,
                                 sizeof(uint32_t) * ppW, n, hipMemcpyDeviceToDevice, st), "checkpoint");
        // values are slot-major: the first ppW slots are one contiguous block
        HIP_TRY(cp(c.ppVal, s->H.ppVal, sizeof(double) * 9 * n * ppW), "checkpoint");
    }
    if (pwW > 0) {
        HIP_TRY(hipMemcpy2DAsync(c.pwKey, sizeof(uint32_t) * pwW, s->H.pwKey, sizeof(uint32_t) * PW_SLOTS,
                                 sizeof(uint32_t) * pwW, n, hipMemcpyDeviceToDevice, st), "checkpoint");
        HIP_TRY(hipMemcpy2DAsync(c.pwVal, sizeof(double) * 9 * pwW, s->H.pwVal, sizeof(double) * 9 * PW_SLOTS,
                                 sizeof(double) * 9 * pwW, n, hipMemcpyDeviceToDevice, st), "checkpoint");
    }
    c.step = s->stepIndex;
    return DEM_OK;
}

int restore_checkpoint(dem_solver* s, const Checkpoint& c)
{
    const long n = s->n;
    hipStream_t st = s->stream;
    s->recFresh = false;
    s->vlistValid = false;  // the candidate lists are rebuilt from the restored centres
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    };
    HIP_TRY(cp(s->dState[s->cur], c.d, sizeof(double) * NDF * n), "restore checkpoint");
    HIP_TRY(cp(s->H.ppOcc, c.ppOcc, sizeof(unsigned long long) * n), "restore checkpoint");
    HIP_TRY(cp(s->H.ppSeen, c.ppSeen, sizeof(unsigned long long) * 2 * n), "restore checkpoint");
    HIP_TRY(cp(s->H.pwOcc, c.pwOcc, sizeof(unsigned) * n), "restore checkpoint");
    HIP_TRY(cp(s->H.pwSeen, c.pwSeen, sizeof(unsigned) * 2 * n), "restore checkpoint");
    if (c.ppW > 0) {
        HIP_TRY(hipMemcpy2DAsync(s->H.ppKey, sizeof(uint32_t) * PP_SLOTS, c.ppKey, sizeof(uint32_t) * c.ppW,
                                 sizeof(uint32_t) * c.ppW, n, hipMemcpyDeviceToDevice, st), "restore checkpoint");
        HIP_TRY(cp(s->H.ppVal, c.ppVal, sizeof(double) * 9 * n * c.ppW), "restore checkpoint");
    }
    if (c.pwW > 0) {
        HIP_TRY(hipMemcpy2DAsync(s->H.pwKey, sizeof(uint32_t) * PW_SLOTS, c.pwKey, sizeof(uint32_t) * c.pwW,
                                 sizeof(uint32_t) * c.pwW, n, hipMemcpyDeviceToDevice, st), "restore checkpoint");
        HIP_TRY(hipMemcpy2DAsync(s->H.pwVal, sizeof(double) * 9 * PW_SLOTS, c.pwVal, sizeof(double) * 9 * c.pwW,
                                 sizeof(double) * 9 * c.pwW, n, hipMemcpyDeviceToDevice, st), "restore checkpoint");
    }
    s->stepIndex = c.step;
    s->pendingFinish = false;
    s->vlistValid = false;
    return DEM_OK;
}

int diff_high(dem_solver* s, DiffRun& R, int& ppW, int& pwW)
{
    int h[2];
    HIP_TRY(hipMemcpyAsync(h, R.dHigh, sizeof(h), hipMemcpyDeviceToHost, s->stream), "download high-water");
    HIP_TRY(hipStreamSynchronize(s->stream), "download high-water");
    ppW = std::min(PP_SLOTS, std::max(h[0], 1));
    pwW = std::min(PW_SLOTS, std::max(h[1], 1));
    return DEM_OK;
}

// Puts the rows in cell order once, before a differentiable run records
// anything: input rows may carry no locality, and the run keeps its rows
// fixed from here on (every tape entry is a row).  Only with empty history
// tables (no checkpoint loaded), whose rows are then reset to the rows.
int diff_reorder(dem_solver* s)
{
    const long n = s->n;
    hipStream_t st = s->stream;
    Soa S = s->soa();
    hipLaunchKernelGGL(k_step_init, dim3(1), dim3(BLOCK), 0, st, s->sc, s->sceneAcc, s->nscenes);
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->dSp, 0.0, s->offs, s->sc, s->boxPart,
                       s->partScene, s->sceneAcc, (const double*)nullptr, 0.0, (double4*)nullptr,
                       (uint32_t*)nullptr, (int4*)nullptr, (double*)nullptr);
    hipLaunchKernelGGL(k_grid_setup, dim3(s->nscenes), dim3(BLOCK), 0, st, s->sc, s->dSp, s->boxPart,
                       s->partScene, s->sceneAcc, s->dGrid, s->nscenes, s->cellCap, s->gridDone);
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->dGrid, s->keyA, s->valA);
    TRY(launch_check("grid keys"));
    TRY(radix_sort(st, s->keyA, s->valA, s->keyB, s->valB, n, s->rsCounts, s->rsBlocks, &s->sc->nbits,
                   s->rsDigits));
    hipLaunchKernelGGL(k_fixup, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->keyA, s->valA, n,
                       S.i + (long)I_ORIG * n);
    Soa D{s->dState[1 - s->cur], s->iState[1 - s->cur], n};
    hipLaunchKernelGGL(k_permute, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, D, s->valA);
    s->cur = 1 - s->cur;
    s->recFresh = false;
    hipLaunchKernelGGL(k_identity, dim3(blocks_for(n)), dim3(BLOCK), 0, st, D.i + (long)I_HIST * n, n);
    return launch_check("differentiable run row order");
}

// One step.  more: the run's next step follows with no checkpoint in
// between, so the law kernel may finish this step and advance the next one
// itself (the forward run without tapes of the pair law).
int diff_step(dem_solver* s, double dt, bool more = false)
{
    const long n = s->n;
    // the advance also writes the rows' contact records (runs take
    // zero-offset single spheres), so the evaluation packs none; the law
    // kernel of the previous step may have advanced this one already
    if (!s->advancedAhead)
        hipLaunchKernelGGL(k_advance, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->soa(), dt, s->recs());
    s->advancedAhead = false;
    s->recFresh = true;
    s->nextDt = more ? dt : 0.0;
    const int rc = dem_compute_forces(s, nullptr, nullptr, 0);
    s->nextDt = 0.0;
    TRY(rc);
    if (!s->advancedAhead)
        hipLaunchKernelGGL(k_finish, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->soa(), dt);
    return launch_check("differentiable step");
}

void diff_count(dem_solver* s, DiffRun& R, int step)
{
    hipLaunchKernelGGL(k_diff_count, dim3(1), dim3(64), 0, s->stream, s->sc, R.dWallCount, R.dPP, R.dPW, R.dHigh,
                       step);
}

// Grows a tape buffer to hold `need` records of `per` elements (contents not kept).
template <typename T>
int ensure(T*& p, long long& cap, long long need, size_t per, const char* what)
{
    if (p && need <= cap) return DEM_OK;
    if (p) hipFree(p);
    p = nullptr;
    const long long c = need + need / 4 + 1024;
    HIP_TRY(hipMalloc((void**)&p, sizeof(T) * per * (size_t)c), what);
    cap = c;
    return DEM_OK;
}

}  // namespace

extern "C" {

int dem_diff_forward(dem_solver* s, int steps, int checkpoint_interval, int plate_wall, double* height,
                     double* plate_load)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    s->scalarsReset = false;  // the run's own kernels write the step counters
    if (steps < 1 || checkpoint_interval < 1) return set_error(DEM_E_INVALID, "steps and checkpoint interval must be positive");
    if (s->stepIndex != 0) return set_error(DEM_E_INVALID, "a differentiable run starts from a freshly initialized solver");
    if (s->nb > 0 || s->offs || s->universal || s->clump)
        return set_error(DEM_E_UNSUPPORTED, "differentiable runs cover spheres, walls and gravity (no bodies, clump offsets or universal gravitation)");
    for (const SceneParams& q : s->hsp)
        if (plate_wall < 0 || plate_wall >= q.nwall)
            return set_error(DEM_E_INVALID, "plate wall index %d is not a wall of every scene", plate_wall);
    if (!s->diff) s->diff = new DiffRun();
    DiffRun& R = *s->diff;
    R.release();
    R.steps = steps;
    R.K = checkpoint_interval;
    R.plate = plate_wall;
    s->reorderInterval = 0;  // rows stay put (after diff_reorder): every tape entry is a row
    s->keepRows = true;      // (and candidate-list rebuilds leave them in place)
    s->verlet = s->refC != nullptr && !s->verletBroken;
    s->countAlways = true;   // checkpoints size history copies by the high-water marks
    s->vlistValid = false;
    s->recFresh = false;
    const long n = s->n;
    const double dt = s->pp.dt;
    if (!s->historyLoaded) TRY(diff_reorder(s));
    TRY(dalloc(R.dPP, steps + 1, "hipMalloc(run counts)"));
    TRY(dalloc(R.dPW, steps + 1, "hipMalloc(run counts)"));
    TRY(dalloc(R.dHigh, 2, "hipMalloc(run counts)"));
    TRY(dalloc(R.dWallCount, 1, "hipMalloc(run counts)"));
    TRY(dalloc(R.plateAcc, n, "hipMalloc(plate load)"));
    HIP_TRY(hipMemsetAsync(R.dHigh, 0, sizeof(int) * 2, s->stream), "clear");
    HIP_TRY(hipMemsetAsync(R.dWallCount, 0, sizeof(unsigned), s->stream), "clear");
    HIP_TRY(hipMemsetAsync(R.plateAcc, 0, sizeof(double) * n, s->stream), "clear");
    TRY(save_checkpoint(s, R.init, 1, 1));
    s->tape = Tape{};
    s->tape.wallCount = R.dWallCount;
    s->tape.plateAcc = R.plateAcc;
    s->tape.plate = plate_wall;
    int rc = dem_compute_forces(s, nullptr, nullptr, 0);
    if (rc == DEM_OK) {
        diff_count(s, R, 0);
        hipLaunchKernelGGL(k_init_half, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->soa(), 0.5 * dt);
        const int nseg = (steps + R.K - 1) / R.K;
        R.cps.resize(nseg);
        for (int step = 1; step <= steps && rc == DEM_OK; ++step) {
            if ((step - 1) % R.K == 0) {
                int ppW, pwW;
                rc = diff_high(s, R, ppW, pwW);
                if (rc == DEM_OK) rc = save_checkpoint(s, R.cps[(step - 1) / R.K], ppW, pwW);
            }
            // the step before a checkpoint (and the last) finishes on its own,
            // so the checkpoint holds a finished state
            if (rc == DEM_OK) rc = diff_step(s, dt, step < steps && step % R.K != 0);
            if (rc == DEM_OK) diff_count(s, R, step);
        }
    }
    s->tape = Tape{};
    TRY(rc);
    int ppW, pwW;
    TRY(diff_high(s, R, ppW, pwW));
    TRY(save_checkpoint(s, R.fin, ppW, pwW));
    R.ppCount.assign(steps + 1, 0);
    R.pwCount.assign(steps + 1, 0);
    HIP_TRY(hipMemcpy(R.ppCount.data(), R.dPP, sizeof(long long) * (steps + 1), hipMemcpyDeviceToHost), "download counts");
    HIP_TRY(hipMemcpy(R.pwCount.data(), R.dPW, sizeof(long long) * (steps + 1), hipMemcpyDeviceToHost), "download counts");
    // observables, summed per scene in input-row order (a scene's input rows
    // and rows are the same range)
    std::vector<double> zpos(n), acc(n);
    std::vector<int> orig(n), row(n);
    HIP_TRY(hipMemcpy(zpos.data(), s->dState[s->cur] + (size_t)F_Z * n, sizeof(double) * n, hipMemcpyDeviceToHost),
            "download heights");
    HIP_TRY(hipMemcpy(acc.data(), R.plateAcc, sizeof(double) * n, hipMemcpyDeviceToHost), "download plate load");
    HIP_TRY(hipMemcpy(orig.data(), s->iState[s->cur] + (size_t)I_ORIG * n, sizeof(int) * n, hipMemcpyDeviceToHost),
            "download rows");
    for (long p = 0; p < n; ++p) row[orig[p]] = (int)p;
    for (int k = 0; k < s->nscenes; ++k) {
        const SceneParams& q = s->hsp[k];
        double hz = 0.0, f = 0.0;
        for (long i = q.row0; i < q.row1; ++i) {
            hz += zpos[row[i]];
            f += acc[row[i]];
        }
        if (height) height[k] = hz / (double)(q.row1 - q.row0);
        if (plate_load) plate_load[k] = f / (double)steps;
    }
    R.ready = true;
    return DEM_OK;
}

int dem_diff_backward(dem_solver* s, const double* d_height, const double* d_plate_load, double* d_mech)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    s->scalarsReset = false;
    if (!s->diff || !s->diff->ready) return set_error(DEM_E_INVALID, "dem_diff_backward needs a completed dem_diff_forward");
    if (!d_height || !d_plate_load || !d_mech) return set_error(DEM_E_INVALID, "null argument");
    DiffRun& R = *s->diff;
    const long n = s->n;
    const int S_ = s->nscenes;
    const double dt = s->pp.dt;
    hipStream_t st = s->stream;
    const int K = R.K;
    if (!R.lam) {
        TRY(dalloc(R.lam, (size_t)NLAM * n, "hipMalloc(adjoint state)"));
        TRY(dalloc(R.z, (size_t)NZ * n, "hipMalloc(adjoint scratch)"));
        TRY(dalloc(R.lamH, (size_t)PP_SLOTS * 9 * n, "hipMalloc(history adjoint)"));
        TRY(dalloc(R.lamHW, (size_t)PW_SLOTS * 9 * n, "hipMalloc(history adjoint)"));
        TRY(dalloc(R.grads, (size_t)8 * S_, "hipMalloc(material adjoint)"));
        TRY(dalloc(R.seeds, (size_t)2 * S_, "hipMalloc(seeds)"));
        TRY(dalloc(R.snaps, (size_t)(K + 1) * NSNAP * n, "hipMalloc(tape)"));
        TRY(dalloc(R.occT, (size_t)K * n, "hipMalloc(tape)"));
        TRY(dalloc(R.pwOccT, (size_t)K * n, "hipMalloc(tape)"));
    }
    std::vector<double> seedHost(2 * S_);
    for (int k = 0; k < S_; ++k) {
        seedHost[k] = d_height[k];
        seedHost[S_ + k] = d_plate_load[k] / (double)R.steps;
    }
    HIP_TRY(hipMemcpyAsync(R.seeds, seedHost.data(), sizeof(double) * 2 * S_, hipMemcpyHostToDevice, st), "upload seeds");
    HIP_TRY(hipMemsetAsync(R.lamH, 0, sizeof(double) * PP_SLOTS * 9 * n, st), "clear");
    HIP_TRY(hipMemsetAsync(R.lamHW, 0, sizeof(double) * PW_SLOTS * 9 * n, st), "clear");
    HIP_TRY(hipMemsetAsync(R.grads, 0, sizeof(double) * 8 * S_, st), "clear");
    hipLaunchKernelGGL(k_adj_seed, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->soa(), s->dSp, R.seeds, R.lam);
    const Walls W{s->nw, s->wallBuf, s->wallBuf + 3 * s->nw};
    const int nb = blocks_for(n);
    const double* plateSeed = R.seeds + S_;

    // record one step's tape (contacts of `step` into slot k of the segment)
    auto record = [&](int k, long long pairOff, long long wallOff) {
        s->tape = Tape{};
        s->tape.ppOcc = R.occT + (size_t)k * n;
        s->tape.pwOcc = R.pwOccT + (size_t)k * n;
        s->tape.pair = R.pairT + pairOff;
        s->tape.pairPrev = R.pairPrevT + pairOff * 9;
        s->tape.wall = R.wallT + wallOff;
        s->tape.wallPrev = R.wallPrevT + wallOff * 9;
        s->tape.wallCount = R.dWallCount;
        s->tape.plate = R.plate;
    };
    auto reverse = [&](int k, int step, long long pairOff, long long wallOff, bool initial) {
        const double* pre = R.snaps + (size_t)k * NSNAP * n;
        const double* post = R.snaps + (size_t)(k + 1) * NSNAP * n;
        hipLaunchKernelGGL(k_adj_prep, dim3(nb), dim3(BLOCK), 0, st, s->soa(), pre, post, dt, initial, s->recs(),
                           R.lam, R.z);
        const long long np = R.ppCount[step], nw = R.pwCount[step];
        if (np > 0)
            hipLaunchKernelGGL(k_adj_pairs, dim3((unsigned)((np + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, (long)np,
                               R.pairT + pairOff, R.pairPrevT + pairOff * 9, R.occT + (size_t)k * n, s->recs(),
                               s->dMatPP, initial, n, R.z, R.lamH, R.grads);
        if (nw > 0)
            hipLaunchKernelGGL(k_adj_walls, dim3((unsigned)((nw + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, (long)nw,
                               R.wallT + wallOff, R.wallPrevT + wallOff * 9, R.pwOccT + (size_t)k * n, s->recs(),
                               s->dMatPW, s->dSp, W, initial, R.plate, plateSeed, n, R.z, R.lamHW, R.grads);
        if (!initial)
            hipLaunchKernelGGL(k_adj_body, dim3(nb), dim3(BLOCK), 0, st, s->soa(), pre, dt, R.lam, R.z);
        return launch_check("reverse step");
    };

    int rc = DEM_OK;
    const int nseg = (int)R.cps.size();
    for (int c = nseg - 1; c >= 0 && rc == DEM_OK; --c) {
        const int s0 = 1 + c * K, s1 = std::min(R.steps, s0 + K - 1), len = s1 - s0 + 1;
        long long needP = 0, needW = 0;
        std::vector<long long> po(len), wo(len);
        for (int k = 0; k < len; ++k) {
            po[k] = needP;
            wo[k] = needW;
            needP += R.ppCount[s0 + k];
            needW += R.pwCount[s0 + k];
        }
        TRY(ensure(R.pairT, R.pairCap, needP, 1, "hipMalloc(contact tape)"));
        TRY(ensure(R.pairPrevT, R.pairPrevCap, needP, 9, "hipMalloc(contact tape)"));
        TRY(ensure(R.wallT, R.wallCap, needW, 1, "hipMalloc(wall tape)"));
        TRY(ensure(R.wallPrevT, R.wallPrevCap, needW, 9, "hipMalloc(wall tape)"));
        TRY(restore_checkpoint(s, R.cps[c]));
        for (int k = 0; k < len && rc == DEM_OK; ++k) {
            hipLaunchKernelGGL(k_snap, dim3(nb), dim3(BLOCK), 0, st, s->soa(), R.snaps + (size_t)k * NSNAP * n);
            record(k, po[k], wo[k]);
            rc = diff_step(s, dt);
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, R.dWallCount, 1L);
        }
        s->tape = Tape{};
        if (rc != DEM_OK) break;
        hipLaunchKernelGGL(k_snap, dim3(nb), dim3(BLOCK), 0, st, s->soa(), R.snaps + (size_t)len * NSNAP * n);
        for (int k = len - 1; k >= 0 && rc == DEM_OK; --k) rc = reverse(k, s0 + k, po[k], wo[k], false);
    }
    if (rc == DEM_OK) {
        // first evaluation (dt = 0) and half-step initialisation
        TRY(ensure(R.pairT, R.pairCap, R.ppCount[0], 1, "hipMalloc(contact tape)"));
        TRY(ensure(R.pairPrevT, R.pairPrevCap, R.ppCount[0], 9, "hipMalloc(contact tape)"));
        TRY(ensure(R.wallT, R.wallCap, R.pwCount[0], 1, "hipMalloc(wall tape)"));
        TRY(ensure(R.wallPrevT, R.wallPrevCap, R.pwCount[0], 9, "hipMalloc(wall tape)"));
        TRY(restore_checkpoint(s, R.init));
        hipLaunchKernelGGL(k_snap, dim3(nb), dim3(BLOCK), 0, st, s->soa(), R.snaps);
        record(0, 0, 0);
        rc = dem_compute_forces(s, nullptr, nullptr, 0);
        s->tape = Tape{};
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, R.dWallCount, 1L);
        if (rc == DEM_OK) {
            hipLaunchKernelGGL(k_init_half, dim3(nb), dim3(BLOCK), 0, st, s->soa(), 0.5 * dt);
            hipLaunchKernelGGL(k_snap, dim3(nb), dim3(BLOCK), 0, st, s->soa(), R.snaps + (size_t)NSNAP * n);
            rc = reverse(0, 0, 0, 0, true);
        }
    }
    s->tape = Tape{};
    // leave the solver at the end of the run
    TRY(restore_checkpoint(s, R.fin));
    TRY(rc);
    std::vector<double> g(8 * S_);
    HIP_TRY(hipMemcpyAsync(g.data(), R.grads, sizeof(double) * 8 * S_, hipMemcpyDeviceToHost, st), "download gradients");
    HIP_TRY(hipStreamSynchronize(st), "download gradients");
    std::vector<Material> mpp(S_), mpw(S_);
    HIP_TRY(hipMemcpy(mpp.data(), s->dMatPP, sizeof(Material) * S_, hipMemcpyDeviceToHost), "download materials");
    HIP_TRY(hipMemcpy(mpw.data(), s->dMatPW, sizeof(Material) * S_, hipMemcpyDeviceToHost), "download materials");
    for (int k = 0; k < S_; ++k) {
        const double* a = &g[8 * k];
        double* o = d_mech + 8 * k;
        // [mu, epsN, epsS, kN] for the particle-particle law, then the particle-wall law
        o[0] = a[0]; o[1] = a[1] / mpp[k].epsN; o[2] = a[2] / mpp[k].epsS; o[3] = a[3];
        o[4] = a[4]; o[5] = a[5] / mpw[k].epsN; o[6] = a[6] / mpw[k].epsS; o[7] = a[7];
    }
    return DEM_OK;
}

}  // extern "C"

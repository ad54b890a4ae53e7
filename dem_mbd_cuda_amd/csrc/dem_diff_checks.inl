// dem_diff_checks.inl -- host entry points that evaluate one contact law and
// its reverse-mode derivative on the CPU (the same DEM_FN code the kernels
// run), so the hand-written derivatives can be checked against finite
// differences without a device.  Included by dem_hip.hip inside extern "C".
//
// Packed layouts (doubles):
//   body  [21]: pos3 vel3 vh3 w3 wh3 off3 mass r status
//   mech  [9] : epsS mu epsN cohesion beta muT muR kN dt
//   grads body [15]: pos vel vh w wh;  material [4]: mu logN logS kN

namespace {

dem::Body unpack_body(const double* b)
{
    dem::Body B;
    B.pos = dem::mk(b[0], b[1], b[2]);
    B.vel = dem::mk(b[3], b[4], b[5]);
    B.vh = dem::mk(b[6], b[7], b[8]);
    B.w = dem::mk(b[9], b[10], b[11]);
    B.wh = dem::mk(b[12], b[13], b[14]);
    B.off = dem::mk(b[15], b[16], b[17]);
    B.mass = b[18];
    B.r = b[19];
    B.status = (int)b[20];
    B.cid = 0;
    return B;
}

dem::Material unpack_mech(const double* m)
{
    dem_mech q{};
    q.epsS = m[0]; q.mu = m[1]; q.epsN = m[2]; q.cohesion = m[3]; q.beta = m[4]; q.muT = m[5];
    q.muR = m[6]; q.kN = m[7];
    return make_material(q, m[8]);
}

void pack_grad(const dem::BodyGrad& g, double* o)
{
    const dem::V3 v[5] = {g.pos, g.vel, g.vh, g.w, g.wh};
    for (int k = 0; k < 5; ++k) {
        o[3 * k] = v[k].x;
        o[3 * k + 1] = v[k].y;
        o[3 * k + 2] = v[k].z;
    }
}

}  // namespace

/* One particle-particle contact: in = body I [21], body J [21], mech [9],
 * stored slot values [9]; out = force on I [3], torque term [3], new slot
 * values [9]. */
int dem_check_pp_law(const double* in, double* out)
{
    if (!in || !out) return set_error(DEM_E_INVALID, "null argument");
    const dem::Body I = unpack_body(in), J = unpack_body(in + 21);
    const dem::Material m = unpack_mech(in + 42);
    double slot[9];
    for (int k = 0; k < 9; ++k) slot[k] = in[51 + k];
    const dem::V3 delta = dem::sub(J.pos, I.pos);
    dem::V3 cf, tq;
    dem::pp_law(I, J, delta, dem::dot(delta, delta), m, slot, cf, tq);
    out[0] = cf.x; out[1] = cf.y; out[2] = cf.z;
    out[3] = tq.x; out[4] = tq.y; out[5] = tq.z;
    for (int k = 0; k < 9; ++k) out[6 + k] = slot[k];
    return DEM_OK;
}

/* Reverse of dem_check_pp_law: seeds = adjoints of its 15 outputs; grads =
 * body I [15], body J [15], slot values [9], material [4]. */
int dem_check_pp_law_vjp(const double* in, const double* seeds, double* grads)
{
    if (!in || !seeds || !grads) return set_error(DEM_E_INVALID, "null argument");
    const dem::Body I = unpack_body(in), J = unpack_body(in + 21);
    const dem::Material m = unpack_mech(in + 42);
    dem::BodyGrad Ib = dem::zero_body_grad(), Jb = dem::zero_body_grad();
    double prevb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    dem::MatGrad g{0.0, 0.0, 0.0, 0.0};
    dem::pp_law_vjp(I, J, m, in + 51, dem::mk(seeds[0], seeds[1], seeds[2]),
                    dem::mk(seeds[3], seeds[4], seeds[5]), seeds + 6, Ib, Jb, prevb, g);
    pack_grad(Ib, grads);
    pack_grad(Jb, grads + 15);
    for (int k = 0; k < 9; ++k) grads[30 + k] = prevb[k];
    grads[39] = g.mu; grads[40] = g.logN; grads[41] = g.logS; grads[42] = g.kN;
    return DEM_OK;
}

/* One particle-wall contact: in = body [21], outward normal [3], origin [3],
 * mech [9], slot values [9]; out = force [3], torque [3], new slot values [9]. */
int dem_check_pw_law(const double* in, double* out)
{
    if (!in || !out) return set_error(DEM_E_INVALID, "null argument");
    const dem::Body I = unpack_body(in);
    const dem::V3 outward = dem::mk(in[21], in[22], in[23]), origin = dem::mk(in[24], in[25], in[26]);
    const dem::Material m = unpack_mech(in + 27);
    double slot[9];
    for (int k = 0; k < 9; ++k) slot[k] = in[36 + k];
    const double depth = I.r - dem::dot(dem::sub(I.pos, origin), outward);
    dem::V3 f = dem::zero3(), t = dem::zero3();
    dem::pw_law(I, outward, depth, m, slot, f, t);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    out[3] = t.x; out[4] = t.y; out[5] = t.z;
    for (int k = 0; k < 9; ++k) out[6 + k] = slot[k];
    return DEM_OK;
}

/* Reverse of dem_check_pw_law: grads = body [15], slot values [9], material [4]. */
int dem_check_pw_law_vjp(const double* in, const double* seeds, double* grads)
{
    if (!in || !seeds || !grads) return set_error(DEM_E_INVALID, "null argument");
    const dem::Body I = unpack_body(in);
    const dem::V3 outward = dem::mk(in[21], in[22], in[23]), origin = dem::mk(in[24], in[25], in[26]);
    const dem::Material m = unpack_mech(in + 27);
    dem::BodyGrad Ib = dem::zero_body_grad();
    double prevb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    dem::MatGrad g{0.0, 0.0, 0.0, 0.0};
    dem::pw_law_vjp(I, outward, origin, m, in + 36, dem::mk(seeds[0], seeds[1], seeds[2]),
                    dem::mk(seeds[3], seeds[4], seeds[5]), seeds + 6, Ib, prevb, g);
    pack_grad(Ib, grads);
    for (int k = 0; k < 9; ++k) grads[15 + k] = prevb[k];
    grads[24] = g.mu; grads[25] = g.logN; grads[26] = g.logS; grads[27] = g.kN;
    return DEM_OK;
}

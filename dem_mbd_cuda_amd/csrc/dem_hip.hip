// dem_hip.hip -- libdemhip.so: the FP64 DEM step on gfx950 behind include/dem_hip.h.
//
// Step pipeline (one HIP stream per solver, no host synchronisation in a
// pure-DEM step):
//   k_cells      component centre, integer cell, 21-bit range check, cell
//                bounding box (wave shuffles + one atomic per block)
//   k_keys       bbox-relative linear cell index (same order as the
//                reference's packed 64-bit key) -> radix sort payload
//   radix sort   stable LSD sort, 8-bit digits, passes beyond the key's
//                significant bits are skipped on the device
//   k_fixup      within a cell, order by input row (the reference's stable
//                sort of an index sequence orders ties by component index)
//   k_permute    every reorder_interval steps the whole particle SoA is
//                gathered into cell order so every later access is coalesced
//   k_cellstart  dense cell -> first sorted index table (when it fits)
//   k_posr       packed (x, y, z, r) and component id in cell order
//   k_pack       per-particle contact record (one 160-byte line per row)
//   k_scan       27-cell candidate scan, nine contiguous column ranges per
//                particle, the block's neighbourhood staged in LDS column by
//                column, then history expiry and one slot per contact
//   k_off_*      prefix sum of list lengths, compacted (owner, entry) list
//   k_pairs      particle-particle law, one thread per contact
//   k_tri_sift   (coupled runs) rows whose cell holds no triangle entry take
//                the no-contact path; the others go to a work list
//   k_law_pt     per listed row and body: triangle manifold, manifold law,
//                microstate terms, read-only triangle contacts
//   k_law        walls, field, ordered sum of the pair results, triangle
//                accumulator, accelerations
//   k_reduce     deterministic row sums for the body loads and diagnostics
//
// History is indexed by input row, so it never moves when particles are
// reordered; occupancy and "seen this step / last step" live in bit masks
// instead of the reference's per-slot lastSeen words.
#include "dem_device.hpp"
#include "dem_adjoint.hpp"

#include "../../include/dem_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace dem;

namespace {

thread_local char g_error[512];

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr, what)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_error(DEM_E_DEVICE, "%s: %s", what, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// Scratch memory of dem_compute_forces_noalloc: the per-step triangle-grid
// buffers of a call that may not allocate.  Every device the library runs on
// holds one copy for the life of the process; host code reaches it only
// through the address scratch_on_current_device returns for that device.
static __device__ __align__(256) unsigned char scratch[DEM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchLock;
void* g_scratchPtr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_on_current_device(void** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxScratchDevices)
        return set_error(DEM_E_DEVICE, "device %d is beyond the no-allocation scratch table", dev);
    std::lock_guard<std::mutex> guard(g_scratchLock);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(DEM_E_DEVICE, "hipGetSymbolAddress(scratch): null address");
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return DEM_OK;
}

constexpr int PP_SLOTS = DEM_PARTICLE_HISTORY_SLOTS;
constexpr int PW_SLOTS = DEM_WALL_HISTORY_SLOTS;
constexpr int PT_SLOTS = DEM_BODY_HISTORY_SLOTS;
constexpr int BODY_ROW = 29;  // per (body, row): 6 load components, Fz, 22 microstate terms
constexpr int MAX_LIST = PP_SLOTS;  // contact-list capacity per component
constexpr long kSkinSteps = 1024;   // steps a candidate list is sized to last (adaptive skin)
constexpr int SLOT_BITS = 6;        // list entry = partner row << SLOT_BITS | slot
static_assert((1 << SLOT_BITS) == PP_SLOTS, "slot field width");
constexpr long MAX_PARTICLES = (1L << (31 - SLOT_BITS)) - 1;
constexpr int PAIR_BLOCKS = 8192;   // default grid of the per-contact law (grid-stride)
constexpr int BLOCK = 256;

// Double fields of the particle SoA (one array of n doubles each).
enum DField {
    F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_WX, F_WY, F_WZ, F_QW, F_QX, F_QY, F_QZ,
    F_VHX, F_VHY, F_VHZ, F_WHX, F_WHY, F_WHZ, F_HQW, F_HQX, F_HQY, F_HQZ,
    F_AX, F_AY, F_AZ, F_AWX, F_AWY, F_AWZ,
    F_MASS, F_IX, F_IY, F_IZ, F_CR, F_BX, F_BY, F_BZ,
    F_CPX, F_CPY, F_CPZ, F_OFX, F_OFY, F_OFZ,
    NDF
};
// Int fields.
// I_ORIG: input row (output order); I_HIST: the particle's row in the contact
// history tables (the input row, or the row of the last candidate-list
// rebuild, when the history moves with the particles).
enum IField { I_PID, I_STATUS, I_CID, I_ORIG, I_CX, I_CY, I_CZ, I_VALID, I_SCENE, I_HIST, NIF };

struct Soa {
    double* d;  // [NDF][n]
    int* i;     // [NIF][n]
    long n;
    __device__ __forceinline__ double& f(int field, long p) const { return d[field * n + p]; }
    __device__ __forceinline__ int& k(int field, long p) const { return i[field * n + p]; }
};

// A solver holds one scene or a batch of independent scenes.  Scene s owns
// the contiguous rows [row0, row1) (contiguous before and after any
// reordering, since the cell index below is scene-major), its walls
// [wall0, wall0 + nwall) and its own cell size.
struct SceneParams {
    double inv;   // 1 / cell size
    double maxr;  // the scene's largest component radius
    long row0, row1;
    int wall0, nwall;
};

// Distance along one axis from coordinate x to the slab of cells with index
// c (cell size h); 0 inside.  A candidate scan skips a cell whose box lies
// farther from the centre than any partner's reach: with cells assigned as
// floor(x / h) in floating point, a centre lies in its cell's box up to a
// relative 1e-16, far inside the margin cell_margin() adds.
__device__ __forceinline__ double slab_gap(double x, long c, double h)
{
    const double lo = (double)c * h, hi = (double)(c + 1) * h;
    return fmax(0.0, fmax(lo - x, x - hi));
}
__device__ __forceinline__ double cell_margin(double4 me, double h)
{
    return 1.0e-9 * (fabs(me.x) + fabs(me.y) + fabs(me.z) + h);
}

// Per-step grid of one scene: cell bounding box and the scene's first index
// in the global linear cell numbering.
struct SceneGrid {
    int lo[3], hi[3];
    unsigned long long dims[3];
    unsigned long long ncells;
    unsigned long long base;
};

// Per-step device scalars.
struct Scalars {
    unsigned long long ncells;  // all scenes
    int nbits;
    int mbits;          // key bits of the Morton row order (candidate-list rebuilds)
    int dense;
    int grid_overflow;  // components out of the 21-bit cell range; kept across steps until the host clears it
    int list_overflow;
    int pp_overflow, pw_overflow, pt_overflow;  // pw+pt feed the "body" counter
    int pp_high, body_high;
    int particle_contacts, body_contacts;
    unsigned long long body_max_depth_bits;
    int tri_overflow;
    int pad;
    int tri_lo[3], tri_hi[3];  // cell box of the triangle grid entries
    long long total_pairs;
    int need_rebuild;   // some component moved more than skin/2 since the last rebuild
    int cand_overflow;  // a rebuild listed more partners than CAND_MAX
    int tri_work;       // rows listed by k_tri_sift (their cell holds triangle entries)
    unsigned long long tri_total;  // triangle grid entries
    // full-scan block passes (k_scan, diagnostics evaluations only, since
    // initialisation): candidates staged in LDS / read from HBM
    unsigned long long scan_lds, scan_hbm;
};

struct History {
    unsigned long long* ppOcc;  // [n]
    unsigned long long* ppSeen; // [2][n]
    uint32_t* ppKey;            // [n][64]
    double* ppVal;              // [64][n][9] (slot-major: a slot of consecutive rows is contiguous)
    unsigned* pwOcc;            // [n]
    unsigned* pwSeen;           // [2][n]
    uint32_t* pwKey;            // [n][16]
    double* pwVal;              // [n][16][9]
    unsigned long long* ptOcc;  // [n] (only with bodies)
    unsigned long long* ptSeen; // [2][n]
    uint32_t* ptKey;            // [n][64]
    double* ptVal;              // [n][64][9]
    long rows;                  // n
};

// Particle-pair history values of slot `slot` of history row `o`.  Slot-major,
// so the contacts of a wave (consecutive rows, mostly the low slots) read and
// write a few contiguous runs instead of a partial line per row.
__device__ __forceinline__ double* pp_slot(const History& H, long o, int slot)
{
    return H.ppVal + ((long)slot * H.rows + o) * 9;
}

struct Walls {
    int n;
    const double* o;  // [w][3]
    const double* nrm;
};

struct TriSet {
    int nt, nb;
    const int* body;        // [nt]
    const double* local;    // [nt][15]
    double* world;          // [nt][15]
    double* aabb;           // [nt][6]
    const double* pose;     // [nb][13]
    int entries;
    // entries sorted by cell: keys are the cell's index in the cell box lo +
    // [0, dims) (x-major, the packed key's order), values the triangles
    const unsigned long long* tkey;
    const int* ttri;
    int lo[3];
    long long dims[3];
    const unsigned* start;  // [cells + 1] first entry of each cell, or null (binary search)
};

struct Outputs {
    double* contact;  // [6][n] by physical index
    double* wall;     // [6][n]
    double* body;     // [6][n]
    double* field;    // [3][n]
    double* bodyRows;            // [nb][nr][BODY_ROW] per (body, component input row): load, Fz, microstate
    unsigned char* bodyTouched;  // [nb][nr] the row holds values for the body (else they are zero)
    double* bodyParticle;  // [nb][n_orig][6] (optional)
    double* bodyDetail;    // [nb][n_orig][19] (optional)
    int* regime;           // [nb][n] by physical index
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ unsigned long long pack_cell(int x, int y, int z)
{
    const int bias = 1 << 20;
    const unsigned long long mask = (1ull << 21) - 1ull;
    return (((unsigned long long)(long long)(x + bias) & mask) << 42) |
           (((unsigned long long)(long long)(y + bias) & mask) << 21) |
           ((unsigned long long)(long long)(z + bias) & mask);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

// Orders a wave's LDS accesses across its lanes (no block barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-wide int reductions finished with one atomic per wave.
__device__ __forceinline__ void atomic_add_wave(int* dst, int v)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}
__device__ __forceinline__ void atomic_max_wave(int* dst, int v)
{
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst, v);
}
__device__ __forceinline__ void atomic_max_wave_u64(unsigned long long* dst, unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst, v);
}
// Appends `item` for the lanes with `on` set to list[*count ...], one atomic
// per wave; the order within the list is unspecified.  block_append does the
// same with one atomic per block; every thread of the block calls it.
__device__ __forceinline__ void block_append(bool on, int item, int* count, int* list)
{
    __shared__ int wcount[BLOCK / 64], wbase[BLOCK / 64 + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(on);
    if (lane == 0) wcount[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            wbase[w] = total;
            total += wcount[w];
        }
        wbase[BLOCK / 64] = total ? atomicAdd(count, total) : 0;
    }
    __syncthreads();
    if (on) list[wbase[BLOCK / 64] + wbase[wave] + __popcll(m & ((1ull << lane) - 1ull))] = item;
}
__device__ __forceinline__ void wave_append(bool on, int item, int* count, int* list)
{
    const unsigned long long m = __ballot(on);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    if (on) list[base + __popcll(m & ((1ull << lane) - 1ull))] = item;
}

// Component view of physical particle p (updateComponentsKernel semantics,
// particle:116-147, folded into the consumer).
__device__ __forceinline__ Body load_body(const Soa& S, long p, bool offs)
{
    Body B;
    const double qw = S.f(F_QW, p), qx = S.f(F_QX, p), qy = S.f(F_QY, p), qz = S.f(F_QZ, p);
    B.w = rot(qw, qx, qy, qz, mk(S.f(F_WX, p), S.f(F_WY, p), S.f(F_WZ, p)));
    B.wh = rot(qw, qx, qy, qz, mk(S.f(F_WHX, p), S.f(F_WHY, p), S.f(F_WHZ, p)));
    B.mass = S.f(F_MASS, p);
    B.r = S.f(F_CR, p);
    B.status = S.k(I_STATUS, p);
    B.cid = (uint32_t)S.k(I_CID, p);
    if (offs) {
        const V3 o = mk(S.f(F_OFX, p), S.f(F_OFY, p), S.f(F_OFZ, p));
        B.off = o;
        B.pos = mk(S.f(F_CPX, p), S.f(F_CPY, p), S.f(F_CPZ, p));
        B.vel = mk(S.f(F_VX, p) + B.w.y * o.z - B.w.z * o.y, S.f(F_VY, p) + B.w.z * o.x - B.w.x * o.z,
                   S.f(F_VZ, p) + B.w.x * o.y - B.w.y * o.x);
        B.vh = mk(S.f(F_VHX, p) + B.wh.y * o.z - B.wh.z * o.y,
                  S.f(F_VHY, p) + B.wh.z * o.x - B.wh.x * o.z,
                  S.f(F_VHZ, p) + B.wh.x * o.y - B.wh.y * o.x);
    } else {
        B.off = zero3();
        B.pos = mk(S.f(F_X, p), S.f(F_Y, p), S.f(F_Z, p));
        B.vel = mk(S.f(F_VX, p), S.f(F_VY, p), S.f(F_VZ, p));
        B.vh = mk(S.f(F_VHX, p), S.f(F_VHY, p), S.f(F_VHZ, p));
    }
    return B;
}

// ------------------------------------------------------------ step kernels
// The per-step counters and the per-scene cell box accumulators, reset by one
// block before a force evaluation: by k_step_init, or at the end of the
// advance that precedes it (which then saves a launch).
__device__ __forceinline__ void reset_step_scalars(Scalars* sc, int* sceneAcc, int nscenes)
{
    for (int i = threadIdx.x; i < 6 * nscenes; i += BLOCK) sceneAcc[i] = (i % 6) < 3 ? INT_MAX : INT_MIN;
    if (threadIdx.x == 0) {
        // grid_overflow is not reset here: it stays set until the host reads it
        sc->list_overflow = 0;
        sc->need_rebuild = 0;
        sc->cand_overflow = 0;
        sc->pp_high = 0;
        sc->body_high = 0;
        sc->particle_contacts = 0;
        sc->body_contacts = 0;
        sc->body_max_depth_bits = 0ull;
        sc->tri_overflow = 0;
        sc->tri_work = 0;
        sc->tri_total = 0ull;
        for (int a = 0; a < 3; ++a) {
            sc->tri_lo[a] = INT_MAX;
            sc->tri_hi[a] = INT_MIN;
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_step_init(Scalars* sc, int* sceneAcc, int nscenes)
{
    reset_step_scalars(sc, sceneAcc, nscenes);
}

// Owner state copied to each of its component rows (clumps), so that the
// contact pipeline runs over components with the per-row code of single
// spheres: the component centre and velocities follow from the owner's
// pose and the body-frame offset (updateComponentsKernel, particle:116-147).
__global__ void __launch_bounds__(BLOCK) k_expand(Soa S, Soa R, const int* owner)
{
    const long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= R.n) return;
    const long o = owner[c];
    constexpr int fields[] = {F_X,   F_Y,   F_Z,   F_VX,  F_VY,  F_VZ,  F_WX,  F_WY,  F_WZ,  F_QW,
                              F_QX,  F_QY,  F_QZ,  F_VHX, F_VHY, F_VHZ, F_WHX, F_WHY, F_WHZ, F_MASS};
#pragma unroll
    for (int f : fields) R.f(f, c) = S.f(f, o);
    R.k(I_STATUS, c) = S.k(I_STATUS, o);
}

// A row's cell and identity as the candidate-list scan reads them, one
// 16-byte load per partner: the cell's three coordinates, each offset by
// 2^20 into 21 bits, packed in x (low word) and y (high word), with bit 63
// set for a component outside the hash range (adjacent to no cell); the
// input row in z and the component id in w.
__device__ __forceinline__ int4 pack_meta(bool ok, int cx, int cy, int cz, int orig, uint32_t cid)
{
    const unsigned long long c =
        ok ? ((unsigned long long)(uint32_t)(cx + (1 << 20)) << 42) |
                 ((unsigned long long)(uint32_t)(cy + (1 << 20)) << 21) | (unsigned long long)(uint32_t)(cz + (1 << 20))
           : 1ull << 63;
    return make_int4((int)(uint32_t)c, (int)(uint32_t)(c >> 32), orig, (int)cid);
}
__device__ __forceinline__ unsigned long long meta_cell(int4 m)
{
    return ((unsigned long long)(uint32_t)m.y << 32) | (uint32_t)m.x;
}

// component centre + cell + range check + bbox (grid:51-71), cell size of
// the particle's scene.  The cell bounding box of each scene is reduced per
// block (wave shuffles, then LDS) into part[block][6] when the block lies in
// one scene; a block that straddles scenes adds its lanes to the scenes'
// accumulators with atomics.  k_grid_setup folds both.  With posr (rows in
// cell order: sorted position = row) it also writes k_posr's outputs, which
// a step without rebuild then uses as they are.  inv1: the inverse cell size
// of a single scene (0 with several), so the cell waits on no other load.
__global__ void __launch_bounds__(BLOCK) k_cells(Soa S, const SceneParams* sp, double inv1, bool offs, Scalars* sc,
                                                 int* part, int* partScene, int* sceneAcc, const double* refC,
                                                 double halfSkin2, double4* posr, uint32_t* cidS, int4* meta,
                                                 double* dispMax)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    int scene = -1;
    int moved = 0;
    double disp2 = 0.0;
    if (p < S.n) {
        double inv = inv1;
        if (inv1 > 0.0) {
            scene = 0;
        } else {
            scene = S.k(I_SCENE, p);
            inv = sp[scene].inv;
        }
        double cx = S.f(F_X, p), cy = S.f(F_Y, p), cz = S.f(F_Z, p);
        if (offs) {
            const V3 o = rot(S.f(F_QW, p), S.f(F_QX, p), S.f(F_QY, p), S.f(F_QZ, p),
                             mk(S.f(F_BX, p), S.f(F_BY, p), S.f(F_BZ, p)));
            S.f(F_OFX, p) = o.x; S.f(F_OFY, p) = o.y; S.f(F_OFZ, p) = o.z;
            cx = cx + o.x; cy = cy + o.y; cz = cz + o.z;
            S.f(F_CPX, p) = cx; S.f(F_CPY, p) = cy; S.f(F_CPZ, p) = cz;
        }
        if (refC) {
            const double ex = cx - refC[p], ey = cy - refC[S.n + p], ez = cz - refC[2 * S.n + p];
            disp2 = ex * ex + ey * ey + ez * ez;
            moved = disp2 > halfSkin2 ? 1 : 0;
        }
        const double fx = floor(cx * inv), fy = floor(cy * inv), fz = floor(cz * inv);
        const double minCell = -(double)(1 << 20), maxCell = (double)((1 << 20) - 1);
        const bool ok = isfinite(fx) && isfinite(fy) && isfinite(fz) && fx >= minCell &&
                        fx <= maxCell && fy >= minCell && fy <= maxCell && fz >= minCell &&
                        fz <= maxCell;
        if (ok) {
            const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
            S.k(I_CX, p) = ix; S.k(I_CY, p) = iy; S.k(I_CZ, p) = iz;
            lo[0] = hi[0] = ix; lo[1] = hi[1] = iy; lo[2] = hi[2] = iz;
        } else {
            atomicAdd(&sc->grid_overflow, 1);
        }
        S.k(I_VALID, p) = ok ? 1 : 0;
        if (posr) {
            posr[p] = make_double4(cx, cy, cz, S.f(F_CR, p));
            const uint32_t cid = (uint32_t)S.k(I_CID, p);
            cidS[p] = cid;
            meta[p] = pack_meta(ok, lo[0], lo[1], lo[2], S.k(I_ORIG, p), cid);
        }
    }
    if (refC) atomic_max_wave(&sc->need_rebuild, moved);
    if (dispMax) {
        // the block's largest displacement since the rebuild (read by the
        // host now and then to size the skin)
        __shared__ double dmax[BLOCK / 64];
        disp2 = wave_max(disp2);
        if ((threadIdx.x & 63) == 0) dmax[threadIdx.x >> 6] = disp2;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = dmax[0];
            for (int w = 1; w < BLOCK / 64; ++w) m = fmax(m, dmax[w]);
            dispMax[blockIdx.x] = m;
        }
    }
    __shared__ int red[BLOCK / 64][8];
    const int wave = threadIdx.x >> 6;
    {
        const int smin = wave_min(scene < 0 ? INT_MAX : scene), smax = wave_max(scene);
        if ((threadIdx.x & 63) == 0) { red[wave][6] = smin; red[wave][7] = smax; }
    }
    for (int a = 0; a < 3; ++a) {
        const int l = wave_min(lo[a]), h = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) { red[wave][a] = l; red[wave][3 + a] = h; }
    }
    __syncthreads();
    int smin = red[0][6], smax = red[0][7];
    for (int w = 1; w < BLOCK / 64; ++w) { smin = min(smin, red[w][6]); smax = max(smax, red[w][7]); }
    if (smin == smax) {
        if (threadIdx.x < 6) {
            int v = red[0][threadIdx.x];
            for (int w = 1; w < BLOCK / 64; ++w)
                v = threadIdx.x < 3 ? min(v, red[w][threadIdx.x]) : max(v, red[w][threadIdx.x]);
            part[(long)blockIdx.x * 6 + threadIdx.x] = v;
        }
        if (threadIdx.x == 0) partScene[blockIdx.x] = smin;
    } else {
        // a block over several scenes: each wave reduces every scene it
        // holds and adds it with one set of atomics (lanes of the same scene
        // are contiguous, so a wave holds few)
        const int lane = threadIdx.x & 63;
        const int wmin = wave_min(scene < 0 ? INT_MAX : scene), wmax = wave_max(scene);
        for (int sc_ = wmin; sc_ <= wmax; ++sc_) {
            const bool mine = scene == sc_;
            int v[6];
            for (int a = 0; a < 3; ++a) {
                v[a] = wave_min(mine ? lo[a] : INT_MAX);
                v[3 + a] = wave_max(mine ? hi[a] : INT_MIN);
            }
            if (lane == 0 && v[0] <= v[3])
                for (int a = 0; a < 3; ++a) {
                    atomicMin(&sceneAcc[sc_ * 6 + a], v[a]);
                    atomicMax(&sceneAcc[sc_ * 6 + 3 + a], v[3 + a]);
                }
        }
        if (threadIdx.x == 0) partScene[blockIdx.x] = -1;
    }
}

// The largest of the blocks' displacements (k_cells' dispMax) into *out.
__global__ void __launch_bounds__(BLOCK) k_disp_max(const double* dispMax, int nblocks, double* out)
{
    __shared__ double w[BLOCK / 64];
    double m = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += BLOCK) m = fmax(m, dispMax[b]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < BLOCK / 64; ++k) m = fmax(m, w[k]);
        *out = m;
    }
}

template <typename T>
__device__ __forceinline__ T block_exclusive_scan_t(T v, T* total)
{
    __shared__ T wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    T before = 0, all = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Folds each scene's box (one block per scene) and sizes its grid; the last
// block to finish numbers the scenes' cells consecutively and sizes the
// global cell table.
__global__ void __launch_bounds__(BLOCK) k_grid_setup(Scalars* sc, const SceneParams* sp, const int* part,
                                                      const int* partScene, const int* sceneAcc,
                                                      SceneGrid* grid, int nscenes, unsigned long long cap,
                                                      unsigned* done)
{
    const int s = blockIdx.x;
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = sceneAcc[s * 6 + a]; hi[a] = sceneAcc[s * 6 + 3 + a]; }
    const long b0 = sp[s].row0 / BLOCK, b1 = sp[s].row1 > sp[s].row0 ? (sp[s].row1 - 1) / BLOCK : b0 - 1;
    for (long b = b0 + threadIdx.x; b <= b1; b += BLOCK)
        if (partScene[b] == s)
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], part[b * 6 + a]);
                hi[a] = max(hi[a], part[b * 6 + 3 + a]);
            }
    __shared__ int red[BLOCK / 64][6];
    __shared__ bool last;
    const int wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        const int l = wave_min(lo[a]), h = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) { red[wave][a] = l; red[wave][3 + a] = h; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long nc = 1ull;
        SceneGrid& g = grid[s];
        for (int a = 0; a < 3; ++a) {
            int l = red[0][a], h = red[0][3 + a];
            for (int w = 1; w < BLOCK / 64; ++w) { l = min(l, red[w][a]); h = max(h, red[w][3 + a]); }
            if (l > h) { l = 0; h = 0; }
            g.lo[a] = l;
            g.hi[a] = h;
            g.dims[a] = (unsigned long long)((long long)h - l + 1);
            nc *= g.dims[a];
        }
        g.ncells = nc;
        __threadfence();
        last = atomicAdd(done, 1u) == (unsigned)nscenes - 1u;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    unsigned long long carry = 0ull;
    bool wrapped = false;
    for (int b = 0; b < nscenes; b += BLOCK) {
        const int i = b + threadIdx.x;
        const unsigned long long v = i < nscenes ? grid[i].ncells : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_exclusive_scan_t(v, &tot);
        if (i < nscenes) grid[i].base = carry + ex;
        if (carry + tot < carry) wrapped = true;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        if (wrapped) atomicAdd(&sc->grid_overflow, 1);
        sc->ncells = carry;
        const unsigned long long top = carry - 1ull;
        sc->nbits = top == 0ull ? 0 : 64 - __clzll((long long)top);
        sc->dense = (carry + 1ull <= cap) ? 1 : 0;
        *done = 0u;
    }
}

__device__ __forceinline__ unsigned long long rel_cell(const SceneGrid& g, int cx, int cy, int cz)
{
    return g.base + ((unsigned long long)(cx - g.lo[0]) * g.dims[1] + (unsigned long long)(cy - g.lo[1])) *
                        g.dims[2] +
           (unsigned long long)(cz - g.lo[2]);
}

// Global linear cell index (scene-major, then the reference's x, y, z key
// order within the scene).  A particle that failed the range check keeps its
// scene's first cell (the step reports the error).
__global__ void __launch_bounds__(BLOCK) k_keys(Soa S, const SceneGrid* grid, unsigned long long* key,
                                                int* val)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    const SceneGrid& g = grid[S.k(I_SCENE, p)];
    key[p] = S.k(I_VALID, p) ? rel_cell(g, S.k(I_CX, p), S.k(I_CY, p), S.k(I_CZ, p)) : g.base;
    val[p] = (int)p;
}

// ------------------------------------------------------------ radix sort
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = BLOCK * RS_ITEMS;

// A pass over bits [shift, shift+8) runs only if the key has significant
// bits there; *nbits is computed on the device, so no host round trip.
__device__ __forceinline__ bool pass_active(const int* nbits, int shift)
{
    return shift < *nbits;
}

__global__ void __launch_bounds__(BLOCK) k_rs_hist(const unsigned long long* key, long n, int shift,
                                                   const int* nbits, unsigned* counts, int nblocks)
{
    if (!pass_active(nbits, shift)) return;
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long e = base + (long)r * BLOCK + threadIdx.x;
        if (e < n) atomicAdd(&h[(key[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (wave prefix sums with
// shuffles, then the four wave totals through LDS).  Returns the exclusive
// prefix; *total receives the block sum.
constexpr int SCAN_ITEMS = 8;  // values per thread in the one-block scans
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* total)
{
    __shared__ unsigned wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned before = 0, all = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Block-wide maximum (wave maxima through LDS); every thread of the block
// calls it and receives the result.
__device__ __forceinline__ int block_max(int v)
{
    __shared__ int wmax[BLOCK / 64];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    int m = wmax[0];
    for (int w = 1; w < BLOCK / 64; ++w) m = max(m, wmax[w]);
    __syncthreads();
    return m;
}

// The block's sum of the rows' contact counts, for the list offsets
// (bsum[b] covers rows b * BLOCK ...), and the block's history high-water
// mark in bsum[gridDim.x + b] (k_off_scan folds both, so no contact counter
// is an atomic); every thread of the block calls it.
__device__ __forceinline__ void block_count(int stored, int high, unsigned* bsum, long b)
{
    unsigned tot;
    block_exclusive_scan((unsigned)stored, &tot);
    const int h = block_max(high);
    if (threadIdx.x == 0) {
        bsum[b] = tot;
        bsum[gridDim.x + b] = (unsigned)h;
    }
}

// atomicMax of the block's maximum of v (one atomic per block, none if 0).
__device__ __forceinline__ void atomic_max_block(int* dst, int v)
{
    v = block_max(v);
    if (threadIdx.x == 0 && v) atomicMax(dst, v);
}

// Digit-major counts [256][nblocks]: block d turns row d into exclusive
// offsets within the digit and records the digit total.
__global__ void __launch_bounds__(BLOCK) k_rs_scan_rows(unsigned* counts, int nblocks, int shift,
                                                        const int* nbits, unsigned* digitTotal)
{
    if (!pass_active(nbits, shift)) return;
    unsigned* row = counts + (long)blockIdx.x * nblocks;
    unsigned carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += BLOCK) {
        const int b = b0 + threadIdx.x;
        const unsigned v = b < nblocks ? row[b] : 0u;
        unsigned tot;
        const unsigned ex = block_exclusive_scan(v, &tot);
        if (b < nblocks) row[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) digitTotal[blockIdx.x] = carry;
}

// Exclusive scan of the 256 digit totals into digit bases.
__global__ void __launch_bounds__(BLOCK) k_rs_scan_base(unsigned* digitTotal, int shift, const int* nbits)
{
    if (!pass_active(nbits, shift)) return;
    unsigned tot;
    const unsigned ex = block_exclusive_scan(digitTotal[threadIdx.x], &tot);
    __syncthreads();
    digitTotal[256 + threadIdx.x] = ex;
}

// After the pass slots the data sits in the second buffer when an odd number
// of passes ran; move it back so callers always read the first buffer.
__global__ void __launch_bounds__(BLOCK) k_rs_copyback(unsigned long long* ka, int* va,
                                                       const unsigned long long* kb, const int* vb,
                                                       long n, const int* nbits)
{
    const int passes = (*nbits + 7) / 8;
    if ((passes & 1) == 0) return;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
        ka[i] = kb[i];
        va[i] = vb[i];
    }
}

__global__ void __launch_bounds__(BLOCK) k_rs_scatter(const unsigned long long* kin, const int* vin,
                                                      unsigned long long* kout, int* vout, long n,
                                                      int shift, const int* nbits,
                                                      const unsigned* counts, int nblocks,
                                                      const unsigned* digitBase)
{
    if (!pass_active(nbits, shift)) return;
    __shared__ unsigned base[256];
    __shared__ unsigned wc[BLOCK / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = digitBase[tid] + counts[(long)tid * nblocks + blockIdx.x];
    const long tile = (long)blockIdx.x * RS_TILE;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long e = tile + (long)r * BLOCK + tid;
        const bool valid = e < n;
        const unsigned long long k = valid ? kin[e] : 0ull;
        const unsigned d = (unsigned)((k >> shift) & 255u);
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const unsigned long long b = __ballot(valid && on);
            peers &= on ? b : ~b;
        }
        const int rank = __popcll(peers & below);
        for (int w = 0; w < BLOCK / 64; ++w) wc[w][tid] = 0;
        __syncthreads();
        if (valid && rank == 0) wc[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned pos = base[d] + (unsigned)rank;
            for (int w = 0; w < wave; ++w) pos += wc[w][d];
            kout[pos] = k;
            vout[pos] = vin[e];
        }
        __syncthreads();
        unsigned add_ = 0;
        for (int w = 0; w < BLOCK / 64; ++w) add_ += wc[w][tid];
        base[tid] += add_;
        __syncthreads();
    }
}

// Within each run of equal keys, order the payload by input row.
__global__ void __launch_bounds__(BLOCK) k_fixup(const unsigned long long* key, int* val, long n,
                                                 const int* orig)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n) return;
    const unsigned long long k = key[t];
    if (t > 0 && key[t - 1] == k) return;
    long e = t + 1;
    while (e < n && key[e] == k) ++e;
    for (long a = t + 1; a < e; ++a) {
        const int v = val[a];
        const int o = orig[v];
        long b = a - 1;
        while (b >= t && orig[val[b]] > o) {
            val[b + 1] = val[b];
            --b;
        }
        val[b + 1] = v;
    }
}

// Gather every particle field into sorted order (new physical order).
__global__ void __launch_bounds__(BLOCK) k_permute(Soa src, Soa dst, int* sorted)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= src.n) return;
    const long p = sorted[t];
    for (int f = 0; f < NDF; ++f) dst.f(f, t) = src.f(f, p);
    for (int f = 0; f < NIF; ++f) dst.k(f, t) = src.k(f, p);
}

__global__ void __launch_bounds__(BLOCK) k_identity(int* sorted, long n)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t < n) sorted[t] = (int)t;
}

// Row layout between candidate-list rebuilds: the rows are put in Morton
// order of their cells (16 bits per axis, x y z interleaved) under their
// scene index (so a scene's rows stay contiguous, as the per-scene block
// reductions want), rows of one cell in the grid's order.  Neighbouring cells in all
// three directions then lie close in memory, so the partner gathers of the
// list steps stay in L2.  Any row order gives the same results: the contact
// order comes from the reference's key, the history moves with the rows, and
// the sorted order of the grid is kept in key/sorted (sorted position ->
// row) for the rebuild scan and the debug output.
__device__ __forceinline__ unsigned long long spread3(unsigned v)
{
    unsigned long long x = v & 0xFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void __launch_bounds__(BLOCK) k_morton_keys(Soa S, const int* sorted, const SceneGrid* grid,
                                                       unsigned long long* key, int* val, Scalars* sc, int bits)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t == 0) sc->mbits = bits;
    if (t >= S.n) return;
    const long p = sorted[t];
    // rows outside the hash range last within their scene
    unsigned long long k = ((unsigned long long)S.k(I_SCENE, p) << 48) | ((1ull << 48) - 1ull);
    if (S.k(I_VALID, p)) {
        const int scene = S.k(I_SCENE, p);
        const SceneGrid& g = grid[scene];
        const unsigned x = (unsigned)(S.k(I_CX, p) - g.lo[0]), y = (unsigned)(S.k(I_CY, p) - g.lo[1]),
                       z = (unsigned)(S.k(I_CZ, p) - g.lo[2]);
        k = ((unsigned long long)scene << 48) | (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
    }
    key[t] = k;
    val[t] = (int)t;
}

// perm[j] = the old row that goes to row j; sorted[t] = the new row of the
// row at sorted position t.
__global__ void __launch_bounds__(BLOCK) k_morton_perm(const int* sorted, const int* order, int* perm, long n)
{
    const long j = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) perm[j] = sorted[order[j]];
}
__global__ void __launch_bounds__(BLOCK) k_morton_rank(const int* order, int* sorted, long n)
{
    const long j = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) sorted[order[j]] = (int)j;
}

__device__ __forceinline__ long lower_bound_u64(const unsigned long long* k, long n,
                                                unsigned long long key)
{
    long first = 0, length = n;
    while (length > 0) {
        const long half = length >> 1, mid = first + half;
        if (k[mid] < key) { first = mid + 1; length -= half + 1; }
        else length = half;
    }
    return first;
}

__global__ void __launch_bounds__(BLOCK) k_cellstart(const unsigned long long* key, long n,
                                                     const Scalars* sc, unsigned* start,
                                                     unsigned long long cap)
{
    if (!sc->dense) return;
    const unsigned long long nc = sc->ncells;
    for (unsigned long long c = (unsigned long long)blockIdx.x * BLOCK + threadIdx.x; c <= nc && c < cap;
         c += (unsigned long long)gridDim.x * BLOCK)
        start[c] = (unsigned)lower_bound_u64(key, n, c);
}

// Packed contact view of every row, by physical index (the fields
// particleContact reads for both sides, contact:585-609), in two parts: the
// per-step kinematics (component centre, endpoint and half-step velocities,
// world angular velocities) with the mass in one 128-byte line, and the
// row's constants (radius, status, scene) in 16 bytes, so a neighbour costs
// one line and a shared one.
struct alignas(128) Rec {
    double px, py, pz, vx, vy, vz, hx, hy, hz, wx, wy, wz, ux, uy, uz, mass;
};
static_assert(sizeof(Rec) == 128, "contact record layout");
struct alignas(16) RecAux {
    double r;
    int status;
    int scene;
};
static_assert(sizeof(RecAux) == 16, "contact record layout");
struct Recs {
    Rec* main;
    RecAux* aux;
};

__device__ __forceinline__ Body rec_body(const Recs& R, long p)
{
    const Rec& q = R.main[p];
    const RecAux& x = R.aux[p];
    Body B;
    B.pos = mk(q.px, q.py, q.pz);
    B.vel = mk(q.vx, q.vy, q.vz);
    B.vh = mk(q.hx, q.hy, q.hz);
    B.w = mk(q.wx, q.wy, q.wz);
    B.wh = mk(q.ux, q.uy, q.uz);
    B.mass = q.mass;
    B.r = x.r;
    B.status = x.status;
    B.cid = 0u;  // not read by the laws (slots are found by the scans)
    B.off = zero3();
    return B;
}

// Kinematic part of the contact record of a zero-offset row from its state
// values (what k_pack writes for it; the advance kernels call it with the
// values they just stored, saving k_pack's second pass over the state).  The
// constant part stays as the last k_pack wrote it: rows only change place
// in a reorder or a reload, and those clear the records' freshness.
__device__ __forceinline__ void put_rec(const Recs& R, long p, V3 pos, V3 vel, V3 vh, double qw, double qx,
                                        double qy, double qz, V3 w, V3 wh, double mass)
{
    Rec q;
    const V3 ww = rot(qw, qx, qy, qz, w), wwh = rot(qw, qx, qy, qz, wh);
    q.px = pos.x; q.py = pos.y; q.pz = pos.z;
    q.vx = vel.x; q.vy = vel.y; q.vz = vel.z;
    q.hx = vh.x; q.hy = vh.y; q.hz = vh.z;
    q.wx = ww.x; q.wy = ww.y; q.wz = ww.z;
    q.ux = wwh.x; q.uy = wwh.y; q.uz = wwh.z;
    q.mass = mass;
    R.main[p] = q;
}
__device__ __forceinline__ void put_rec(const Recs& R, const Soa& S, long p, V3 pos, V3 vel, V3 vh, double qw,
                                        double qx, double qy, double qz, V3 w, V3 wh)
{
    put_rec(R, p, pos, vel, vh, qw, qx, qy, qz, w, wh, S.f(F_MASS, p));
}

__global__ void __launch_bounds__(BLOCK) k_pack(Soa S, bool offs, Recs R)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    const Body B = load_body(S, p, offs);
    Rec q;
    q.px = B.pos.x; q.py = B.pos.y; q.pz = B.pos.z;
    q.vx = B.vel.x; q.vy = B.vel.y; q.vz = B.vel.z;
    q.hx = B.vh.x; q.hy = B.vh.y; q.hz = B.vh.z;
    q.wx = B.w.x; q.wy = B.w.y; q.wz = B.w.z;
    q.ux = B.wh.x; q.uy = B.wh.y; q.uz = B.wh.z;
    q.mass = B.mass;
    R.main[p] = q;
    RecAux x;
    x.r = B.r;
    x.status = B.status;
    x.scene = S.k(I_SCENE, p);
    R.aux[p] = x;
}

// Contact centre, radius and component id in sorted order (what the
// candidate scan reads per candidate).
// With clumps it also records each row's sorted position and the owner of
// the t-th sorted row (the scan skips partners of the same owner).
__global__ void __launch_bounds__(BLOCK) k_posr(Soa S, const int* sorted, bool offs, double4* posr,
                                                uint32_t* cidS, int4* meta, const int* owner, int* ownS,
                                                int* rank)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= S.n) return;
    const long p = sorted[t];
    if (ownS) {
        ownS[t] = owner[p];
        rank[p] = (int)t;
    }
    double4 v;
    if (offs) { v.x = S.f(F_CPX, p); v.y = S.f(F_CPY, p); v.z = S.f(F_CPZ, p); }
    else { v.x = S.f(F_X, p); v.y = S.f(F_Y, p); v.z = S.f(F_Z, p); }
    v.w = S.f(F_CR, p);
    posr[t] = v;
    cidS[t] = (uint32_t)S.k(I_CID, p);
    if (meta)
        meta[t] = pack_meta(S.k(I_VALID, p), S.k(I_CX, p), S.k(I_CY, p), S.k(I_CZ, p), S.k(I_ORIG, p),
                            (uint32_t)S.k(I_CID, p));
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  This maps the hardware block index to a logical one so that every XCD
// works on one contiguous eighth of the range (neighbouring particles, and
// their neighbours' data, stay in one L2).
__device__ __forceinline__ long xcd_block(long b, long nb)
{
    const long q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
    return x * q + (x < r ? x : r) + i;
}

// History slot acquisition (acquireHistorySlot, contact:190-215) on an
// occupancy mask.  Returns the slot or -1 (overflow).
template <typename M>
__device__ __forceinline__ int acquire(M& occ, M& cur, uint32_t* keys, double* vals, long vstride, int slots,
                                       uint32_t key, int& high, int& overflow)
{
    M m = occ;
    while (m) {
        const int b = (int)__ffsll((unsigned long long)m) - 1;
        if (keys[b] == key) {
            cur |= (M)1 << b;
            high = max(high, b + 1);
            return b;
        }
        m &= m - 1;
    }
    const M full = slots >= (int)(8 * sizeof(M)) ? ~(M)0 : (((M)1 << slots) - 1);
    const M freem = ~occ & full;
    if (!freem) {
        ++overflow;
        return -1;
    }
    const int b = (int)__ffsll((unsigned long long)freem) - 1;
    occ |= (M)1 << b;
    cur |= (M)1 << b;
    keys[b] = key;
    double* v = vals + (long)b * vstride;
    for (int q = 0; q < 9; ++q) v[q] = 0.0;
    high = max(high, b + 1);
    return b;
}

// What a differentiable run records of one step for the reverse sweep
// (every pointer null when nothing is recorded).
struct Tape {
    unsigned long long* ppOcc;  // [n] particle-particle occupancy after expiry
    unsigned* pwOcc;            // [n] particle-wall occupancy after expiry
    int2* pair;                 // [contacts] (row, partner row << SLOT_BITS | slot)
    double* pairPrev;           // [contacts][9] slot values the law read
    int4* wall;                 // [wall contacts] (row, wall within the scene, slot, 0)
    double* wallPrev;           // [wall contacts][9]
    unsigned* wallCount;        // wall-contact append counter (counted even with wall == null)
    double* plateAcc;           // [n] force of the plate wall on the particle (z), summed over steps
    int plate;                  // plate wall index within each scene
};

struct LawArgs {
    Soa S;                       // owners
    Soa R;                       // contact rows: the owners, or their components (clumps)
    const int* rank;             // [rows] sorted position of each contact row (clumps)
    const int* ownerStart;       // [owners + 1] first component of each owner (clumps)
    const double4* posr;         // [rows] centre and radius in cell order
    Recs rec;
    const int* sorted;
    const int* ncount;
    const unsigned* off;
    const double* pairF;
    const double* rowF;          // [rows][6] contact sums of rows summed in k_pairs (single spheres)
    History H;
    Walls W;
    TriSet T;
    Outputs O;
    Scalars* sc;
    Material pt;                 // particle-triangle (single scene with bodies)
    const Material* matPW;       // particle-wall, per scene
    const SceneParams* sp;
    bool first;                  // first force evaluation: no history advance (dt = 0)
    double g[3];
    int gravity;
    int universal;
    const int* byorig;
    int step;
    bool offs;
    int nb;
    Tape tape;
    bool counters;  // per-step statistics (contacts, high-water marks) wanted
    double finishDt, nextDt;  // k_law<.., true>: this step's finish and the next step's advance
    int* sceneAcc;            // k_law<.., true>: per-scene grid boxes reset for the next evaluation
    int nscenes;
    bool identRows;  // sorted is the identity (candidate-list steps): k_law reads no index
    bool oneScene;   // a single scene: its walls are all of W
};

// History lookup with the first 16 keys of the owner held in registers (kept
// in step with memory when a slot is reassigned): a hit among them needs no
// dependent load.  Same result as acquire().  find_pp is its lookup among the
// occupied slots (-1: a new contact), new_pp its allocation of a new one.
__device__ __forceinline__ int find_pp(unsigned long long occ, const uint32_t (&k16)[16], const uint32_t* keys,
                                       uint32_t key)
{
    unsigned match = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) match |= (k16[i] == key ? 1u : 0u) << i;
    match &= (unsigned)(occ & 0xFFFFull);
    if (match) return __ffs(match) - 1;
    unsigned long long m = occ & ~0xFFFFull;
    while (m) {
        const int c = (int)__ffsll(m) - 1;
        if (keys[c] == key) return c;
        m &= m - 1;
    }
    return -1;
}
__device__ __forceinline__ int new_pp(unsigned long long& occ, unsigned long long& cur, uint32_t (&k16)[16],
                                      uint32_t* keys, double* vals, long vstride, uint32_t key, int& high,
                                      int& overflow)
{
    const unsigned long long freem = ~occ;
    if (!freem) {
        ++overflow;
        return -1;
    }
    const int b = (int)__ffsll(freem) - 1;
    occ |= 1ull << b;
    cur |= 1ull << b;
    keys[b] = key;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (i == b) k16[i] = key;  // keep the register copy current (a reused slot)
    double* v = vals + (long)b * vstride;
    for (int q = 0; q < 9; ++q) v[q] = 0.0;
    high = max(high, b + 1);
    return b;
}
__device__ __forceinline__ int acquire_pp(unsigned long long& occ, unsigned long long& cur,
                                          uint32_t (&k16)[16], uint32_t* keys, double* vals, long vstride,
                                          uint32_t key, int& high, int& overflow)
{
    const int b = find_pp(occ, k16, keys, key);
    if (b >= 0) {
        cur |= 1ull << b;
        high = max(high, b + 1);
        return b;
    }
    return new_pp(occ, cur, k16, keys, vals, vstride, key, high, overflow);
}

// History expiry at the start of the owner's step (expireHistory,
// contact:217-229; the caller reads the row's expired occupancy and first 16
// keys ahead of its scan, row_history) and one slot per contact in list order
// (acquireHistorySlot, contact:190-215).  nbr[c][t] holds the sorted position
// of the c-th contact on entry and (partner row << SLOT_BITS | slot) on
// return, contacts without a slot (overflow) dropped; returns the entries
// kept.  The contacts' ids and rows are gathered four at a time.
template <typename Order>
__device__ __forceinline__ int assign_batched(long t, long n, long o, int found, unsigned long long occ,
                                              uint32_t (&k16)[16], const Order& sorted, const uint32_t* cidS,
                                              const History& H, int step, int* nbr, int& high, int& over)
{
    unsigned long long cur = 0;
    uint32_t* keys = H.ppKey + o * PP_SLOTS;
    double* vals = pp_slot(H, o, 0);
    int stored = 0;
    for (int c0 = 0; c0 < found; c0 += 4) {
        int kq[4], rq[4];
        uint32_t cq[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) kq[j] = nbr[(long)min(c0 + j, found - 1) * n + t];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cq[j] = cidS[kq[j]];
            rq[j] = sorted[kq[j]];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (c0 + j >= found) break;
            const int slot = acquire_pp(occ, cur, k16, keys, vals, H.rows * 9, cq[j] + 1u, high, over);
            if (slot < 0) continue;  // slot overflow: counted, no force
            nbr[(long)stored * n + t] = (rq[j] << SLOT_BITS) | slot;
            ++stored;
        }
    }
    H.ppOcc[o] = occ;
    H.ppSeen[(long)(step & 1) * n + o] = cur;
    return stored;
}

// The row's expired occupancy and first 16 history keys.
__device__ __forceinline__ unsigned long long row_history(const History& H, long n, long o, int step,
                                                          uint32_t (&k16)[16])
{
    const uint4* kv = reinterpret_cast<const uint4*>(H.ppKey + o * PP_SLOTS);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint4 v = kv[i];
        k16[4 * i] = v.x; k16[4 * i + 1] = v.y; k16[4 * i + 2] = v.z; k16[4 * i + 3] = v.w;
    }
    return H.ppOcc[o] & H.ppSeen[(long)((step + 1) & 1) * n + o];
}

// More overlapping partners than a contact list holds (more than there are
// history slots, so the reference runs out of slots): expiry, then each
// contact's slot acquired while `visit` walks the partners again in
// candidate order, as the reference's loop does; a contact left without a
// slot is counted (pp_overflow) and dropped.  At most PP_SLOTS contacts can
// hold a slot, so the list never overflows.
// The identity as a sorted order (candidate-list steps: rows are their own
// sorted positions), for the helpers that map a sorted position to a row.
struct IdentOrder {
    __device__ int operator[](long k) const { return (int)k; }
};
__device__ __forceinline__ IdentOrder ident_order() { return IdentOrder{}; }

template <typename Order, typename Visit>
__device__ __forceinline__ int assign_inline(const Soa& S, long t, long p, const Order& sorted, const uint32_t* cidS,
                             const History& H, int step, int* nbr, unsigned long long* occTape, int& high,
                             int& over, Visit&& visit)
{
    const long n = S.n;
    const long o = S.k(I_HIST, p);
    unsigned long long occ = H.ppOcc[o] & H.ppSeen[(long)((step + 1) & 1) * n + o];
    if (occTape) occTape[o] = occ;
    unsigned long long cur = 0;
    uint32_t* keys = H.ppKey + o * PP_SLOTS;
    double* vals = pp_slot(H, o, 0);
    int stored = 0;
    visit([&](long k) {
        const int slot = acquire(occ, cur, keys, vals, H.rows * 9, PP_SLOTS, cidS[k] + 1u, high, over);
        if (slot < 0) return;
        nbr[(long)stored * n + t] = (sorted[k] << SLOT_BITS) | slot;
        ++stored;
    });
    H.ppOcc[o] = occ;
    H.ppSeen[(long)(step & 1) * n + o] = cur;
    return stored;
}

// The overlapping candidates of sorted row t in the reference's cell and
// entry order (see k_scan), handed to f one at a time.
template <bool CLUMP, typename F>
__device__ __forceinline__ void scan_overlaps(const Soa& S, long t, long p, int mine, const double4* posr,
                                              const int* ownS, const unsigned long long* key,
                                              const unsigned* start, const SceneGrid* grid, bool dense, F&& f)
{
    const long n = S.n;
    const double4 me = posr[t];
    const SceneGrid& g = grid[S.k(I_SCENE, p)];
    const long X = (long)g.dims[0], Y = (long)g.dims[1], Z = (long)g.dims[2];
    const long cx = S.k(I_CX, p) - g.lo[0], cy = S.k(I_CY, p) - g.lo[1], cz = S.k(I_CZ, p) - g.lo[2];
    const long zlo = cz > 0 ? cz - 1 : 0, zhi = cz + 1 < Z ? cz + 1 : Z - 1;
    for (long dx = -1; dx <= 1; ++dx) {
        const long xx = cx + dx;
        if (xx < 0 || xx >= X) continue;
        for (long dy = -1; dy <= 1; ++dy) {
            const long yy = cy + dy;
            if (yy < 0 || yy >= Y) continue;
            const unsigned long long col = g.base + ((unsigned long long)xx * Y + yy) * Z;
            long b, e;
            if (dense) {
                b = start[col + zlo];
                e = start[col + zhi + 1];
            } else {
                b = lower_bound_u64(key, n, col + zlo);
                e = lower_bound_u64(key, n, col + zhi + 1);
            }
            for (long k = b; k < e; ++k) {
                if (CLUMP ? ownS[k] == mine : k == t) continue;
                const double4 c = posr[k];
                const double dx_ = c.x - me.x, dy_ = c.y - me.y, dz_ = c.z - me.z;
                const double d2 = dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
                const double rs = me.w + c.w;
                if (d2 >= rs * rs || d2 <= 1.0e-28) continue;
                f(k);
            }
        }
    }
}

// 27-cell candidate scan (particleContact's overlap test, contact:579-591, in
// the reference's cell and entry order: the nine (dx, dy) columns, each one
// contiguous run of the sorted array over dz = -1..1) followed by the
// particle-particle history: expiry at the start of the owner's step
// (expireHistory, contact:217-229) and one slot per contact in candidate
// order (acquireHistorySlot, contact:190-215).  The scan first records the
// overlapping candidates, so the slot lookups run once per contact rather
// than once per candidate iteration of the wave.  Leaves the owner's contact
// list as entries (partner row << SLOT_BITS | slot), k-major.  CLUMP: rows
// are clump components; partners of the row's own owner are skipped
// (computeContactsKernel compares owners, contact:1138).
//
// The block's cell neighbourhood is staged in LDS.  A block's 256 sorted
// rows lie in a few (x, y) columns of the grid (segments, each a run of z
// cells).  The scan runs in three passes, dx = -1, 0, 1; each
// pass first copies, for every segment and dy = -1..1, the sorted range of
// column (x + dx, y + dy) over the segment's z cells +-1 into LDS (centre and
// radius; with clumps also the owner), then every row tests its three
// columns of that pass from LDS.  The passes and columns come in the
// reference's cell order, so the contacts are found in the reference's
// order.  A block with more segments than STAGE_SEGS, a pass with more
// entries than STAGE_CAP, or a grid without the dense cell table reads its
// candidates from HBM instead.
constexpr int STAGE_SEGS = 4;
constexpr int STAGE_CAP = 1024;
template <bool CLUMP>
__global__ void __launch_bounds__(BLOCK) k_scan(Soa S, const int* sorted, const double4* posr,
                                                    const uint32_t* cidS, const unsigned long long* key,
                                                    const unsigned* start, const SceneGrid* grid, Scalars* sc,
                                                    History H, int step, int* nbr, int* ncount,
                                                    unsigned long long* occTape, bool counters, const int* ownS,
                                                    unsigned* bsum)
{
    __shared__ double4 stage[STAGE_CAP];
    __shared__ int stageOwn[CLUMP ? STAGE_CAP : 1];
    __shared__ unsigned long long colOf[BLOCK];
    __shared__ int segX[STAGE_SEGS], segY[STAGE_SEGS], segZa[STAGE_SEGS], segZb[STAGE_SEGS], segScene[STAGE_SEGS];
    __shared__ long rngB[3 * STAGE_SEGS];
    __shared__ int rngN[3 * STAGE_SEGS], rngOff[3 * STAGE_SEGS + 1];
    __shared__ int nseg, staged;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    const bool dense = sc->dense;
    long p = -1, hrow = 0;
    int mine = 0;
    bool valid = false;
    long cx = 0, cy = 0, cz = 0, X = 0, Y = 0, Z = 0;
    unsigned long long base = 0;
    unsigned long long occ = 0;
    uint32_t k16[16];
    if (t < n) {
        p = sorted[t];
        mine = CLUMP ? ownS[t] : 0;
        valid = S.k(I_VALID, p) != 0;
        // the row's history state, read ahead of the scan
        hrow = S.k(I_HIST, p);
        occ = row_history(H, n, hrow, step, k16);
        if (valid) {
            const SceneGrid& g = grid[S.k(I_SCENE, p)];
            X = (long)g.dims[0]; Y = (long)g.dims[1]; Z = (long)g.dims[2];
            cx = S.k(I_CX, p) - g.lo[0]; cy = S.k(I_CY, p) - g.lo[1]; cz = S.k(I_CZ, p) - g.lo[2];
            base = g.base;
        }
    }
    // segments: runs of rows in one (scene, x, y) column
    const unsigned long long col = valid ? base + ((unsigned long long)cx * Y + cy) * Z : ~0ull;
    colOf[threadIdx.x] = col;
    if (threadIdx.x == 0) nseg = 0;
    __syncthreads();
    const bool head = valid && (threadIdx.x == 0 || colOf[threadIdx.x - 1] != col);
    const bool tail = valid && (threadIdx.x == BLOCK - 1 || colOf[threadIdx.x + 1] != col);
    unsigned nsegs;
    const unsigned si = block_exclusive_scan(head ? 1u : 0u, &nsegs);
    const int seg = head ? (int)si : (int)si - 1;  // the row's segment (valid rows)
    if (head && si < STAGE_SEGS) {
        segX[si] = (int)cx; segY[si] = (int)cy; segZa[si] = (int)cz;
        segScene[si] = S.k(I_SCENE, p);
    }
    if (tail && seg < STAGE_SEGS) segZb[seg] = (int)cz;
    const bool useLds = dense && nsegs <= (unsigned)STAGE_SEGS;
    __syncthreads();
    int found = 0, stored = 0, high = 0, over = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        if (useLds) {
            // the pass's ranges: segment-major, dy = -1, 0, 1
            if (threadIdx.x < 3 * (int)nsegs) {
                const int sg = threadIdx.x / 3, dy = threadIdx.x % 3 - 1;
                const SceneGrid& g = grid[segScene[sg]];
                const long gX = (long)g.dims[0], gY = (long)g.dims[1], gZ = (long)g.dims[2];
                const long xx = segX[sg] + dx, yy = segY[sg] + dy;
                long b = 0, cnt = 0;
                if (xx >= 0 && xx < gX && yy >= 0 && yy < gY) {
                    const unsigned long long c0 = g.base + ((unsigned long long)xx * gY + yy) * gZ;
                    const long zlo = segZa[sg] > 0 ? segZa[sg] - 1 : 0;
                    const long zhi = segZb[sg] + 1 < gZ ? segZb[sg] + 1 : gZ - 1;
                    b = start[c0 + zlo];
                    cnt = (long)start[c0 + zhi + 1] - b;
                }
                rngB[threadIdx.x] = b;
                rngN[threadIdx.x] = (int)cnt;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int o = 0;
                for (int r = 0; r < 3 * (int)nsegs; ++r) {
                    rngOff[r] = o;
                    o += rngN[r];
                }
                rngOff[3 * nsegs] = o;
                staged = o <= STAGE_CAP ? o : -1;
            }
            __syncthreads();
            const int total = staged;
            for (int i = threadIdx.x; i < total; i += BLOCK) {
                int r = 0;
                while (rngOff[r + 1] <= i) ++r;
                const long k = rngB[r] + (i - rngOff[r]);
                stage[i] = posr[k];
                if (CLUMP) stageOwn[i] = ownS[k];
            }
            __syncthreads();
        }
        if (valid) {
            const double4 me = posr[t];
            const long xx = cx + dx;
            if (xx >= 0 && xx < X) {
                const long zlo = cz > 0 ? cz - 1 : 0, zhi = cz + 1 < Z ? cz + 1 : Z - 1;
                // the pass's three column ranges, read together (empty off the grid)
                long rb[3], re[3];
                for (int dy = -1; dy <= 1; ++dy) {
                    const long yy = cy + dy;
                    long b = 0, e = 0;
                    if (yy >= 0 && yy < Y) {
                        const unsigned long long c0 = base + ((unsigned long long)xx * Y + yy) * Z;
                        if (dense) {
                            b = start[c0 + zlo];
                            e = start[c0 + zhi + 1];
                        } else {
                            b = lower_bound_u64(key, n, c0 + zlo);
                            e = lower_bound_u64(key, n, c0 + zhi + 1);
                        }
                    }
                    rb[dy + 1] = b;
                    re[dy + 1] = e;
                }
                for (int dy = -1; dy <= 1; ++dy) {
                    const long b = rb[dy + 1], e = re[dy + 1];
                    if (b >= e) continue;
                    // the overlap test of one candidate, in the reference's order
                    auto hit = [&](const double4& c) {
                        const double dx_ = c.x - me.x, dy_ = c.y - me.y, dz_ = c.z - me.z;
                        const double d2 = dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
                        const double rs = me.w + c.w;
                        return !(d2 >= rs * rs || d2 <= 1.0e-28);
                    };
                    if (useLds && staged >= 0) {
                        // staged range: 32-bit stage indices, sorted row = i - shift
                        const int r = seg * 3 + dy + 1;
                        const int shift = rngOff[r] - (int)rngB[r];
                        // (the row itself fails the overlap test: its distance is 0)
                        #pragma unroll 2
                        for (int i = (int)b + shift, ie = (int)e + shift; i < ie; ++i) {
                            if (CLUMP && stageOwn[i] == mine) continue;
                            if (!hit(stage[i])) continue;
                            if (found < MAX_LIST) nbr[(long)found * n + t] = i - shift;
                            ++found;
                        }
                    } else {
                        for (long k = b; k < e; ++k) {
                            if (CLUMP && ownS[k] == mine) continue;
                            if (!hit(posr[k])) continue;
                            if (found < MAX_LIST) nbr[(long)found * n + t] = (int)k;
                            ++found;
                        }
                    }
                }
            }
        }
        if (counters && threadIdx.x == 0)
            atomicAdd(useLds && staged >= 0 ? &sc->scan_lds : &sc->scan_hbm, 1ull);
        if (useLds) __syncthreads();  // the next pass overwrites the stage
    }
    if (t < n) {
        if (found > MAX_LIST) {
            stored = assign_inline(S, t, p, sorted, cidS, H, step, nbr, occTape, high, over, [&](auto&& acq) {
                scan_overlaps<CLUMP>(S, t, p, mine, posr, ownS, key, start, grid, dense, acq);
            });
        } else {
            if (occTape) occTape[hrow] = occ;
            stored = assign_batched(t, n, hrow, found, occ, k16, sorted, cidS, H, step, nbr, high, over);
        }
        ncount[t] = stored;
    }
    block_count(stored, high, bsum, blockIdx.x);
    atomic_add_wave(&sc->pp_overflow, over);
}

// Compacts the k-major contact lists into one array ordered by owner (CSR):
// exclusive prefix sum of the list lengths in three passes (block sums, one
// block scanning the sums, per-block offsets), the last also writing each
// contact's (owner, entry) at its position.

// The scan's block sums (bsum[0, nblocks)) become block offsets; the total
// is the step's stored contact count and the block high-water marks
// (bsum[nblocks, 2 nblocks)) fold into the step's high-water mark.
__global__ void __launch_bounds__(BLOCK) k_off_scan(unsigned* bsum, int nblocks, Scalars* sc)
{
    unsigned carry = 0;
    int high = 0;
    for (long base = 0; base < nblocks; base += SCAN_TILE) {
        const long b = base + (long)threadIdx.x * SCAN_ITEMS;
        unsigned v[SCAN_ITEMS];
        unsigned sum = 0;
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = b + j < nblocks ? bsum[b + j] : 0u;
            sum += v[j];
            if (b + j < nblocks) high = max(high, (int)bsum[nblocks + b + j]);
        }
        unsigned tot;
        unsigned run = carry + block_exclusive_scan(sum, &tot);
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            if (b + j < nblocks) bsum[b + j] = run;
            run += v[j];
        }
        carry += tot;
    }
    high = block_max(high);
    if (threadIdx.x == 0) {
        sc->total_pairs = carry;
        sc->particle_contacts = (int)carry;
        sc->pp_high = high;
    }
}

// The block's entries are one contiguous run of the compacted list: they are
// put in order in LDS and written out in full lines (a block with more
// entries than the stage writes each row's run directly).
constexpr int OFF_STAGE = 2048;
__global__ void __launch_bounds__(BLOCK) k_off_apply(const int* cnt, long n, const unsigned* bsum,
                                                     const int* nbr, unsigned* off, int2* pairs)
{
    __shared__ int2 stage[OFF_STAGE];
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const int c = t < n ? cnt[t] : 0;
    unsigned tot;
    const unsigned ex = block_exclusive_scan((unsigned)c, &tot);
    const unsigned first = bsum[blockIdx.x];
    if (t < n) off[t] = first + ex;
    if (tot <= (unsigned)OFF_STAGE) {
        for (int k = 0; k < c; ++k) stage[ex + k] = make_int2((int)t, nbr[(long)k * n + t]);
        __syncthreads();
        for (unsigned i = threadIdx.x; i < tot; i += BLOCK) pairs[first + i] = stage[i];
    } else {
        for (int k = 0; k < c; ++k) pairs[first + ex + k] = make_int2((int)t, nbr[(long)k * n + t]);
    }
}

// A row's contacts are summed inside k_pairs when its list [off, off + cnt)
// lies in one aligned 64-entry chunk, i.e. in one wave's iteration.
__device__ __forceinline__ bool row_in_chunk(long base, int cnt)
{
    return cnt > 0 && (base >> 6) == ((base + cnt - 1) >> 6);
}

// particleContact (contact:591-706) for every listed contact, one thread per
// contact (grid-stride over the compacted list), so uneven contact counts do
// not idle lanes.  With rowF, a row whose list lies in one wave's chunk is
// summed here by its first contact's lane, from +0.0 in list order (the order
// and start of the owner's own sum), into rowF[row]; the other rows' contacts
// write their force and torque term to pairF for the owner to add.
// ROWS: candidate-list steps, where the list's owner index is the row and
// every row's history sits in its own row (the identity order and I_HIST of
// the rebuilds), so no index is looked up.  ONE: a single scene, whose law
// is uniform across the grid, passed by value (scalar registers).
// (kept at 3 waves per SIMD: at most 168 VGPRs)
template <bool ROWS, bool ONE>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) k_pairs(Soa S, const int* sorted, Recs rec,
                                                 const int2* pairs, const unsigned* off, const int* ncount,
                                                 History H, const Material* mats, Material mat1,
                                                 bool first, bool offs, const Scalars* sc, double* pairF,
                                                 double* rowF, int2* tapePair, double* tapePrev)
{
    __shared__ double stage[BLOCK * 6];
    const long total = sc->total_pairs;
    const int lane = threadIdx.x & 63;
    double* ws = stage + (threadIdx.x - lane) * 6;  // this wave's 64 results
    const long stride = (long)gridDim.x * BLOCK;
    long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    // the list entry of the lane's next contact is fetched one iteration
    // ahead, so each iteration starts with its gathers
    int2 next = e < total ? pairs[e] : make_int2(0, 0);
    for (long e0 = (long)blockIdx.x * BLOCK; e0 < total; e0 += stride, e += stride) {
        V3 cf = zero3(), tq = zero3();
        long row = -1;
        int len = 0;
        const int2 pr = next;
        if (e < total) {
        const long p = ROWS ? (long)pr.x : (long)sorted[pr.x];
        const long q = pr.y >> SLOT_BITS;
        const int slot = pr.y & (PP_SLOTS - 1);
        Body I = rec_body(rec, p);
        if (offs) I.off = mk(S.f(F_OFX, p), S.f(F_OFY, p), S.f(F_OFZ, p));
        const Body J = rec_body(rec, q);
        const long o = ROWS ? p : (long)S.k(I_HIST, p);
        Material m = ONE ? mat1 : mats[rec.aux[p].scene];
        if (first) m.dt = 0.0;  // the first evaluation advances no history
        // the slot's nine values are read once, kept in registers through
        // the law and written back once
        double* hv = pp_slot(H, o, slot);
        double h[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) h[k] = hv[k];
        long base = 0;
        int cnt = 0;
        if (rowF) {
            base = off[pr.x];
            cnt = ncount[pr.x];
        }
        if (e + stride < total) next = pairs[e + stride];
        const V3 delta = sub(J.pos, I.pos);
        const double d2 = dot(delta, delta);
        if (tapePair) {
            tapePair[e] = make_int2((int)p, pr.y);
            for (int k = 0; k < 9; ++k) tapePrev[e * 9 + k] = h[k];
        }
        pp_law(I, J, delta, d2, m, h, cf, tq);
#pragma unroll
        for (int k = 0; k < 9; ++k) hv[k] = h[k];
        bool whole = false;
        if (rowF) {
            whole = row_in_chunk(base, cnt);
            if (whole && e == base) {
                row = pr.x;
                len = cnt;
            }
        }
        if (!whole) {
            double* out = pairF + e * 6;
            out[0] = cf.x; out[1] = cf.y; out[2] = cf.z;
            out[3] = tq.x; out[4] = tq.y; out[5] = tq.z;
        }
        }
        if (rowF) {  // block-uniform
            wave_sync();
            double* mine = ws + lane * 6;
            mine[0] = cf.x; mine[1] = cf.y; mine[2] = cf.z;
            mine[3] = tq.x; mine[4] = tq.y; mine[5] = tq.z;
            wave_sync();
            if (row >= 0) {
                double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0, f5 = 0.0;
                for (int j = 0; j < len; ++j) {
                    const double* v = ws + (lane + j) * 6;
                    f0 = f0 + v[0]; f1 = f1 + v[1]; f2 = f2 + v[2];
                    f3 = f3 + v[3]; f4 = f4 + v[4]; f5 = f5 + v[5];
                }
                double* out = rowF + row * 6;
                out[0] = f0; out[1] = f1; out[2] = f2;
                out[3] = f3; out[4] = f4; out[5] = f5;
            }
        }
    }
}

#include "dem_verlet.inl"

__device__ __forceinline__ Tri load_tri(const TriSet& T, int t)
{
    const double* w = T.world + (long)t * 15;
    Tri r;
    r.a = mk(w[3], w[4], w[5]);
    r.b = mk(w[6], w[7], w[8]);
    r.c = mk(w[9], w[10], w[11]);
    r.n = mk(w[12], w[13], w[14]);
    r.body = T.body[t];
    return r;
}

__device__ __forceinline__ Pose load_pose(const TriSet& T, int b)
{
    const double* q = T.pose + (long)b * 13;
    Pose P;
    P.c = mk(q[0], q[1], q[2]);
    P.qw = q[3]; P.qx = q[4]; P.qy = q[5]; P.qz = q[6];
    P.v = mk(q[7], q[8], q[9]);
    P.w = rot(q[3], q[4], q[5], q[6], mk(q[10], q[11], q[12]));
    return P;
}

__device__ __forceinline__ long upper_bound_u64(const unsigned long long* k, long n,
                                                unsigned long long key)
{
    long first = 0, length = n;
    while (length > 0) {
        const long half = length >> 1, mid = first + half;
        if (key < k[mid]) length = half;
        else { first = mid + 1; length -= half + 1; }
    }
    return first;
}

// Entries [b, e) of the triangle grid for cell (x, y, z).
__device__ __forceinline__ void tri_range(const TriSet& T, int x, int y, int z, long& b, long& e)
{
    b = e = 0;
    if (T.entries <= 0) return;
    const long long dx = (long long)x - T.lo[0], dy = (long long)y - T.lo[1], dz = (long long)z - T.lo[2];
    if (dx < 0 || dy < 0 || dz < 0 || dx >= T.dims[0] || dy >= T.dims[1] || dz >= T.dims[2]) return;
    const unsigned long long c = ((unsigned long long)dx * T.dims[1] + dy) * T.dims[2] + dz;
    if (T.start) {
        b = T.start[c];
        e = T.start[c + 1];
    } else {
        b = lower_bound_u64(T.tkey, T.entries, c);
        e = upper_bound_u64(T.tkey, T.entries, c);
    }
}

// Field force on owner row p: gravity, then the O(N^2) universal
// gravitation over the owners in input-row order (contact:1083-1106), compiled
// in only with UNIVERSAL (its loop costs the kernel registers otherwise).
template <bool UNIVERSAL = true>
__device__ __forceinline__ V3 field_force(const LawArgs& A, long p, double mass)
{
    const Soa& S = A.S;
    V3 fF = zero3();
    if (A.gravity) fF = mk(mass * A.g[0], mass * A.g[1], mass * A.g[2]);
    if (UNIVERSAL && A.universal) {
        const V3 pi = mk(S.f(F_X, p), S.f(F_Y, p), S.f(F_Z, p));
        for (long oj = 0; oj < S.n; ++oj) {
            const long j = A.byorig[oj];
            if (j == p) continue;
            const V3 d = sub(pi, mk(S.f(F_X, j), S.f(F_Y, j), S.f(F_Z, j)));
            const double d2 = dot(d, d);
            if (d2 > 1.0e-28)
                fF = add(fF, scl(d, kGravityConstant * mass * S.f(F_MASS, j) / (d2 * sqrt(d2))));
        }
    }
    return fF;
}

// Walls of one contact row r (history row hr) in wall order (wallContact,
// contact:975-1063), after the row's wall-history expiry (expireHistory,
// contact:217-229: slots not seen last step are dropped); adds to wF / wT.
// TAPE: the differentiable run's records are compiled in (k_law's variant
// that finishes the step never runs with a tape).  ONE: a single scene, whose
// walls and wall law are the same for every lane (scalar registers).
template <bool TAPE = true, bool ONE = false>
__device__ __forceinline__ void wall_pass(const LawArgs& A, const Body& I, long r, long hr, int scene, V3& wF,
                                          V3& wT, int& bodyHigh, int& pwOver)
{
    const History& H = A.H;
    const long nr = A.R.n;
    const int s = A.step;
    unsigned pwOcc = H.pwOcc[hr] & H.pwSeen[(long)((s + 1) & 1) * nr + hr];
    unsigned pwCur = 0;
    if (TAPE && A.tape.pwOcc) A.tape.pwOcc[hr] = pwOcc;
    // walls of the row's scene (history key: the scene's wall index + 1)
    const int w0 = ONE ? 0 : A.sp[scene].wall0, nw = ONE ? A.W.n : A.sp[scene].nwall;
    Material mw = A.matPW[ONE ? 0 : scene];
    if (A.first) mw.dt = 0.0;
    for (int w = 0; w < nw; ++w) {
        const int g = w0 + w;
        const V3 outward = mk(A.W.nrm[3 * g], A.W.nrm[3 * g + 1], A.W.nrm[3 * g + 2]);
        const V3 origin = mk(A.W.o[3 * g], A.W.o[3 * g + 1], A.W.o[3 * g + 2]);
        const double distance = dot(sub(I.pos, origin), outward);
        const double depth = I.r - distance;
        if (depth <= 0.0) continue;
        const int slot = acquire(pwOcc, pwCur, H.pwKey + hr * PW_SLOTS, H.pwVal + hr * PW_SLOTS * 9, 9L,
                                 PW_SLOTS, (uint32_t)(w + 1), bodyHigh, pwOver);
        if (slot < 0) continue;
        double* hv = H.pwVal + (hr * PW_SLOTS + slot) * 9;
        if (TAPE && A.tape.wallCount) {
            // one tape entry per wall contact, reserved for the wave's
            // contacting lanes with one atomic (the lowest lane's)
            const unsigned long long act = __ballot(1);
            const int lane = threadIdx.x & 63, lead = __ffsll((long long)act) - 1;
            unsigned first = 0u;
            if (lane == lead) first = atomicAdd(A.tape.wallCount, (unsigned)__popcll(act));
            first = __builtin_amdgcn_readlane(first, lead);
            const unsigned e = first + (unsigned)__popcll(act & ((1ull << lane) - 1ull));
            if (A.tape.wall) {
                A.tape.wall[e] = make_int4((int)r, w, slot, 0);
                for (int k = 0; k < 9; ++k) A.tape.wallPrev[(long)e * 9 + k] = hv[k];
            }
        }
        const V3 cf = pw_law(I, outward, depth, mw, hv, wF, wT);
        if (TAPE && A.tape.plateAcc && w == A.tape.plate && !A.first) A.tape.plateAcc[r] += cf.z;
    }
    H.pwOcc[hr] = pwOcc;
    H.pwSeen[(long)(s & 1) * nr + hr] = pwCur;
}

// observeBodyMicrostateKernel (contact:1226-1356) for one (row, body) in
// contact: the 22 microstate terms from the row's manifold (normal, centre,
// distance, area total) and its manifold slot after this step's update,
// written to micro[f * stride] (the [body][22][t] rows reduced in cell order).
__device__ __forceinline__ void micro_terms(const Body& I, const Pose& P, const Material& m, V3 normal, V3 cc,
                                            double distance, double total, const double* sv, double* ms,
                                            long stride)
{
    const V3 arm = sub(cc, P.c);
    const V3 bv = add(P.v, cross(P.w, arm));
    const V3 rv = sub(bv, I.vel);
    const double cN = damping(m.logN, m.kN, I.mass);
    const double radius = I.r;
    const double brs = m.beta * radius * m.beta * radius;
    const double depth = radius - distance;
    V3 fn = add(scl(normal, -m.kN * depth), scl(normal, cN * dot(rv, normal)));
    fn = add(fn, scl(normal, m.cohesion * 4.0 * brs));
    const double fnn = len(fn);
    const double msc = m.beta * radius * fnn;
    const double tlim = m.mu * fnn;
    const double wlim = m.mu * m.muT * msc;
    const V3 th = hget(sv, 1), wh = hget(sv, 2);
    const V3 axisZ = rot(P.qw, P.qx, P.qy, P.qz, mk(0.0, 0.0, 1.0));
    const V3 radial = sub(arm, scl(axisZ, dot(arm, axisZ)));
    const double tu = tlim > 1.0e-30 ? fmin(1.0, len(th) / tlim) : 0.0;
    const double wu = wlim > 1.0e-30 ? fmin(1.0, len(wh) / wlim) : 0.0;
    const V3 btw = scl(wh, -total), bnl = scl(fn, -total), bth = scl(th, -total);
    const V3 btt = cross(arm, bth);
    const double v[22] = {fnn * total, fnn * radius * total, fnn * len(radial) * total,
                          btw.x, btw.y, btw.z, len(wh) * total, wlim * total, len(th) * total,
                          tlim * total, wu * total, tu * total, total, bnl.x, bnl.y, bnl.z,
                          bth.x, bth.y, bth.z, btt.x, btt.y, btt.z};
    for (int f = 0; f < 22; ++f) ms[(long)f * stride] = v[f];
}

// Per-triangle geometry of one contact row for tri_pass, visited body by
// body in entry order (a cell's entries are in triangle order and a body's
// triangles are contiguous in body order, so each body's triangles form one
// run of the cell's entries).  manifold(body, f) calls f(n, scale, distance,
// centre) for the body's triangles with a positive area scale; contacts(...)
// then calls g(q) with the read-only contact of each of them.
//
// EntryGeo evaluates the geometry in the lane: the run's entries whose
// bounding box meets the sphere (the contact circle lies on it) as bits of a
// mask (the first 64; a longer run's rest is visited entry by entry), then
// the area scale of each, lowest bit first.  pt_law hits exactly the entries
// with a positive scale (its depth test follows from r^2 - d^2 > 1e-30).
struct EntryGeo {
    const TriSet& T;
    V3 pos;
    double r;
    long e0, ks, ke, kx;
    unsigned long long hits;
    __device__ EntryGeo(const TriSet& T_, V3 p, double r_, long b0, long e0_)
        : T(T_), pos(p), r(r_), e0(e0_), ks(b0), ke(b0), kx(b0), hits(0ull) {}
    template <class F>
    __device__ __forceinline__ void manifold(int body, F&& f)
    {
        unsigned long long cand = 0ull;
        ke = ks;
        for (; ke < e0; ++ke) {
            const int tr = T.ttri[ke];
            if (T.body[tr] != body) break;
            if (ke - ks < 64 && ball_meets_box(pos, r, T.aabb + (long)tr * 6)) cand |= 1ull << (ke - ks);
        }
        kx = ke < ks + 64 ? ke : ks + 64;
        hits = 0ull;
        for (unsigned long long m = cand, kt = (unsigned long long)kx;;) {
            long k;
            if (m) { k = ks + __builtin_ctzll(m); m &= m - 1ull; }
            else if ((long)kt < ke) k = (long)kt++;
            else break;
            const int tr = T.ttri[k];
            const double* box = T.aabb + (long)tr * 6;
            if (!ball_meets_box(pos, r, box)) continue;
            const Tri tri = load_tri(T, tr);
            const double scv = tri_area_scale(pos, r, tri, box);
            if (scv <= 0.0) continue;
            if (k - ks < 64) hits |= 1ull << (k - ks);
            const double distance = dot(tri.n, sub(pos, tri.a));
            f(tri.n, scv, distance, sub(pos, scl(tri.n, distance)));
        }
    }
    template <class G>
    __device__ __forceinline__ void contacts(const Body& I, const Pose& P, const Material& ro, const double* sv, G&& g)
    {
        for (unsigned long long m = hits, kt = (unsigned long long)kx;;) {
            long k;
            if (m) { k = ks + __builtin_ctzll(m); m &= m - 1ull; }
            else if ((long)kt < ke) k = (long)kt++;
            else break;
            const PtOut q = pt_law(I, load_tri(T, T.ttri[k]), P, ro, sv);
            if (q.hit) g(q);
        }
    }
    __device__ __forceinline__ void next() { ks = ke; }
};

// HitGeo reads the row's hits as k_tri_area stored them: the entries with a
// positive area scale as a mask over the cell's entries, and (scale, plane
// distance) per hit in entry order; normal and centre are formed as above.
struct HitGeo {
    const TriSet& T;
    V3 pos;
    long b0;
    unsigned long long rest, run;  // hits not yet visited; the current body's run
    const double2* val;
    int rank, runRank;
    __device__ HitGeo(const TriSet& T_, V3 p, long b0_, unsigned long long hits, const double2* v)
        : T(T_), pos(p), b0(b0_), rest(hits), run(0ull), val(v), rank(0), runRank(0) {}
    __device__ __forceinline__ V3 normal(int tr) const
    {
        const double* w = T.world + (long)tr * 15;
        return mk(w[12], w[13], w[14]);
    }
    template <class F>
    __device__ __forceinline__ void manifold(int body, F&& f)
    {
        run = 0ull;
        runRank = rank;
        while (rest) {
            const int tr = T.ttri[b0 + __builtin_ctzll(rest)];
            if (T.body[tr] != body) break;
            const V3 n = normal(tr);
            const double2 v = val[rank];
            f(n, v.x, v.y, sub(pos, scl(n, v.y)));
            run |= rest & (~rest + 1ull);
            rest &= rest - 1ull;
            ++rank;
        }
    }
    template <class G>
    __device__ __forceinline__ void contacts(const Body& I, const Pose& P, const Material& ro, const double* sv, G&& g)
    {
        int k = runRank;
        for (unsigned long long m = run; m; m &= m - 1ull, ++k) {
            const V3 n = normal(T.ttri[b0 + __builtin_ctzll(m)]);
            const double2 v = val[k];
            g(pt_contact(I, n, v.y, sub(I.pos, scl(n, v.y)), v.x, P, ro, sv));
        }
    }
    __device__ __forceinline__ void next() {}
};

// Particle-triangle part of computeContactsKernel (contact:1148-1208) for
// contact row r (history row hr, owner input row o, component input row u),
// over the row's cell entries with geometry from `geo`: per body, the
// area-weighted manifold advances one history slot, then every triangle
// contributes its read-only contact.  Adds the row's force and torque to
// bF / bT and the owner's per-body outputs at o; writes the row's body load,
// Fz and microstate terms at u (reduced over the component input rows later)
// and its regime state at r; depthBits keeps the largest contact depth (as
// bits) for one atomic per wave.
template <class Geo>
__device__ __forceinline__ void tri_pass_geo(const LawArgs& A, const Body& I, long r, long u, long hr, long o,
                                             bool listed, Geo& geo, V3& bF, V3& bT, int& bodyHigh, int& ptOver,
                                             int& bcount, unsigned long long& depthBits)
{
    const Soa& R = A.R;
    const long nr = R.n, no = A.S.n;
    const int s = A.step;
    const History& H = A.H;
    unsigned long long ptOcc = H.ptOcc[hr] & H.ptSeen[(long)((s + 1) & 1) * nr + hr];
    unsigned long long ptCur = 0;
    const TriSet& T = A.T;
    if (listed) {
        for (int body = 0; body < A.nb; ++body) {
            double total = 0.0, wD = 0.0;
            V3 wN = zero3(), wC = zero3();
            geo.manifold(body, [&](V3 n, double scv, double distance, V3 cc) {
                total += scv;
                wN = add(wN, scl(n, -scv));
                wC = add(wC, scl(cc, scv));
                wD += distance * scv;
            });
            double bl[6] = {0, 0, 0, 0, 0, 0};
            double fz = 0.0;
            double* br = A.O.bodyRows + ((long)body * nr + u) * BODY_ROW;
            bool observed = false;
            if (total > 0.0) {
                const V3 mN = unit(wN);
                const V3 mC = scl(wC, 1.0 / total);
                const double mD = wD / total;
                const int slot = acquire(ptOcc, ptCur, H.ptKey + hr * PT_SLOTS, H.ptVal + hr * PT_SLOTS * 9, 9L,
                                         PT_SLOTS, (uint32_t)(body + 1), bodyHigh, ptOver);
                if (slot >= 0) {
                    double* sv = H.ptVal + (hr * PT_SLOTS + slot) * 9;
                    const Pose P = load_pose(T, body);
                    bool slide = false, twisty = false;
                    manifold_law(I, P, A.pt, mN, mC, mD, sv, &slide, &twisty);
                    A.O.regime[(long)body * nr + r] = 1 | (slide ? 2 : 0) | (twisty ? 4 : 0);
                    micro_terms(I, P, A.pt, mN, mC, mD, total, sv, br + 7, 1);
                    observed = true;
                    Material ro = A.pt;
                    ro.dt = 0.0;
                    geo.contacts(I, P, ro, sv, [&](const PtOut& q) {
                        bF = add(bF, q.pf);
                        bT = add(bT, q.ptq);
                        const V3 btq = add(add(q.bnt, q.bft), add(q.btw, q.brl));
                        bl[0] += q.bnf.x + q.bfr.x;
                        bl[1] += q.bnf.y + q.bfr.y;
                        bl[2] += q.bnf.z + q.bfr.z;
                        bl[3] += btq.x;
                        bl[4] += btq.y;
                        bl[5] += btq.z;
                        fz += q.bnf.z + q.bfr.z;
                        ++bcount;
                        const unsigned long long db = (unsigned long long)__double_as_longlong(q.depth);
                        depthBits = db > depthBits ? db : depthBits;
                        if (A.O.bodyParticle) {
                            double* d = A.O.bodyParticle + ((long)body * no + o) * 6;
                            d[0] += q.pf.x; d[1] += q.pf.y; d[2] += q.pf.z;
                            d[3] += q.ptq.x; d[4] += q.ptq.y; d[5] += q.ptq.z;
                        }
                        if (A.O.bodyDetail) {
                            double* d = A.O.bodyDetail + ((long)body * no + o) * 19;
                            const V3 gr[6] = {q.bnf, q.bfr, q.bnt, q.bft, q.btw, q.brl};
                            for (int g = 0; g < 6; ++g) {
                                d[3 * g] += gr[g].x;
                                d[3 * g + 1] += gr[g].y;
                                d[3 * g + 2] += gr[g].z;
                            }
                            if (d[18] < q.depth) d[18] = q.depth;
                        }
                    });
                }
            }
            // a body with no manifold contributes zeros: its row is marked
            // untouched instead of written
            if (total > 0.0) {
                for (int k = 0; k < 6; ++k) br[k] = bl[k];
                br[6] = fz;
                if (!observed)
                    for (int f = 0; f < 22; ++f) br[7 + f] = 0.0;
            }
            A.O.bodyTouched[(long)body * nr + u] = total > 0.0 ? 1 : 0;
            geo.next();
        }
    } else {
        for (int body = 0; body < A.nb; ++body) A.O.bodyTouched[(long)body * nr + u] = 0;
    }
    H.ptOcc[hr] = ptOcc;
    H.ptSeen[(long)(s & 1) * nr + hr] = ptCur;
}

// tri_pass_geo with the geometry evaluated in the lane (rows of any size).
__device__ __forceinline__ void tri_pass(const LawArgs& A, const Body& I, long r, long u, long hr, long o,
                                         V3& bF, V3& bT, int& bodyHigh, int& ptOver, int& bcount,
                                         unsigned long long& depthBits)
{
    const TriSet& T = A.T;
    long b0 = 0, e0 = 0;
    const bool listed = T.nt > 0 && T.entries > 0 && A.R.k(I_VALID, r);
    if (listed) tri_range(T, A.R.k(I_CX, r), A.R.k(I_CY, r), A.R.k(I_CZ, r), b0, e0);
    EntryGeo geo(T, I.pos, I.r, b0, e0);
    tri_pass_geo(A, I, r, u, hr, o, listed, geo, bF, bT, bodyHigh, ptOver, bcount, depthBits);
}

// A row's orientation, angular velocity and mass at the end of its force
// evaluation (what its acceleration was computed from).
struct RowSpin {
    double qw, qx, qy, qz, wx, wy, wz, mass;
};

__device__ __forceinline__ RowSpin load_spin(const Soa& S, long p)
{
    return RowSpin{S.f(F_QW, p), S.f(F_QX, p), S.f(F_QY, p), S.f(F_QZ, p),
                   S.f(F_WX, p), S.f(F_WY, p), S.f(F_WZ, p), S.f(F_MASS, p)};
}

// advanceHalfVelocityKernel of the previous step followed by
// advancePositionKernel of the next one (particle:213-265) for row p, in one
// pass over its state, given the row's accelerations (a, aw) of the force
// evaluation in between: the reference's loop calls them back to back with
// only host work (the body solve) in between.  KEEP_HALF_QUAT: also store the
// half-step quaternion rate (F_HQ*), which only k_advance reads; the law
// kernel's fused pass leaves it out, since the run's next step rewrites it
// in its own finish before any advance could read it.
template <bool KEEP_HALF_QUAT>
__device__ __forceinline__ void finish_advance_spin(const Soa& S, long p, double dtf, double dt, const Recs& R,
                                                    const double (&a)[6], const RowSpin& r)
{
    const double ax = a[0], ay = a[1], az = a[2], awx = a[3], awy = a[4], awz = a[5];
    double vhx = S.f(F_VHX, p) + ax * dtf, vhy = S.f(F_VHY, p) + ay * dtf, vhz = S.f(F_VHZ, p) + az * dtf;
    double whx = S.f(F_WHX, p) + awx * dtf, why = S.f(F_WHY, p) + awy * dtf, whz = S.f(F_WHZ, p) + awz * dtf;
    double qw = r.qw, qx = r.qx, qy = r.qy, qz = r.qz;
    const double h = 0.5 * dtf;
    double dw, dx, dy, dz;
    quat_dot(qw, qx, qy, qz, r.wx, r.wy, r.wz, dw, dx, dy, dz);
    double hw = qw + dw * h, hx = qx + dx * h, hy = qy + dy * h, hz = qz + dz * h;
    quat_unit(hw, hx, hy, hz);
    double gw, gx, gy, gz;
    quat_dot(hw, hx, hy, hz, whx, why, whz, gw, gx, gy, gz);
    S.f(F_VHX, p) = vhx; S.f(F_VHY, p) = vhy; S.f(F_VHZ, p) = vhz;
    S.f(F_WHX, p) = whx; S.f(F_WHY, p) = why; S.f(F_WHZ, p) = whz;
    if (KEEP_HALF_QUAT) { S.f(F_HQW, p) = gw; S.f(F_HQX, p) = gx; S.f(F_HQY, p) = gy; S.f(F_HQZ, p) = gz; }
    V3 pos = mk(S.f(F_X, p), S.f(F_Y, p), S.f(F_Z, p)), vel, w;
    if (S.k(I_STATUS, p) != -1) {
        pos = mk(pos.x + vhx * dt, pos.y + vhy * dt, pos.z + vhz * dt);
        S.f(F_X, p) = pos.x; S.f(F_Y, p) = pos.y; S.f(F_Z, p) = pos.z;
        qw = qw + gw * dt; qx = qx + gx * dt; qy = qy + gy * dt; qz = qz + gz * dt;
        quat_unit(qw, qx, qy, qz);
        S.f(F_QW, p) = qw; S.f(F_QX, p) = qx; S.f(F_QY, p) = qy; S.f(F_QZ, p) = qz;
        vel = mk(vhx + ax * (0.5 * dt), vhy + ay * (0.5 * dt), vhz + az * (0.5 * dt));
        w = mk(whx + awx * (0.5 * dt), why + awy * (0.5 * dt), whz + awz * (0.5 * dt));
        S.f(F_VX, p) = vel.x; S.f(F_VY, p) = vel.y; S.f(F_VZ, p) = vel.z;
        S.f(F_WX, p) = w.x; S.f(F_WY, p) = w.y; S.f(F_WZ, p) = w.z;
    } else {
        if (!R.main) return;
        vel = mk(S.f(F_VX, p), S.f(F_VY, p), S.f(F_VZ, p));
        w = mk(r.wx, r.wy, r.wz);
    }
    if (R.main) put_rec(R, p, pos, vel, mk(vhx, vhy, vhz), qw, qx, qy, qz, w, mk(whx, why, whz), r.mass);
}

__device__ __forceinline__ void finish_advance_row(const Soa& S, long p, double dtf, double dt, const Recs& R,
                                                   double ax, double ay, double az, double awx, double awy,
                                                   double awz)
{
    const double a[6] = {ax, ay, az, awx, awy, awz};
    finish_advance_spin<true>(S, p, dtf, dt, R, a, load_spin(S, p));
}

// Force outputs of owner row p (contact:1211-1220) and its acceleration
// (accelerationKernel, particle:71-101), also returned in acc, with the
// orientation and angular velocity it was computed from in spin.  STORE:
// write the force outputs and the acceleration.  The law kernel's fused pass
// (a step that another step of the same run follows) keeps them in
// registers: the next step's evaluation rewrites every row's force outputs
// and acceleration before anything reads them.
template <bool STORE = true>
__device__ __forceinline__ void owner_out(const LawArgs& A, long p, double m, V3 cF, V3 cT, V3 fF, V3 wF, V3 wT,
                                          V3 bF, V3 bT, double (&acc)[6], RowSpin& spin)
{
    const Soa& S = A.S;
    const long n = S.n;
    if (STORE) {
        double* oc = A.O.contact;
        oc[0 * n + p] = cF.x; oc[1 * n + p] = cF.y; oc[2 * n + p] = cF.z;
        oc[3 * n + p] = cT.x; oc[4 * n + p] = cT.y; oc[5 * n + p] = cT.z;
        double* ow = A.O.wall;
        ow[0 * n + p] = wF.x; ow[1 * n + p] = wF.y; ow[2 * n + p] = wF.z;
        ow[3 * n + p] = wT.x; ow[4 * n + p] = wT.y; ow[5 * n + p] = wT.z;
        double* of = A.O.field;
        of[0 * n + p] = fF.x; of[1 * n + p] = fF.y; of[2 * n + p] = fF.z;
    }
    const V3 tF = add(add(cF, fF), add(wF, bF));
    const V3 tT = add(cT, add(wT, bT));
    acc[0] = tF.x / m;
    acc[1] = tF.y / m;
    acc[2] = tF.z / m;
    if (STORE) {
        S.f(F_AX, p) = acc[0];
        S.f(F_AY, p) = acc[1];
        S.f(F_AZ, p) = acc[2];
    }
    const double qw = S.f(F_QW, p), qx = S.f(F_QX, p), qy = S.f(F_QY, p), qz = S.f(F_QZ, p);
    const double tx = (1.0 - 2.0 * (qy * qy + qz * qz)) * tT.x + 2.0 * (qx * qy + qw * qz) * tT.y +
                      2.0 * (qx * qz - qw * qy) * tT.z;
    const double ty = 2.0 * (qx * qy - qw * qz) * tT.x + (1.0 - 2.0 * (qx * qx + qz * qz)) * tT.y +
                      2.0 * (qy * qz + qw * qx) * tT.z;
    const double tz = 2.0 * (qx * qz + qw * qy) * tT.x + 2.0 * (qy * qz - qw * qx) * tT.y +
                      (1.0 - 2.0 * (qx * qx + qy * qy)) * tT.z;
    const double wx = S.f(F_WX, p), wy = S.f(F_WY, p), wz = S.f(F_WZ, p);
    const double ix = S.f(F_IX, p), iy = S.f(F_IY, p), iz = S.f(F_IZ, p);
    acc[3] = (tx - (iz - iy) * wz * wy) / ix;
    acc[4] = (ty - (ix - iz) * wx * wz) / iy;
    acc[5] = (tz - (iy - ix) * wy * wx) / iz;
    if (STORE) {
        S.f(F_AWX, p) = acc[3];
        S.f(F_AWY, p) = acc[4];
        S.f(F_AWZ, p) = acc[5];
    }
    spin = RowSpin{qw, qx, qy, qz, wx, wy, wz, m};
}
__device__ __forceinline__ void owner_out(const LawArgs& A, long p, double m, V3 cF, V3 cT, V3 fF, V3 wF, V3 wT,
                                          V3 bF, V3 bT, double (&acc)[6])
{
    RowSpin spin;
    owner_out<true>(A, p, m, cF, cT, fF, wF, wT, bF, bT, acc, spin);
}

// computeContactsKernel (contact:1065-1221) for one single-sphere owner,
// plus the acceleration (particle:71-101).  UNIVERSAL: runs with universal
// gravitation (without it the kernel fits 3 waves per SIMD instead of 2).
// FINISH: a step of a multi-step run that another step follows: the row's
// finish of this step and advance of the next (finish_advance_row) follow its
// acceleration here, from registers, instead of in k_finish_advance (no row
// reads another row's state or record in this kernel), and the first block
// resets the next evaluation's step counters (this evaluation's are not read:
// it collects no diagnostics).  Such a step stores neither its force outputs
// nor its accelerations nor the half-step quaternion rate: the run's next
// step rewrites all three before anything reads them.  Kept at 3 waves per
// SIMD (at most 168 VGPRs) without universal gravitation.  ONE: a single
// scene (walls and wall law wave-uniform).  COUNTED: the differentiable
// forward run's fused pass, which keeps this evaluation's counters for the
// host (the next evaluation resets them) and records the wall tape (contact
// count and plate load).
template <bool UNIVERSAL, bool FINISH, bool ONE = false, bool COUNTED = false>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(UNIVERSAL ? 2 : 3))) k_law(LawArgs A)
{
    if (FINISH && !COUNTED && blockIdx.x == 0) reset_step_scalars(A.sc, A.sceneAcc, A.nscenes);
    const long t = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    const Soa& S = A.S;
    const long n = S.n;
    int bodyHigh = 0, pwOver = 0;
    if (t < n) {
        const long p = A.identRows ? t : (long)A.sorted[t];
        const double mass = S.f(F_MASS, p);
        V3 cF = zero3(), cT = zero3(), wF = zero3(), wT = zero3(), bF = zero3(), bT = zero3();
        const V3 fF = field_force<UNIVERSAL>(A, p, mass);
        // walls: the depth test of wall_pass on the packed centre and radius;
        // the full contact record is read only by rows that touch a wall (a
        // single scene's walls are uniform across the grid, so the test waits
        // on no per-row load but the centre)
        const int scene = ONE ? 0 : S.k(I_SCENE, p);
        const long hr = A.identRows ? p : (long)S.k(I_HIST, p);  // rows hold their own histories on list steps
        const double4 cr = A.posr[t];
        bool touch = false;
        auto test_wall = [&](int g) {
            const V3 outward = mk(A.W.nrm[3 * g], A.W.nrm[3 * g + 1], A.W.nrm[3 * g + 2]);
            const V3 origin = mk(A.W.o[3 * g], A.W.o[3 * g + 1], A.W.o[3 * g + 2]);
            touch |= cr.w - dot(sub(mk(cr.x, cr.y, cr.z), origin), outward) > 0.0;
        };
        if (ONE || A.oneScene) {
            // wall index uniform across the wave: the planes come through
            // the scalar cache
            for (int g = 0; g < A.W.n; ++g) test_wall(g);
        } else {
            const int w0 = A.sp[scene].wall0, nw = A.sp[scene].nwall;
            for (int w = 0; w < nw; ++w) test_wall(w0 + w);
        }
        if (touch) {
            Body I = rec_body(A.rec, p);
            if (A.offs) I.off = mk(S.f(F_OFX, p), S.f(F_OFY, p), S.f(F_OFZ, p));
            wall_pass<!FINISH || COUNTED, ONE>(A, I, p, hr, scene, wF, wT, bodyHigh, pwOver);
        } else {
            // expiry only: no wall slot is seen this step
            const History& H = A.H;
            const unsigned occ = H.pwOcc[hr] & H.pwSeen[(long)((A.step + 1) & 1) * n + hr];
            if ((!FINISH || COUNTED) && A.tape.pwOcc) A.tape.pwOcc[hr] = occ;
            H.pwOcc[hr] = occ;
            H.pwSeen[(long)(A.step & 1) * n + hr] = 0u;
        }
        // particle pairs, summed in candidate order from the per-contact results
        const int cnt = A.ncount[t];
        const long base = A.off[t];
        if (A.rowF && row_in_chunk(base, cnt)) {
            const double* rf = A.rowF + t * 6;
            cF = mk(rf[0], rf[1], rf[2]);
            cT = mk(rf[3], rf[4], rf[5]);
        } else {
            for (int c = 0; c < cnt; ++c) {
                const double* pf = A.pairF + (base + c) * 6;
                cF = add(cF, mk(pf[0], pf[1], pf[2]));
                cT = add(cT, mk(pf[3], pf[4], pf[5]));
            }
        }
        // triangle contacts were evaluated by k_law_pt (separate accumulator)
        if (A.nb > 0) {
            const double* ob_ = A.O.body;
            bF = mk(ob_[0 * n + p], ob_[1 * n + p], ob_[2 * n + p]);
            bT = mk(ob_[3 * n + p], ob_[4 * n + p], ob_[5 * n + p]);
        }
        double acc[6];
        RowSpin spin;
        owner_out<!FINISH>(A, p, mass, cF, cT, fF, wF, wT, bF, bT, acc, spin);
        if (FINISH) finish_advance_spin<false>(S, p, A.finishDt, A.nextDt, A.rec, acc, spin);
    }
    if (A.counters) atomic_max_block(&A.sc->body_high, bodyHigh);
    atomic_add_wave(&A.sc->pw_overflow, pwOver);
}

// Single-sphere owners whose cell holds no triangle entry: tri_pass's
// no-contact path (triangle-slot expiry, zero body accumulator, body rows
// marked untouched).  The others are listed in `work` (sorted positions) for
// k_tri_area, k_law_pt and k_regime.
__global__ void __launch_bounds__(BLOCK) k_tri_sift(LawArgs A, int* work)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const Soa& S = A.S;
    const long n = S.n;
    bool near = false;
    if (t < n) {
        const long p = A.sorted[t];
        long b0 = 0, e0 = 0;
        if (A.T.nt > 0 && A.T.entries > 0 && S.k(I_VALID, p))
            tri_range(A.T, S.k(I_CX, p), S.k(I_CY, p), S.k(I_CZ, p), b0, e0);
        near = e0 > b0;
        if (!near) {
            const long u = S.k(I_ORIG, p);
            for (int body = 0; body < A.nb; ++body) A.O.bodyTouched[(long)body * n + u] = 0;
            const History& H = A.H;
            const long hr = S.k(I_HIST, p);
            H.ptOcc[hr] = H.ptOcc[hr] & H.ptSeen[(long)((A.step + 1) & 1) * n + hr];
            H.ptSeen[(long)(A.step & 1) * n + hr] = 0ull;
            for (int k = 0; k < 6; ++k) A.O.body[(long)k * n + p] = 0.0;
        }
    }
    block_append(near, (int)t, &A.sc->tri_work, work);
}

// Triangle geometry of the rows listed by k_tri_sift, 16 lanes per row (four
// rows per wave; waves stride over the list), one lane per entry of the
// row's cell in chunks of 16: the bounding-box test and area scale of
// EntryGeo for all of the row's triangles at once.  Stores the row's hit
// mask (entries with a positive area scale) and (scale, plane distance) per
// hit in entry order; a row with more than 64 entries or TRI_HITS hits gets
// hitN = -1 and is evaluated in the lane by k_law_pt.
constexpr int TRI_HITS = 16;
constexpr int TRI_AREA_BLOCKS = 2048;
constexpr int TRI_GROUP = 16;
__global__ void __launch_bounds__(BLOCK) k_tri_area(LawArgs A, const int* work, unsigned long long* hitMask,
                                                    int* hitN, double2* hitVal)
{
    const int lane = threadIdx.x & 63, g = lane / TRI_GROUP, l = lane % TRI_GROUP;
    const long count = A.sc->tri_work;
    const long rowsPerPass = (long)gridDim.x * (BLOCK / 64) * (64 / TRI_GROUP);
    for (long i0 = ((long)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) * (64 / TRI_GROUP); i0 < count;
         i0 += rowsPerPass) {
        const long i = i0 + g;
        const bool active = i < count;
        long b0 = 0, e0 = 0;
        double4 cr = make_double4(0.0, 0.0, 0.0, 0.0);
        if (active) {
            const long t = work[i];
            const long p = A.sorted[t];
            tri_range(A.T, A.S.k(I_CX, p), A.S.k(I_CY, p), A.S.k(I_CZ, p), b0, e0);
            cr = A.posr[t];
        }
        const long m = e0 - b0;
        const int chunks = active && m <= 64 ? (int)((m + TRI_GROUP - 1) / TRI_GROUP) : 0;
        const V3 pos = mk(cr.x, cr.y, cr.z);
        unsigned long long rowHits = 0ull;
        int nh = 0;
        const int passes = wave_max(chunks);  // uniform over the wave: the ballots stay converged
        for (int c = 0; c < passes; ++c) {
            const long j = (long)c * TRI_GROUP + l;
            double scv = 0.0, distance = 0.0;
            if (c < chunks && j < m) {
                const int tr = A.T.ttri[b0 + j];
                const double* box = A.T.aabb + (long)tr * 6;
                if (ball_meets_box(pos, cr.w, box)) {
                    const Tri tri = load_tri(A.T, tr);
                    scv = tri_area_scale(pos, cr.w, tri, box);
                    if (scv > 0.0) distance = dot(tri.n, sub(pos, tri.a));
                }
            }
            const unsigned bits = (unsigned)((__ballot(scv > 0.0) >> (g * TRI_GROUP)) & 0xFFFFull);
            if (scv > 0.0) {
                const int rank = nh + __popc(bits & ((1u << l) - 1u));
                if (rank < TRI_HITS) hitVal[i * TRI_HITS + rank] = make_double2(scv, distance);
            }
            rowHits |= (unsigned long long)bits << (c * TRI_GROUP);
            nh += __popc(bits);
        }
        if (active && l == 0) {
            hitMask[i] = rowHits;
            hitN[i] = m <= 64 && nh <= TRI_HITS ? nh : -1;
        }
    }
}

// Triangle contacts (tri_pass_geo) of the single-sphere owners listed by
// k_tri_sift, one per lane in 64-lane blocks (a few hundred waves spread
// over every CU), from the hits k_tri_area stored (or in the lane for the
// rows it left).  Writes the body accumulator (bF, bT) read by k_law.
constexpr int TRI_BLOCK = 64;
__global__ void __launch_bounds__(TRI_BLOCK) k_law_pt(LawArgs A, const int* work,
                                                      const unsigned long long* hitMask, const int* hitN,
                                                      const double2* hitVal)
{
    const long i = (long)blockIdx.x * TRI_BLOCK + threadIdx.x;
    const Soa& S = A.S;
    const long n = S.n;
    int bodyHigh = 0, ptOver = 0, bcount = 0;
    unsigned long long depthBits = 0ull;
    if (i < A.sc->tri_work) {
        const long t = work[i];
        const long p = A.sorted[t];
        Body I = rec_body(A.rec, p);
        if (A.offs) I.off = mk(S.f(F_OFX, p), S.f(F_OFY, p), S.f(F_OFZ, p));
        V3 bF = zero3(), bT = zero3();
        const long hr = S.k(I_HIST, p), o = S.k(I_ORIG, p);
        if (hitN[i] >= 0) {
            long b0, e0;
            tri_range(A.T, S.k(I_CX, p), S.k(I_CY, p), S.k(I_CZ, p), b0, e0);
            HitGeo geo(A.T, I.pos, b0, hitMask[i], hitVal + i * TRI_HITS);
            tri_pass_geo(A, I, p, o, hr, o, true, geo, bF, bT, bodyHigh, ptOver, bcount, depthBits);
        } else {
            tri_pass(A, I, p, o, hr, o, bF, bT, bodyHigh, ptOver, bcount, depthBits);
        }
        double* ob = A.O.body;
        ob[0 * n + p] = bF.x; ob[1 * n + p] = bF.y; ob[2 * n + p] = bF.z;
        ob[3 * n + p] = bT.x; ob[4 * n + p] = bT.y; ob[5 * n + p] = bT.z;
    }
    atomic_add_wave(&A.sc->body_contacts, bcount);
    atomic_max_wave(&A.sc->body_high, bodyHigh);
    atomic_add_wave(&A.sc->pt_overflow, ptOver);
    atomic_max_wave_u64(&A.sc->body_max_depth_bits, depthBits);
}

// computeContactsKernel (contact:1065-1221) for owners made of several
// sphere components (clumps): one thread per owner visits its components in
// (owner, id) order, each with its walls, its particle contacts (the
// per-contact results of k_pairs, in candidate order) and its triangles;
// each kind of contact is summed in one accumulator across the components,
// as the reference's per-owner loop does.  Then the acceleration
// (particle:71-101).  Owners keep their input rows (no reordering).
__global__ void __launch_bounds__(BLOCK) k_law_clump(LawArgs A)
{
    const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
    const Soa& S = A.S;
    const Soa& R = A.R;
    int bodyHigh = 0, pwOver = 0, ptOver = 0, bcount = 0;
    unsigned long long depthBits = 0ull;
    if (i < S.n) {
        const double mass = S.f(F_MASS, i);
        V3 cF = zero3(), cT = zero3(), wF = zero3(), wT = zero3(), bF = zero3(), bT = zero3();
        const V3 fF = field_force(A, i, mass);
        for (long c = A.ownerStart[i]; c < A.ownerStart[i + 1]; ++c) {
            const long t = A.rank[c];
            Body I = rec_body(A.rec, c);
            I.off = mk(R.f(F_OFX, c), R.f(F_OFY, c), R.f(F_OFZ, c));
            const long hr = R.k(I_HIST, c);
            wall_pass(A, I, c, hr, R.k(I_SCENE, c), wF, wT, bodyHigh, pwOver);
            const int cnt = A.ncount[t];
            const long base = A.off[t];
            for (int k = 0; k < cnt; ++k) {
                const double* pf = A.pairF + (base + k) * 6;
                cF = add(cF, mk(pf[0], pf[1], pf[2]));
                cT = add(cT, mk(pf[3], pf[4], pf[5]));
            }
            if (A.nb > 0) tri_pass(A, I, c, c, hr, i, bF, bT, bodyHigh, ptOver, bcount, depthBits);
        }
        if (A.nb > 0) {
            double* ob = A.O.body;
            const long n = S.n;
            ob[0 * n + i] = bF.x; ob[1 * n + i] = bF.y; ob[2 * n + i] = bF.z;
            ob[3 * n + i] = bT.x; ob[4 * n + i] = bT.y; ob[5 * n + i] = bT.z;
        }
        double acc[6];
        owner_out(A, i, mass, cF, cT, fF, wF, wT, bF, bT, acc);
    }
    atomic_add_wave(&A.sc->body_contacts, bcount);
    if (A.counters) atomic_max_block(&A.sc->body_high, bodyHigh);
    atomic_add_wave(&A.sc->pw_overflow, pwOver);
    atomic_add_wave(&A.sc->pt_overflow, ptOver);
    atomic_max_wave_u64(&A.sc->body_max_depth_bits, depthBits);
}

// regime counters (accumulateBodyRegimeKernel, contact:530-567) of body
// blockIdx.y: integer sums, so counting per thread, then per block, with one
// atomic per block and counter gives the same result in any order.  With
// `work`, over the rows listed by k_tri_sift (no other row is in contact);
// else over every row.
constexpr int REGIME_BLOCKS = 64;
__global__ void __launch_bounds__(BLOCK) k_regime(const int* regime, const int* sorted, long n,
                                                  const int* work, const Scalars* sc,
                                                  unsigned long long* instant,
                                                  unsigned long long* interval, int accumulate)
{
    const int body = blockIdx.y;
    const long lim = work ? (long)sc->tri_work : n;
    // total, primary regime (stick / slide / twist), slide, twist, both
    unsigned c[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < lim; i += (long)gridDim.x * BLOCK) {
        const int st = regime[(long)body * n + (work ? (long)sorted[work[i]] : i)];
        if ((st & 1) == 0) continue;
        const bool sl = (st & 2) != 0, tw = (st & 4) != 0;
        ++c[0];
        ++c[tw ? 3 : (sl ? 2 : 1)];
        c[4] += sl;
        c[5] += tw;
        c[6] += sl && tw;
    }
    __shared__ unsigned part[BLOCK / 64][7];
    for (int k = 0; k < 7; ++k) {
        const unsigned v = wave_sum(c[k]);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        unsigned long long v = 0;
        for (int w = 0; w < BLOCK / 64; ++w) v += part[w][threadIdx.x];
        if (v) {
            atomicAdd(&instant[body * 7 + threadIdx.x], v);
            if (accumulate) atomicAdd(&interval[body * 7 + threadIdx.x], v);
        }
    }
}

constexpr int RED_TILE = 4096;

// Deterministic sums of the body rows over the component input rows, in a
// fixed tree independent of scheduling: tiles of RED_TILE rows; in a tile,
// each thread sums rows tid, tid + BLOCK, ... in order, the wave folds its
// partials with xor shuffles (32, 16, ..., 1), and the waves' results are
// added in order.  One block per (tile, body) over all BODY_ROW quantities
// of the body at once: a touched
// element's values are one contiguous read; an untouched one adds nothing
// (its values are zero, and a sum that starts at +0.0 never becomes -0.0, so
// adding +0.0 would change no bit).  partial[row][tile] with rows in result
// order: 6 nb loads, nb Fz, 22 nb microstate terms.
__global__ void __launch_bounds__(BLOCK) k_reduce_body(const double* rows, const unsigned char* touched, long n,
                                                       int nb, double* partial, int nblocks)
{
    const int body = blockIdx.y;
    const long base = (long)blockIdx.x * RED_TILE;
    double acc[BODY_ROW];
    for (int j = 0; j < BODY_ROW; ++j) acc[j] = 0.0;
    // the thread's touched elements as bits (independent flag loads), then
    // their values, lowest bit first: the order of the additions is kept
    unsigned mask = 0;
#pragma unroll
    for (int k = 0; k < RED_TILE / BLOCK; ++k) {
        const long e = base + (long)k * BLOCK + threadIdx.x;
        if (e < n && touched[(long)body * n + e]) mask |= 1u << k;
    }
    for (; mask; mask &= mask - 1u) {
        const long e = base + (long)__builtin_ctz(mask) * BLOCK + threadIdx.x;
        const double* v = rows + ((long)body * n + e) * BODY_ROW;
        for (int j = 0; j < BODY_ROW; ++j) acc[j] += v[j];
    }
    __shared__ double ws[BODY_ROW][BLOCK / 64];
    for (int j = 0; j < BODY_ROW; ++j) {
        const double w = wave_sum(acc[j]);
        if ((threadIdx.x & 63) == 0) ws[j][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < BODY_ROW) {
        const int j = threadIdx.x;
        double v = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) v += ws[j][w];
        const int row = j < 6 ? body * 6 + j : (j == 6 ? 6 * nb + body : 7 * nb + body * 22 + (j - 7));
        partial[(long)row * nblocks + blockIdx.x] = v;
    }
}

// out[row] = sum of the row's partials in block order; rows from accFrom on
// are also added to acc[row - accFrom] when acc is given.
// One wave per row: the lanes load 64 partials at a time, lane 0 adds them in
// block order from the lanes' registers.
__global__ void __launch_bounds__(64) k_reduce_final(const double* partial, int nrows, int nblocks, double* out,
                                                     double* acc, int accFrom)
{
    const int row = blockIdx.x;
    if (row >= nrows) return;  // whole block
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int b0 = 0; b0 < nblocks; b0 += 64) {
        const double v = b0 + lane < nblocks ? partial[(long)row * nblocks + b0 + lane] : 0.0;
        const int k = nblocks - b0 < 64 ? nblocks - b0 : 64;
        for (int j = 0; j < k; ++j) s += __shfl(v, j, 64);
    }
    if (lane == 0) {
        out[row] = s;
        if (acc && row >= accFrom) acc[row - accFrom] += s;
    }
}

// ------------------------------------------------------------ leapfrog
__device__ __forceinline__ void half_quat(const Soa& S, long p, double h)
{
    double qw = S.f(F_QW, p), qx = S.f(F_QX, p), qy = S.f(F_QY, p), qz = S.f(F_QZ, p);
    double dw, dx, dy, dz;
    quat_dot(qw, qx, qy, qz, S.f(F_WX, p), S.f(F_WY, p), S.f(F_WZ, p), dw, dx, dy, dz);
    qw += dw * h; qx += dx * h; qy += dy * h; qz += dz * h;
    quat_unit(qw, qx, qy, qz);
    double a, b, c, d;
    quat_dot(qw, qx, qy, qz, S.f(F_WHX, p), S.f(F_WHY, p), S.f(F_WHZ, p), a, b, c, d);
    S.f(F_HQW, p) = a; S.f(F_HQX, p) = b; S.f(F_HQY, p) = c; S.f(F_HQZ, p) = d;
}

// initializeLeapfrogKernel (particle:185-211)
__global__ void __launch_bounds__(BLOCK) k_init_half(Soa S, double h)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    S.f(F_VHX, p) = S.f(F_VX, p) + S.f(F_AX, p) * h;
    S.f(F_VHY, p) = S.f(F_VY, p) + S.f(F_AY, p) * h;
    S.f(F_VHZ, p) = S.f(F_VZ, p) + S.f(F_AZ, p) * h;
    S.f(F_WHX, p) = S.f(F_WX, p) + S.f(F_AWX, p) * h;
    S.f(F_WHY, p) = S.f(F_WY, p) + S.f(F_AWY, p) * h;
    S.f(F_WHZ, p) = S.f(F_WZ, p) + S.f(F_AWZ, p) * h;
    half_quat(S, p, h);
}

// advancePositionKernel (particle:213-238); with R, also the row's contact
// record for the force evaluation that follows.
__global__ void __launch_bounds__(BLOCK) k_advance(Soa S, double dt, Recs R)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    const V3 vh = mk(S.f(F_VHX, p), S.f(F_VHY, p), S.f(F_VHZ, p));
    const V3 wh = mk(S.f(F_WHX, p), S.f(F_WHY, p), S.f(F_WHZ, p));
    V3 pos = mk(S.f(F_X, p), S.f(F_Y, p), S.f(F_Z, p));
    double qw = S.f(F_QW, p), qx = S.f(F_QX, p), qy = S.f(F_QY, p), qz = S.f(F_QZ, p);
    V3 vel, w;
    if (S.k(I_STATUS, p) != -1) {
        pos = mk(pos.x + vh.x * dt, pos.y + vh.y * dt, pos.z + vh.z * dt);
        S.f(F_X, p) = pos.x; S.f(F_Y, p) = pos.y; S.f(F_Z, p) = pos.z;
        qw = qw + S.f(F_HQW, p) * dt; qx = qx + S.f(F_HQX, p) * dt;
        qy = qy + S.f(F_HQY, p) * dt; qz = qz + S.f(F_HQZ, p) * dt;
        quat_unit(qw, qx, qy, qz);
        S.f(F_QW, p) = qw; S.f(F_QX, p) = qx; S.f(F_QY, p) = qy; S.f(F_QZ, p) = qz;
        vel = mk(vh.x + S.f(F_AX, p) * (0.5 * dt), vh.y + S.f(F_AY, p) * (0.5 * dt),
                 vh.z + S.f(F_AZ, p) * (0.5 * dt));
        w = mk(wh.x + S.f(F_AWX, p) * (0.5 * dt), wh.y + S.f(F_AWY, p) * (0.5 * dt),
               wh.z + S.f(F_AWZ, p) * (0.5 * dt));
        S.f(F_VX, p) = vel.x; S.f(F_VY, p) = vel.y; S.f(F_VZ, p) = vel.z;
        S.f(F_WX, p) = w.x; S.f(F_WY, p) = w.y; S.f(F_WZ, p) = w.z;
    } else {
        if (!R.main) return;
        vel = mk(S.f(F_VX, p), S.f(F_VY, p), S.f(F_VZ, p));
        w = mk(S.f(F_WX, p), S.f(F_WY, p), S.f(F_WZ, p));
    }
    if (R.main) put_rec(R, S, p, pos, vel, vh, qw, qx, qy, qz, w, wh);
}

// advanceHalfVelocityKernel (particle:240-265)
__global__ void __launch_bounds__(BLOCK) k_finish(Soa S, double dt)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    S.f(F_VHX, p) += S.f(F_AX, p) * dt;
    S.f(F_VHY, p) += S.f(F_AY, p) * dt;
    S.f(F_VHZ, p) += S.f(F_AZ, p) * dt;
    S.f(F_WHX, p) += S.f(F_AWX, p) * dt;
    S.f(F_WHY, p) += S.f(F_AWY, p) * dt;
    S.f(F_WHZ, p) += S.f(F_AWZ, p) * dt;
    half_quat(S, p, 0.5 * dt);
}

__global__ void __launch_bounds__(BLOCK) k_finish_advance(Soa S, double dtf, double dt, Recs R, Scalars* sc,
                                                          int* sceneAcc, int nscenes)
{
    if (sc && blockIdx.x == 0) reset_step_scalars(sc, sceneAcc, nscenes);
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    finish_advance_row(S, p, dtf, dt, R, S.f(F_AX, p), S.f(F_AY, p), S.f(F_AZ, p), S.f(F_AWX, p), S.f(F_AWY, p),
                       S.f(F_AWZ, p));
}

// ------------------------------------------------------------ triangles
// transformTrianglesKernel (body:71-112)
__global__ void __launch_bounds__(BLOCK) k_transform(TriSet T)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= T.nt) return;
    const double* b = T.pose + (long)T.body[i] * 13;
    const double* l = T.local + (long)i * 15;
    double* w = T.world + (long)i * 15;
    const double qw = b[3], qx = b[4], qy = b[5], qz = b[6];
    const double r00 = 1.0 - 2.0 * (qy * qy + qz * qz), r01 = 2.0 * (qx * qy - qw * qz),
                 r02 = 2.0 * (qx * qz + qw * qy), r10 = 2.0 * (qx * qy + qw * qz),
                 r11 = 1.0 - 2.0 * (qx * qx + qz * qz), r12 = 2.0 * (qy * qz - qw * qx),
                 r20 = 2.0 * (qx * qz - qw * qy), r21 = 2.0 * (qy * qz + qw * qx),
                 r22 = 1.0 - 2.0 * (qx * qx + qy * qy);
    for (int v = 0; v < 5; ++v) {
        const double X = l[3 * v], Y = l[3 * v + 1], Z = l[3 * v + 2];
        double ox = r00 * X + r01 * Y + r02 * Z, oy = r10 * X + r11 * Y + r12 * Z,
               oz = r20 * X + r21 * Y + r22 * Z;
        if (v < 4) { ox += b[0]; oy += b[1]; oz += b[2]; }
        w[3 * v] = ox; w[3 * v + 1] = oy; w[3 * v + 2] = oz;
    }
    double* bb = T.aabb + (long)i * 6;
    double mnx = w[0], mny = w[1], mnz = w[2], mxx = w[0], mxy = w[1], mxz = w[2];
    for (int v = 1; v < 4; ++v) {
        mnx = fmin(mnx, w[3 * v]); mny = fmin(mny, w[3 * v + 1]); mnz = fmin(mnz, w[3 * v + 2]);
        mxx = fmax(mxx, w[3 * v]); mxy = fmax(mxy, w[3 * v + 1]); mxz = fmax(mxz, w[3 * v + 2]);
    }
    bb[0] = mnx; bb[1] = mny; bb[2] = mnz; bb[3] = mxx; bb[4] = mxy; bb[5] = mxz;
}

// countTriangleCellsKernel (contact:409-451): per triangle, the cells of its
// padded bounding box; the box of all of them and the entry total are
// reduced per wave, then one atomic per wave and quantity.
__global__ void __launch_bounds__(BLOCK) k_tri_count(TriSet T, double inv, double pad, int* counts,
                                                     Scalars* sc)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    unsigned long long cnt = 0;
    int over = 0;
    if (i < T.nt) {
        const double* bb = T.aabb + (long)i * 6;
        const double mnx = floor((bb[0] - pad) * inv), mny = floor((bb[1] - pad) * inv),
                     mnz = floor((bb[2] - pad) * inv), mxx = floor((bb[3] + pad) * inv),
                     mxy = floor((bb[4] + pad) * inv), mxz = floor((bb[5] + pad) * inv);
        const double minCell = -(double)(1 << 20), maxCell = (double)((1 << 20) - 1);
        long long nx = 0, ny = 0, nz = 0;
        bool bad = !isfinite(mnx) || !isfinite(mny) || !isfinite(mnz) || !isfinite(mxx) || !isfinite(mxy) ||
                   !isfinite(mxz) || mnx < minCell || mny < minCell || mnz < minCell || mxx > maxCell ||
                   mxy > maxCell || mxz > maxCell;
        if (!bad) {
            nx = (long long)(int)mxx - (int)mnx + 1;
            ny = (long long)(int)mxy - (int)mny + 1;
            nz = (long long)(int)mxz - (int)mnz + 1;
            bad = nx <= 0 || ny <= 0 || nz <= 0 || nx > LLONG_MAX / ny || nx * ny > LLONG_MAX / nz ||
                  nx * ny * nz > INT_MAX;
        }
        if (bad) {
            over = 1;
            counts[i] = 0;
        } else {
            counts[i] = (int)(nx * ny * nz);
            cnt = (unsigned long long)(nx * ny * nz);
            lo[0] = (int)mnx; lo[1] = (int)mny; lo[2] = (int)mnz;
            hi[0] = (int)mxx; hi[1] = (int)mxy; hi[2] = (int)mxz;
        }
    }
    // reduced per wave, then per block: one atomic per block and quantity
    __shared__ int red[BLOCK / 64][7];
    __shared__ unsigned long long redCnt[BLOCK / 64];
    const int wave = threadIdx.x >> 6;
    const bool lead = (threadIdx.x & 63) == 0;
    for (int a = 0; a < 3; ++a) {
        const int l = wave_min(lo[a]), h = wave_max(hi[a]);
        if (lead) { red[wave][a] = l; red[wave][3 + a] = h; }
    }
    const int o = wave_sum(over);
    cnt = wave_sum(cnt);
    if (lead) { red[wave][6] = o; redCnt[wave] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int l[3], h[3], ov = 0;
        unsigned long long total = 0;
        for (int a = 0; a < 3; ++a) { l[a] = INT_MAX; h[a] = INT_MIN; }
        for (int w = 0; w < BLOCK / 64; ++w) {
            for (int a = 0; a < 3; ++a) { l[a] = min(l[a], red[w][a]); h[a] = max(h[a], red[w][3 + a]); }
            ov += red[w][6];
            total += redCnt[w];
        }
        for (int a = 0; a < 3; ++a) {
            if (l[a] != INT_MAX) atomicMin(&sc->tri_lo[a], l[a]);
            if (h[a] != INT_MIN) atomicMax(&sc->tri_hi[a], h[a]);
        }
        if (ov) atomicAdd(&sc->tri_overflow, ov);
        if (total) atomicAdd(&sc->tri_total, total);
    }
}

// fillTriangleCellsKernel (contact:453-476): x -> y -> z nested order.  The
// key is the cell's index in the box of all entries (x-major like the
// reference's packed key, so the sorted order is the same), which needs only
// ceil(log2 cells) bits of sorting.
__global__ void __launch_bounds__(BLOCK) k_tri_fill(TriSet T, double inv, double pad,
                                                    const int* counts, const long long* offsets,
                                                    unsigned long long* key, int* tri)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= T.nt) return;
    const double* bb = T.aabb + (long)i * 6;
    const int mnx = (int)floor((bb[0] - pad) * inv), mny = (int)floor((bb[1] - pad) * inv),
              mnz = (int)floor((bb[2] - pad) * inv), mxx = (int)floor((bb[3] + pad) * inv),
              mxy = (int)floor((bb[4] + pad) * inv), mxz = (int)floor((bb[5] + pad) * inv);
    const long long out = offsets[i];
    int local = 0;
    const int c = counts[i];
    for (int x = mnx; x <= mxx && local < c; ++x)
        for (int y = mny; y <= mxy && local < c; ++y)
            for (int z = mnz; z <= mxz && local < c; ++z) {
                key[out + local] = (((unsigned long long)(x - T.lo[0]) * (unsigned long long)T.dims[1] +
                                     (unsigned long long)(y - T.lo[1])) * (unsigned long long)T.dims[2]) +
                                   (unsigned long long)(z - T.lo[2]);
                tri[out + local] = i;
                ++local;
            }
}

// Triangle grid as a counting sort over the cell box (boxes with a dense
// table): one wave per triangle visits its cells in the x -> y -> z order of
// fillTriangleCellsKernel; COUNT adds 1 to cur[cell], else the triangle is
// placed at the cell's cursor (cur starts at the cell's first entry).
template <bool COUNT>
__global__ void __launch_bounds__(BLOCK) k_tri_cells(TriSet T, double inv, double pad, const int* counts,
                                                     unsigned* cur, int* tri)
{
    const int i = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (i >= T.nt) return;  // whole waves
    const int c = counts[i];
    const double* bb = T.aabb + (long)i * 6;
    const int mnx = (int)floor((bb[0] - pad) * inv), mny = (int)floor((bb[1] - pad) * inv),
              mnz = (int)floor((bb[2] - pad) * inv), mxy = (int)floor((bb[4] + pad) * inv),
              mxz = (int)floor((bb[5] + pad) * inv);
    const int ny = mxy - mny + 1, nz = mxz - mnz + 1;
    for (int l = threadIdx.x & 63; l < c; l += 64) {
        const int x = mnx + l / (ny * nz), y = mny + (l / nz) % ny, z = mnz + l % nz;
        const unsigned long long key =
            (((unsigned long long)(x - T.lo[0]) * (unsigned long long)T.dims[1] + (unsigned long long)(y - T.lo[1])) *
             (unsigned long long)T.dims[2]) + (unsigned long long)(z - T.lo[2]);
        if (COUNT) atomicAdd(&cur[key], 1u);
        else tri[atomicAdd(&cur[key], 1u)] = i;
    }
}

// Within each cell, entries in ascending triangle order (the reference's
// stable sort of entries generated triangle by triangle).  A cell of up to
// CELL_SORT entries is sorted in registers by its thread (odd-even
// transposition network); larger cells are listed for k_tri_cell_order_big.
constexpr int CELL_SORT = 16;
__global__ void __launch_bounds__(BLOCK) k_tri_cell_order(const unsigned* start, long long cells, int* tri,
                                                          int* bigCount, int* big)
{
    const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
    bool large = false;
    if (c < cells) {
        const unsigned b = start[c], m = start[c + 1] - b;
        if (m > (unsigned)CELL_SORT) {
            large = true;
        } else if (m > 1) {
            int v[CELL_SORT];
#pragma unroll
            for (int j = 0; j < CELL_SORT; ++j) v[j] = (unsigned)j < m ? tri[b + j] : INT_MAX;
#pragma unroll
            for (int r = 0; r < CELL_SORT; ++r)
#pragma unroll
                for (int j = r & 1; j + 1 < CELL_SORT; j += 2) {
                    const int lo = min(v[j], v[j + 1]), hi = max(v[j], v[j + 1]);
                    v[j] = lo;
                    v[j + 1] = hi;
                }
#pragma unroll
            for (int j = 0; j < CELL_SORT; ++j)
                if ((unsigned)j < m) tri[b + j] = v[j];
        }
    }
    wave_append(large, (int)c, bigCount, big);
}

// One 64-lane block per listed cell (blocks stride over the list): each
// entry goes to its rank among the cell's (distinct) triangle indices, the
// cell staged in LDS (through `tmp` in global memory for a cell beyond the
// LDS stage).
__global__ void __launch_bounds__(64) k_tri_cell_order_big(const unsigned* start, const int* bigCount,
                                                           const int* big, int* tri, int* tmp)
{
    constexpr unsigned kLds = 4096;  // cells up to this size are ranked from LDS
    __shared__ int sh[kLds];
    const int count = *bigCount;
    for (int k = blockIdx.x; k < count; k += gridDim.x) {
        const long long c = big[k];
        const unsigned b = start[c], m = start[c + 1] - b;
        if (m <= kLds) {
            for (unsigned i = threadIdx.x; i < m; i += 64) sh[i] = tri[b + i];
            __syncthreads();
            for (unsigned i = threadIdx.x; i < m; i += 64) {
                const int v = sh[i];
                unsigned rank = 0;
                for (unsigned j = 0; j < m; ++j) rank += sh[j] < v;
                tri[b + rank] = v;
            }
        } else {
            for (unsigned i = threadIdx.x; i < m; i += 64) {
                const int v = tri[b + i];
                unsigned rank = 0;
                for (unsigned j = 0; j < m; ++j) rank += tri[b + j] < v;
                tmp[b + rank] = v;
            }
            __syncthreads();
            for (unsigned i = threadIdx.x; i < m; i += 64) tri[b + i] = tmp[b + i];
        }
        __syncthreads();
    }
}

// Exclusive scan of n unsigned values: tiles of SCAN_TILE per block (block
// scan of per-thread sums), a one-block scan of the tile totals, then the
// tile offsets added (and, with `copy`, the result duplicated there).
__global__ void __launch_bounds__(BLOCK) k_scan_tiles(const unsigned* in, long n, unsigned* out, unsigned* tileSum)
{
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    unsigned v[SCAN_ITEMS];
    unsigned sum = 0;
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        sum += v[j];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(sum, &total);
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}
__global__ void __launch_bounds__(BLOCK) k_scan_top(unsigned* tileSum, long tiles)
{
    unsigned carry = 0;
    for (long base = 0; base < tiles; base += SCAN_TILE) {
        const long b = base + (long)threadIdx.x * SCAN_ITEMS;
        unsigned v[SCAN_ITEMS];
        unsigned sum = 0;
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = b + j < tiles ? tileSum[b + j] : 0u;
            sum += v[j];
        }
        unsigned total;
        unsigned run = carry + block_exclusive_scan(sum, &total);
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            if (b + j < tiles) tileSum[b + j] = run;
            run += v[j];
        }
        carry += total;
    }
}
__global__ void __launch_bounds__(BLOCK) k_scan_add(unsigned* out, long n, const unsigned* tileOff, unsigned* copy)
{
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
        const unsigned v = out[i] + tileOff[i / SCAN_TILE];
        out[i] = v;
        if (copy) copy[i] = v;
    }
}

// ------------------------------------------------------------ downloads
__global__ void __launch_bounds__(BLOCK) k_gather_state(Soa S, double* out /*[13][n] by row*/)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    const long o = S.k(I_ORIG, p);
    const int fields[13] = {F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_WX, F_WY, F_WZ, F_QW, F_QX, F_QY, F_QZ};
    for (int f = 0; f < 13; ++f) out[(long)f * S.n + o] = S.f(fields[f], p);
}

__global__ void __launch_bounds__(BLOCK) k_gather_rows(Soa S, const double* src, int nrows,
                                                       double* out)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= S.n) return;
    const long o = S.k(I_ORIG, p);
    for (int r = 0; r < nrows; ++r) out[(long)r * S.n + o] = src[(long)r * S.n + p];
}

__global__ void __launch_bounds__(BLOCK) k_byorig(Soa S, int* byorig)
{
    const long p = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (p < S.n) byorig[S.k(I_ORIG, p)] = (int)p;
}

__global__ void __launch_bounds__(BLOCK) k_fill(double* a, long count, double v)
{
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < count; i += (long)gridDim.x * BLOCK)
        a[i] = v;
}
__global__ void __launch_bounds__(BLOCK) k_fill_u64(unsigned long long* a, long count)
{
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < count; i += (long)gridDim.x * BLOCK)
        a[i] = 0ull;
}
__global__ void __launch_bounds__(BLOCK) k_fill_u32(unsigned* a, long count)
{
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < count; i += (long)gridDim.x * BLOCK)
        a[i] = 0u;
}

// debug: reference-space grid = packed key and input row, in sorted order
__global__ void __launch_bounds__(BLOCK) k_debug_grid(Soa S, const int* sorted,
                                                      unsigned long long* keys, int* ids)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= S.n) return;
    const long p = sorted[t];
    keys[t] = S.k(I_VALID, p) ? pack_cell(S.k(I_CX, p), S.k(I_CY, p), S.k(I_CZ, p)) : 0ull;
    ids[t] = S.k(I_ORIG, p);
}

int blocks_for(long n) { return (int)((n + BLOCK - 1) / BLOCK); }

}  // namespace

// ============================================================ solver object
struct DiffRun;                 // differentiable-run state (dem_diff.inl)
void diff_release(DiffRun* r);
struct GraphCache;              // captured step sequences (dem_graph.inl)
void graph_release(GraphCache* g);
// timing events: grid build | contact scan + list compaction | pair law |
// owner sums (triangles, walls, field, acceleration) | body reductions
constexpr int NEV = 6;

// Per-step buffers of the triangle grid: sorted (key, triangle) pairs and
// their radix-sort counts, or the dense cell table with its fill cursors and
// scan partials.
struct TriBufs {
    unsigned long long* keyA = nullptr;
    unsigned long long* keyB = nullptr;
    int* triA = nullptr;
    int* triB = nullptr;
    unsigned* counts = nullptr;  // radix-sort digit counts [256][countBlocks]
    int countBlocks = 0;
    long long cap = 0;           // entries keyA..triB hold
    unsigned* start = nullptr;   // dense cell -> first entry table of the grid's box
    unsigned* cur = nullptr;     // per-cell counts, then fill cursors (startCap entries)
    unsigned* scan = nullptr;    // tile totals of the cell-table scan
    long long startCap = 0;      // entries start and cur hold
};

struct dem_solver {
    bool initialized = false;
    int device = 0;
    hipStream_t stream = nullptr;
    long n = 0;
    int nb = 0, nt = 0, nw = 0;
    bool offs = false;
    double maxRadius = 0.0, meshSize = 0.0, inv = 0.0;
    Material pp{}, pt{}, pw{};
    double g[3] = {0, 0, 0};
    int gravity = 0, universal = 0;
    int stepIndex = 0;
    int reorderInterval = 8;  // steps between physical reorders (DEM_REORDER_INTERVAL)
    int pairBlocks = PAIR_BLOCKS;
    // candidate lists with a skin (pure DEM): rebuild only when needed
    bool verlet = false;       // active
    bool vlistValid = false;   // lists built and not yet invalidated
    double skin = 0.0;
    // skin sized to the motion (no DEM_SKIN given): every so often the host
    // reads the largest displacement since the rebuild and, if a much
    // thinner skin would still last kSkinSteps steps, rebuilds with it; a
    // rebuild that comes sooner than a quarter of that widens the skin
    bool skinAdapt = false;
    double skinMin = 0.0, skinMax = 0.0;
    long sinceRebuild = 0, nextSkinCheck = 0;
    double* dispMax = nullptr;  // [blocks] k_cells' largest displacement since the rebuild
    double* dDisp = nullptr;    // the reduced maximum
    double* hDisp = nullptr;    // pinned copy
    double* refC = nullptr;    // [3][n] centres at the last rebuild
    int* cand = nullptr;       // [CAND_MAX][n]
    int* ncand = nullptr;      // [n]
    int4* meta = nullptr;      // [n] cell and input row in cell order
    int* hFlag = nullptr;      // pinned: displacement / overflow flags
    int* hMirror = nullptr;    // pinned, coherent: [seq, need_rebuild, grid_overflow] written by k_vscan
    int* dMirror = nullptr;    // device address of hMirror
    int mirrorSeq = 0;
    History H2{};              // second history buffer (candidate-list rebuilds move rows)
    long rebuilds = 0;
    // run counters (dem_debug_run_counters): force evaluations, evaluations
    // that sorted the particle grid on the device, steps of dem_run_steps
    // replayed from graphs and launched eagerly
    long long nEvaluations = 0, nSorts = 0, nGraphSteps = 0, nEagerSteps = 0;
    bool lastSorted = false;  // the last evaluation sorted the grid on the device
    bool verletBroken = false;  // a rebuild overflowed CAND_MAX: lists stay off
    bool keepRows = false;      // list rebuilds leave the rows in place (differentiable runs)
    bool capturing = false;   // inside a HIP graph capture: no event records
    bool countAlways = false; // per-step counters on every step (differentiable runs read them)
    GraphCache* graphs = nullptr;
    uint64_t intervalSteps = 0;
    size_t freeBefore = 0;
    bool restored = false;
    bool historyLoaded = false;  // a checkpoint filled the history tables
    bool legacySphere = false;  // one zero-offset component per owner, id = owner id
    // clumps (an owner with several sphere components): the contact pipeline
    // runs over component rows [nc] that carry copies of their owner's state
    bool clump = false;
    long nc = 0;                 // contact rows: components (== n without clumps)
    double* cState = nullptr;    // [NDF][nc] component rows
    int* ciState = nullptr;      // [NIF][nc]
    int* cOwner = nullptr;       // [nc] owner row of each component
    int* ownerStart = nullptr;   // [n + 1] first component of each owner
    int* cRank = nullptr;        // [nc] sorted position of each component
    int* ownS = nullptr;         // [nc] owner of the t-th sorted component

    // device memory
    double* dState[2] = {nullptr, nullptr};
    int* iState[2] = {nullptr, nullptr};
    int cur = 0;
    Scalars* sc = nullptr;
    unsigned long long* keyA = nullptr;
    unsigned long long* keyB = nullptr;
    int* valA = nullptr;
    int* valB = nullptr;
    int* ident = nullptr;                 // [n] identity (list steps index rows directly)
    unsigned long long* mKey = nullptr;   // [n] Morton keys of a candidate-list rebuild
    int* mVal = nullptr;                  // [n] sorted positions in Morton order
    unsigned* rsCounts = nullptr;
    unsigned* rsDigits = nullptr;  // [256] digit totals, [256] digit bases
    int rsBlocks = 0;
    unsigned* cellStart = nullptr;
    int* boxPart = nullptr;
    int* partScene = nullptr;   // [blocks] scene of a one-scene block, -1 otherwise
    int* sceneAcc = nullptr;    // [scenes][6] box of blocks that straddle scenes
    unsigned* gridDone = nullptr;
    int nscenes = 1;
    SceneParams* dSp = nullptr;
    SceneGrid* dGrid = nullptr;
    Material* dMatPP = nullptr;
    Material hMatPP0{};  // the first scene's particle-particle law (k_pairs of one scene)
    Material* dMatPW = nullptr;
    std::vector<SceneParams> hsp;
    unsigned long long cellCap = 0;
    double4* posr = nullptr;
    uint32_t* cidS = nullptr;
    Rec* rec = nullptr;        // contact records, kinematic part
    RecAux* recAux = nullptr;  // contact records, constant part
    Recs recs() const { return Recs{rec, recAux}; }
    bool pendingFinish = false;
    bool recFresh = false;   // the contact records were written by the last advance (zero-offset rows)
    double pendingDt = 0.0;
    int* nbr = nullptr;      // k-major contact lists [MAX_LIST][n]
    int* ncount = nullptr;   // list length per owner (sorted index)
    unsigned* listOff = nullptr;  // [n] owner's first position in the compacted list
    unsigned* listBlk = nullptr;  // [2][blocks] per-block offsets, then per-block high-water marks
    int2* pairs = nullptr;   // compacted (owner, entry) [MAX_LIST * n]
    double* pairF = nullptr; // per-contact force and torque term [MAX_LIST * n][6]
    double* rowF = nullptr;  // [n][6] contact sums of rows summed in k_pairs (single-sphere runs)
    int* byorig = nullptr;
    Tape tape{};  // set by a differentiable run while it records a step
    DiffRun* diff = nullptr;
    History H{};
    double* wallBuf = nullptr;
    Outputs O{};
    // triangles
    int* tBody = nullptr;
    int* triWork = nullptr;  // [nr] sorted positions of the rows near triangles
    unsigned long long* triHitMask = nullptr;  // [nr] per listed row: entries with a positive area scale
    int* triHitN = nullptr;                    // [nr] hits stored, or -1 (evaluated in the lane)
    double2* triHitVal = nullptr;              // [nr][TRI_HITS] (area scale, plane distance)
    double* tLocal = nullptr;
    double* tWorld = nullptr;
    double* tAabb = nullptr;
    double* pose = nullptr;
    double* bodyMass = nullptr;
    std::vector<double> hBodyMass;  // host copy (fixed at initialisation)
    int* tCounts = nullptr;
    long long* tOffsets = nullptr;
    // the triangle grid's per-step buffers: the solver's own, grown on demand
    // (tOwn), and the set this step uses (tUse): tOwn, or one carved out of the
    // device's static scratch array by dem_compute_forces_noalloc
    TriBufs tOwn{};
    TriBufs tUse{};
    bool tInScratch = false;       // tUse points into the scratch array
    int tEntries = 0;
    int tLo[3] = {0, 0, 0};        // cell box of the triangle grid entries
    long long tDims[3] = {0, 0, 0};
    long long tStartCap = 0;       // entries of the table valid this step (0: binary search)
    int* tBits = nullptr;          // device: key bits of the triangle sort
    int* hTri = nullptr;           // pinned: [nt] counts, then the sort bits
    long long* hTriOff = nullptr;  // pinned: [nt] entry offsets
    int* d64 = nullptr;  // device constant 64 (triangle sort bits)
    unsigned long long* regimeInstant = nullptr;
    unsigned long long* regimeInterval = nullptr;
    double* microInstant = nullptr;
    double* microInterval = nullptr;
    double* bodyResult = nullptr;  // [nb*6 + nb]  loads, then fz per body; microInstant follows
    double* redPartial = nullptr;
    int redBlocks = 0;
    double* dl = nullptr;  // download staging [max rows][n]
    // pinned
    double* hPose = nullptr;
    double* hLoads = nullptr;
    Scalars* hScalars = nullptr;
    // cached stats
    dem_step_stats cached{};
    std::vector<uint64_t> cRegInst, cRegInt;
    std::vector<double> cMicInst, cMicInt;
    bool keepBodyParticle = false, keepBodyDetail = false;
    // timing
    hipEvent_t ev[NEV] = {};
    hipEvent_t flagEv = nullptr;  // the scalars copy with the displacement flag has landed
    bool evReady = false;
    bool timing = false;  // record the stage events of dem_debug_last_timings
    bool scalarsReset = false;  // the last advance reset the step counters (no k_step_init needed)
    double nextDt = 0.0;        // set by a multi-step run while another step follows this evaluation
    bool advancedAhead = false; // the last evaluation also finished its step and advanced the next one
    // metadata for checkpoints
    std::vector<int> compIds, ownerIds;  // per component: id, owner id
    std::vector<int> pids;               // per owner: id
    std::vector<double> compR, compB;

    Soa soa() const { return Soa{dState[cur], iState[cur], n}; }
    // the rows the grid, candidate scan and pair law run over
    Soa rows() const { return clump ? Soa{cState, ciState, nc} : soa(); }

    ~dem_solver()
    {
        if (stream) hipStreamSynchronize(stream);
        diff_release(diff);
        graph_release(graphs);
        void* ptrs[] = {dState[0], dState[1], iState[0], iState[1], sc, keyA, keyB, valA, valB, ident, mKey, mVal,
                        rec, recAux, refC, cand, ncand, meta, dispMax, dDisp, rsCounts, rsDigits, cellStart, boxPart, partScene, sceneAcc, gridDone, dSp, dGrid,
                        dMatPP, dMatPW, posr, cidS, nbr, ncount, listOff, listBlk, pairs, pairF, rowF, byorig, H.ppOcc, H.ppSeen, H.ppKey,
                        H.ppVal, H.pwOcc, H.pwSeen, H.pwKey, H.pwVal, H.ptOcc, H.ptSeen, H.ptKey,
                        H.ptVal, H2.ppOcc, H2.ppSeen, H2.ppKey, H2.ppVal, H2.pwOcc, H2.pwSeen, H2.pwKey,
                        H2.pwVal, H2.ptOcc, H2.ptSeen, H2.ptKey, H2.ptVal, wallBuf, O.contact, O.wall, O.body, O.field, O.bodyRows, O.bodyTouched,
                        O.bodyParticle, O.bodyDetail, O.regime, tBody, triWork, triHitMask, triHitN, triHitVal, tLocal, tWorld, tAabb,
                        pose, bodyMass, tCounts, tOffsets, tOwn.keyA, tOwn.keyB, tOwn.triA, tOwn.triB, tOwn.counts, d64,
                        regimeInstant, regimeInterval, microInterval, bodyResult,
                        redPartial, dl, cState, ciState, cOwner, ownerStart, cRank, ownS, tOwn.start, tOwn.cur, tOwn.scan, tBits};
        for (void* p : ptrs)
            if (p) hipFree(p);
        if (hPose) hipHostFree(hPose);
        if (hLoads) hipHostFree(hLoads);
        if (hScalars) hipHostFree(hScalars);
        if (hFlag) hipHostFree(hFlag);
        if (hMirror) hipHostFree(hMirror);
        if (hDisp) hipHostFree(hDisp);
        if (hTri) hipHostFree(hTri);
        if (hTriOff) hipHostFree(hTriOff);
        for (auto& e : ev)
            if (e) hipEventDestroy(e);
        if (flagEv) hipEventDestroy(flagEv);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
int dalloc(T*& p, size_t count, const char* what)
{
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)&p, sizeof(T) * count), what);
    return DEM_OK;
}

#define TRY(x)                 \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

int launch_check(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DEM_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return DEM_OK;
}

// Stable LSD radix sort of (key, val) pairs held in (ka, va) in 8-bit
// digits; (kb, vb) is scratch.  Passes over bits >= *nbits (device) return at
// once on the device, and `passes` bounds the launches when the host knows
// the bits.
int radix_sort(hipStream_t st, unsigned long long* ka, int* va, unsigned long long* kb, int* vb,
               long n, unsigned* counts, int nblocks, const int* nbits_dev, unsigned* digits, int passes = 8)
{
    if (n <= 0) return DEM_OK;
    for (int shift = 0; shift < 8 * passes; shift += 8) {
        const bool even = ((shift / 8) & 1) == 0;
        const unsigned long long* kin = even ? ka : kb;
        const int* vin = even ? va : vb;
        unsigned long long* kout = even ? kb : ka;
        int* vout = even ? vb : va;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(BLOCK), 0, st, kin, n, shift, nbits_dev,
                           counts, nblocks);
        hipLaunchKernelGGL(k_rs_scan_rows, dim3(256), dim3(BLOCK), 0, st, counts, nblocks, shift, nbits_dev,
                           digits);
        hipLaunchKernelGGL(k_rs_scan_base, dim3(1), dim3(BLOCK), 0, st, digits, shift, nbits_dev);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(BLOCK), 0, st, kin, vin, kout, vout, n,
                           shift, nbits_dev, counts, nblocks, digits + 256);
    }
    const int cb = (int)std::min<long>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_rs_copyback, dim3(cb), dim3(BLOCK), 0, st, ka, va, kb, vb, n, nbits_dev);
    return launch_check("radix sort");
}

bool valid_mech(const dem_mech* m) { return m != nullptr; }

Material make_material(const dem_mech& m, double dt)
{
    Material r;
    r.epsS = m.epsS; r.mu = m.mu; r.epsN = m.epsN; r.cohesion = m.cohesion; r.beta = m.beta;
    r.muT = m.muT; r.muR = m.muR; r.kN = m.kN; r.dt = dt;
    r.logN = std::log(m.epsN);
    r.logS = std::log(m.epsS);
    return r;
}

}  // namespace

// ============================================================ C-ABI
extern "C" {

const char* dem_last_error(void) { return g_error; }

int dem_backend_available(void)
{
    int count = 0;
    return hipGetDeviceCount(&count) == hipSuccess && count > 0 ? 1 : 0;
}

const char* dem_device_description(void)
{
    static char desc[256] = {0};
    if (desc[0] == '\0') {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
            snprintf(desc, sizeof desc, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName,
                     prop.multiProcessorCount,
                     (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
        } else {
            snprintf(desc, sizeof desc, "unavailable");
        }
    }
    return desc;
}

int dem_create(dem_solver** out)
{
    if (!out) return set_error(DEM_E_INVALID, "null output pointer");
    *out = new (std::nothrow) dem_solver();
    return *out ? DEM_OK : set_error(DEM_E_INVALID, "out of memory");
}

void dem_destroy(dem_solver* s) { delete s; }

namespace {

// Scene layout of a solver: rows and walls of each scene, per-scene
// particle-particle and particle-wall laws.
struct SceneSpec {
    int count;
    const int64_t* row;   // [count + 1]
    const int32_t* wall;  // [count + 1]
    const dem_mech* pp;   // [count]
    const dem_mech* pw;   // [count]
};

int initialize_impl(dem_solver* s, const dem_particles_view* P, const dem_components_view* C,
                    const dem_walls_view* W, const dem_bodies_view* B, const dem_mech* mpt,
                    const dem_control* ctl, const SceneSpec& spec)
{
    if (s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver already initialized");
    if (P->count <= 0) return set_error(DEM_E_INVALID, "CUDA DEM solver requires particles");
    if (P->count > MAX_PARTICLES)
        return set_error(DEM_E_CAPACITY, "particle count exceeds the contact-list row range (2^25 - 1)");
    const long n = P->count;
    const int nsc = spec.count;
    const int nwall = W ? W->count : 0;
    if (nsc < 1 || spec.row[0] != 0 || spec.row[nsc] != n || spec.wall[0] != 0 || spec.wall[nsc] != nwall)
        return set_error(DEM_E_INVALID, "scene row/wall ranges do not cover the particle and wall views");
    for (int k = 0; k < nsc; ++k) {
        if (spec.row[k + 1] <= spec.row[k] || spec.wall[k + 1] < spec.wall[k])
            return set_error(DEM_E_INVALID, "scene %d has an empty or reversed row range", k);
        if (!valid_mech(&spec.pp[k]) || !valid_mech(&spec.pw[k]))
            return set_error(DEM_E_INVALID, "null argument");
    }
    if (nsc > 1 && B && B->count > 0)
        return set_error(DEM_E_UNSUPPORTED, "batched scenes cannot carry bodies");
    if (nsc > 1 && ctl->universal_gravity)
        return set_error(DEM_E_UNSUPPORTED, "batched scenes cannot use universal gravitation");
    for (int k = 0; k < nsc; ++k) {
        std::vector<int> ids(P->id + spec.row[k], P->id + spec.row[k + 1]);
        std::sort(ids.begin(), ids.end());
        for (size_t i = 1; i < ids.size(); ++i)
            if (ids[i] == ids[i - 1])
                return set_error(DEM_E_INVALID,
                                 "duplicate particle ID %d; CUDA contact history requires stable unique particle IDs",
                                 ids[i]);
    }
    if (C->count <= 0)
        return set_error(DEM_E_INVALID, "CUDA clump topology requires owners and components");
    // components sorted by (owner, id) (backend:121-190)
    std::vector<int> order(C->count);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (C->owner[a] != C->owner[b]) return C->owner[a] < C->owner[b];
        return C->id[a] < C->id[b];
    });
    std::vector<int> perOwner(n, 0), sceneOf(n, 0);
    for (int k = 0; k < nsc; ++k)
        for (int64_t r = spec.row[k]; r < spec.row[k + 1]; ++r) sceneOf[r] = k;
    std::vector<double> maxrScene(nsc, 0.0);
    double maxr = 0.0;
    for (int k : order) {
        const int ow = C->owner[k];
        if (ow < 0 || ow >= n || C->id[k] < 0 || !std::isfinite(C->radius[k]) || C->radius[k] <= 0.0)
            return set_error(DEM_E_INVALID, "invalid native clump component topology");
        ++perOwner[ow];
        maxr = std::max(maxr, C->radius[k]);
        maxrScene[sceneOf[ow]] = std::max(maxrScene[sceneOf[ow]], C->radius[k]);
    }
    for (long i = 0; i < n; ++i)
        if (perOwner[i] < 1 || perOwner[i] > 8)
            return set_error(DEM_E_INVALID, "each CUDA owner requires 1 to 8 components");
    // several components on some owner: the contact pipeline runs over component rows
    const long nc = C->count;
    const bool clump = nc != n;
    if (clump && nsc > 1) return set_error(DEM_E_UNSUPPORTED, "batched scenes cannot carry clumps");
    if (nc > MAX_PARTICLES)
        return set_error(DEM_E_CAPACITY, "component count exceeds the contact-list row range (2^25 - 1)");
    if (nc > (long)INT_MAX / PP_SLOTS)
        return set_error(DEM_E_CAPACITY, "particle count exceeds CUDA contact-history index capacity");
    const long nr = nc;  // contact rows
    if (!dem_backend_available()) return set_error(DEM_E_DEVICE, "no CUDA device is available");

    HIP_TRY(hipSetDevice(0), "hipSetDevice");
    HIP_TRY(hipFree(nullptr), "initialize HIP context");
    size_t total = 0;
    HIP_TRY(hipMemGetInfo(&s->freeBefore, &total), "query HIP memory before allocation");
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : s->ev) HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    HIP_TRY(hipEventCreateWithFlags(&s->flagEv, hipEventDisableTiming), "hipEventCreate");

    s->n = n;
    s->nc = nc;
    s->clump = clump;
    s->nscenes = nsc;
    const double scale = ctl->mesh_scale > 0.0 ? ctl->mesh_scale : 2.5;
    s->maxRadius = maxr;
    s->meshSize = scale * maxr;
    s->inv = 1.0 / s->meshSize;
    s->pp = make_material(spec.pp[0], ctl->dt);
    s->pt = make_material(*mpt, ctl->dt);
    s->pw = make_material(spec.pw[0], ctl->dt);
    s->hsp.resize(nsc);
    std::vector<Material> hpp(nsc), hpw(nsc);
    for (int k = 0; k < nsc; ++k) {
        SceneParams& q = s->hsp[k];
        q.inv = 1.0 / (scale * maxrScene[k]);
        q.maxr = maxrScene[k];
        q.row0 = spec.row[k];
        q.row1 = clump ? nc : spec.row[k + 1];  // grid rows of the scene
        q.wall0 = spec.wall[k];
        q.nwall = spec.wall[k + 1] - spec.wall[k];
        hpp[k] = make_material(spec.pp[k], ctl->dt);
        hpw[k] = make_material(spec.pw[k], ctl->dt);
    }
    s->g[0] = ctl->gravity[0]; s->g[1] = ctl->gravity[1]; s->g[2] = ctl->gravity[2];
    s->gravity = ctl->gravity_enabled ? 1 : 0;
    s->universal = ctl->universal_gravity ? 1 : 0;
    s->keepBodyParticle = ctl->keep_body_particle != 0;
    s->keepBodyDetail = ctl->keep_body_detail != 0;

    // components in (owner, id) order; without clumps component c belongs to
    // owner row c and its fields live in the owner's row
    s->compIds.assign(nc, 0); s->ownerIds.assign(nc, 0); s->compR.assign(nc, 0.0); s->compB.assign(3 * nc, 0.0);
    s->pids.assign(P->id, P->id + n);
    std::vector<int> cown(nc), ostart(n + 1, 0);
    for (long c = 0; c < nc; ++c) {
        const int k = order[c];
        const int ow = C->owner[k];
        cown[c] = ow;
        ++ostart[ow + 1];
        s->compIds[c] = C->id[k]; s->ownerIds[c] = P->id[ow]; s->compR[c] = C->radius[k];
        for (int a = 0; a < 3; ++a) s->compB[3 * c + a] = C->pos_body[3 * k + a];
        if (s->compB[3 * c] != 0.0 || s->compB[3 * c + 1] != 0.0 || s->compB[3 * c + 2] != 0.0) s->offs = true;
    }
    for (long i = 0; i < n; ++i) ostart[i + 1] += ostart[i];
    // host SoA in input row order
    std::vector<double> hd((size_t)NDF * n, 0.0);
    std::vector<int> hi((size_t)NIF * n, 0);
    auto D = [&](int f, long p) -> double& { return hd[(size_t)f * n + p]; };
    auto Iv = [&](int f, long p) -> int& { return hi[(size_t)f * n + p]; };
    if (!clump)
        for (long p = 0; p < n; ++p) {
            D(F_CR, p) = s->compR[p];
            D(F_BX, p) = s->compB[3 * p]; D(F_BY, p) = s->compB[3 * p + 1]; D(F_BZ, p) = s->compB[3 * p + 2];
            Iv(I_CID, p) = s->compIds[p];
        }
    s->legacySphere = !clump;
    for (long p = 0; p < n && s->legacySphere; ++p)
        s->legacySphere = s->compIds[p] == P->id[p] && s->compR[p] == P->radius[p] && s->compB[3 * p] == 0.0 &&
                          s->compB[3 * p + 1] == 0.0 && s->compB[3 * p + 2] == 0.0;
    for (long p = 0; p < n; ++p) {
        D(F_X, p) = P->pos[3 * p]; D(F_Y, p) = P->pos[3 * p + 1]; D(F_Z, p) = P->pos[3 * p + 2];
        D(F_VX, p) = P->vel[3 * p]; D(F_VY, p) = P->vel[3 * p + 1]; D(F_VZ, p) = P->vel[3 * p + 2];
        D(F_WX, p) = P->ang_vel[3 * p]; D(F_WY, p) = P->ang_vel[3 * p + 1]; D(F_WZ, p) = P->ang_vel[3 * p + 2];
        D(F_QW, p) = P->quat[4 * p]; D(F_QX, p) = P->quat[4 * p + 1];
        D(F_QY, p) = P->quat[4 * p + 2]; D(F_QZ, p) = P->quat[4 * p + 3];
        D(F_MASS, p) = P->mass[p];
        D(F_IX, p) = P->inertia[3 * p]; D(F_IY, p) = P->inertia[3 * p + 1]; D(F_IZ, p) = P->inertia[3 * p + 2];
        Iv(I_PID, p) = P->id[p];
        Iv(I_STATUS, p) = P->status[p];
        Iv(I_ORIG, p) = (int)p;
        Iv(I_HIST, p) = (int)p;
        Iv(I_SCENE, p) = sceneOf[p];
    }
    for (int b = 0; b < 2; ++b) {
        TRY(dalloc(s->dState[b], (size_t)NDF * n, "hipMalloc(particle state)"));
        TRY(dalloc(s->iState[b], (size_t)NIF * n, "hipMalloc(particle ints)"));
    }
    HIP_TRY(hipMemcpy(s->dState[0], hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice),
            "upload particle state");
    HIP_TRY(hipMemcpy(s->iState[0], hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice),
            "upload particle ints");
    s->cur = 0;
    if (clump) {
        // component rows: geometry and ids here, owner state copied each step
        std::vector<double> cd((size_t)NDF * nc, 0.0);
        std::vector<int> ci((size_t)NIF * nc, 0);
        for (long c = 0; c < nc; ++c) {
            cd[(size_t)F_CR * nc + c] = s->compR[c];
            cd[(size_t)F_BX * nc + c] = s->compB[3 * c];
            cd[(size_t)F_BY * nc + c] = s->compB[3 * c + 1];
            cd[(size_t)F_BZ * nc + c] = s->compB[3 * c + 2];
            ci[(size_t)I_CID * nc + c] = s->compIds[c];
            ci[(size_t)I_PID * nc + c] = s->ownerIds[c];
            ci[(size_t)I_ORIG * nc + c] = (int)c;  // grid ties ordered by component index
            ci[(size_t)I_HIST * nc + c] = (int)c;
            ci[(size_t)I_SCENE * nc + c] = 0;
        }
        TRY(dalloc(s->cState, cd.size(), "hipMalloc(component rows)"));
        TRY(dalloc(s->ciState, ci.size(), "hipMalloc(component rows)"));
        TRY(dalloc(s->cOwner, nc, "hipMalloc(component owners)"));
        TRY(dalloc(s->ownerStart, n + 1, "hipMalloc(component owners)"));
        TRY(dalloc(s->cRank, nc, "hipMalloc(component ranks)"));
        TRY(dalloc(s->ownS, nc, "hipMalloc(component owners)"));
        HIP_TRY(hipMemcpy(s->cState, cd.data(), sizeof(double) * cd.size(), hipMemcpyHostToDevice), "upload components");
        HIP_TRY(hipMemcpy(s->ciState, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice), "upload components");
        HIP_TRY(hipMemcpy(s->cOwner, cown.data(), sizeof(int) * nc, hipMemcpyHostToDevice), "upload components");
        HIP_TRY(hipMemcpy(s->ownerStart, ostart.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice),
                "upload components");
    }
    TRY(dalloc(s->sc, 1, "hipMalloc(step scalars)"));
    HIP_TRY(hipMemset(s->sc, 0, sizeof(Scalars)), "clear step scalars");
    TRY(dalloc(s->keyA, nr, "hipMalloc(grid keys)"));
    TRY(dalloc(s->keyB, nr, "hipMalloc(grid keys)"));
    TRY(dalloc(s->valA, nr, "hipMalloc(grid ids)"));
    TRY(dalloc(s->valB, nr, "hipMalloc(grid ids)"));
    s->rsBlocks = (int)((nr + RS_TILE - 1) / RS_TILE);
    TRY(dalloc(s->rsCounts, (size_t)256 * s->rsBlocks, "hipMalloc(sort counts)"));
    TRY(dalloc(s->rsDigits, 512, "hipMalloc(sort digit totals)"));
    s->cellCap = (unsigned long long)std::max<long>(8 * nr + 64, 1 << 20);
    TRY(dalloc(s->cellStart, s->cellCap, "hipMalloc(cell table)"));
    TRY(dalloc(s->posr, nr, "hipMalloc(packed positions)"));
    TRY(dalloc(s->cidS, nr, "hipMalloc(sorted component ids)"));
    TRY(dalloc(s->rec, nr, "hipMalloc(contact records)"));
    TRY(dalloc(s->recAux, nr, "hipMalloc(contact records)"));
    TRY(dalloc(s->boxPart, (size_t)6 * blocks_for(nr), "hipMalloc(grid box partials)"));
    TRY(dalloc(s->partScene, blocks_for(nr), "hipMalloc(grid box partials)"));
    TRY(dalloc(s->sceneAcc, (size_t)6 * nsc, "hipMalloc(scene boxes)"));
    TRY(dalloc(s->gridDone, 1, "hipMalloc(grid counter)"));
    HIP_TRY(hipMemset(s->gridDone, 0, sizeof(unsigned)), "clear grid counter");
    TRY(dalloc(s->dSp, nsc, "hipMalloc(scene table)"));
    TRY(dalloc(s->dGrid, nsc, "hipMalloc(scene grids)"));
    TRY(dalloc(s->dMatPP, nsc, "hipMalloc(scene materials)"));
    TRY(dalloc(s->dMatPW, nsc, "hipMalloc(scene materials)"));
    HIP_TRY(hipMemcpy(s->dSp, s->hsp.data(), sizeof(SceneParams) * nsc, hipMemcpyHostToDevice), "upload scenes");
    HIP_TRY(hipMemcpy(s->dMatPP, hpp.data(), sizeof(Material) * nsc, hipMemcpyHostToDevice), "upload scenes");
    s->hMatPP0 = hpp[0];
    HIP_TRY(hipMemcpy(s->dMatPW, hpw.data(), sizeof(Material) * nsc, hipMemcpyHostToDevice), "upload scenes");
    TRY(dalloc(s->nbr, (size_t)MAX_LIST * nr, "hipMalloc(contact lists)"));
    TRY(dalloc(s->ncount, nr, "hipMalloc(contact counts)"));
    TRY(dalloc(s->listOff, nr + 1, "hipMalloc(contact list offsets)"));
    TRY(dalloc(s->listBlk, (size_t)2 * blocks_for(nr), "hipMalloc(contact list offsets)"));
    TRY(dalloc(s->pairs, (size_t)MAX_LIST * nr, "hipMalloc(compacted contacts)"));
    TRY(dalloc(s->pairF, (size_t)6 * MAX_LIST * nr, "hipMalloc(contact forces)"));
    if (!clump) TRY(dalloc(s->rowF, (size_t)6 * nr, "hipMalloc(contact sums)"));
    {
        // candidate lists (single-sphere owners): skin = DEM_SKIN (default 0.2) x the
        // smallest scene's largest radius; the +-2-cell rebuild scan must cover r_i + r_j + skin
        const char* e = getenv("DEM_SKIN");
        const double frac = e ? atof(e) : 0.2;
        double rmin = maxrScene[0];
        for (int k = 1; k < nsc; ++k) rmin = std::min(rmin, maxrScene[k]);
        s->skin = frac * rmin;
        bool ok = frac > 0.0 && !clump;
        for (int k = 0; k < nsc && ok; ++k) ok = 2.0 / s->hsp[k].inv >= 2.0 * maxrScene[k] + s->skin;
        if (ok && !e) {
            // the skin follows the motion unless DEM_SKIN fixes it (or DEM_SKIN_ADAPT=0)
            const char* a = getenv("DEM_SKIN_ADAPT");
            s->skinAdapt = !(a && atoi(a) == 0);
            s->skinMin = 0.00625 * rmin;
            s->skinMax = 0.4 * rmin;
            for (int k = 0; k < nsc; ++k) s->skinMax = std::min(s->skinMax, 2.0 / s->hsp[k].inv - 2.0 * maxrScene[k]);
            s->skinMax = std::max(s->skinMax, s->skin);
        }
        if (ok) {
            TRY(dalloc(s->refC, (size_t)3 * n, "hipMalloc(candidate lists)"));
            TRY(dalloc(s->cand, (size_t)CAND_MAX * n, "hipMalloc(candidate lists)"));
            TRY(dalloc(s->ncand, n, "hipMalloc(candidate lists)"));
            TRY(dalloc(s->meta, n, "hipMalloc(candidate lists)"));
            if (s->skinAdapt) {
                TRY(dalloc(s->dispMax, blocks_for(n), "hipMalloc(candidate lists)"));
                TRY(dalloc(s->dDisp, 1, "hipMalloc(candidate lists)"));
                HIP_TRY(hipHostMalloc((void**)&s->hDisp, sizeof(double), hipHostMallocDefault), "hipHostMalloc");
            }
            TRY(dalloc(s->ident, n, "hipMalloc(candidate lists)"));
            TRY(dalloc(s->mKey, n, "hipMalloc(candidate lists)"));
            TRY(dalloc(s->mVal, n, "hipMalloc(candidate lists)"));
            hipLaunchKernelGGL(k_identity, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->ident, n);
            HIP_TRY(hipHostMalloc((void**)&s->hFlag, sizeof(int) * 2, hipHostMallocDefault), "hipHostMalloc");
            s->hFlag[0] = s->hFlag[1] = 0;
            HIP_TRY(hipHostMalloc((void**)&s->hMirror, sizeof(int) * 4, hipHostMallocMapped | hipHostMallocCoherent),
                    "hipHostMalloc");
            std::memset(s->hMirror, 0, sizeof(int) * 4);
            HIP_TRY(hipHostGetDevicePointer((void**)&s->dMirror, s->hMirror, 0), "hipHostGetDevicePointer");
            s->verlet = true;
        }
    }
    TRY(dalloc(s->byorig, n, "hipMalloc(row index)"));
    // history
    s->H.rows = nr;
    TRY(dalloc(s->H.ppOcc, nr, "hipMalloc(particle history)"));
    TRY(dalloc(s->H.ppSeen, 2 * nr, "hipMalloc(particle history)"));
    TRY(dalloc(s->H.ppKey, (size_t)PP_SLOTS * nr, "hipMalloc(particle history keys)"));
    TRY(dalloc(s->H.ppVal, (size_t)PP_SLOTS * 9 * nr, "hipMalloc(particle history values)"));
    TRY(dalloc(s->H.pwOcc, nr, "hipMalloc(wall history)"));
    TRY(dalloc(s->H.pwSeen, 2 * nr, "hipMalloc(wall history)"));
    TRY(dalloc(s->H.pwKey, (size_t)PW_SLOTS * nr, "hipMalloc(wall history keys)"));
    TRY(dalloc(s->H.pwVal, (size_t)PW_SLOTS * 9 * nr, "hipMalloc(wall history values)"));
    HIP_TRY(hipMemset(s->H.ppOcc, 0, sizeof(unsigned long long) * nr), "clear history");
    HIP_TRY(hipMemset(s->H.ppSeen, 0, sizeof(unsigned long long) * 2 * nr), "clear history");
    HIP_TRY(hipMemset(s->H.pwOcc, 0, sizeof(unsigned) * nr), "clear history");
    HIP_TRY(hipMemset(s->H.pwSeen, 0, sizeof(unsigned) * 2 * nr), "clear history");
    HIP_TRY(hipMemset(s->H.ppKey, 0, sizeof(uint32_t) * PP_SLOTS * nr), "clear history");
    HIP_TRY(hipMemset(s->H.pwKey, 0, sizeof(uint32_t) * PW_SLOTS * nr), "clear history");
    // a second set of tables: reorders (candidate-list rebuilds and the
    // full scan's periodic cell ordering) move each row's history with it
    if (!clump) {
        s->H2.rows = nr;
        TRY(dalloc(s->H2.ppOcc, nr, "hipMalloc(particle history)"));
        TRY(dalloc(s->H2.ppSeen, 2 * nr, "hipMalloc(particle history)"));
        TRY(dalloc(s->H2.ppKey, (size_t)PP_SLOTS * nr, "hipMalloc(particle history keys)"));
        TRY(dalloc(s->H2.ppVal, (size_t)PP_SLOTS * 9 * nr, "hipMalloc(particle history values)"));
        TRY(dalloc(s->H2.pwOcc, nr, "hipMalloc(wall history)"));
        TRY(dalloc(s->H2.pwSeen, 2 * nr, "hipMalloc(wall history)"));
        TRY(dalloc(s->H2.pwKey, (size_t)PW_SLOTS * nr, "hipMalloc(wall history keys)"));
        TRY(dalloc(s->H2.pwVal, (size_t)PW_SLOTS * 9 * nr, "hipMalloc(wall history values)"));
    }
    // walls
    s->nw = W ? W->count : 0;
    TRY(dalloc(s->wallBuf, (size_t)6 * std::max(s->nw, 1), "hipMalloc(walls)"));
    if (s->nw > 0) {
        HIP_TRY(hipMemcpy(s->wallBuf, W->origin, sizeof(double) * 3 * s->nw, hipMemcpyHostToDevice), "upload walls");
        HIP_TRY(hipMemcpy(s->wallBuf + 3 * s->nw, W->normal, sizeof(double) * 3 * s->nw, hipMemcpyHostToDevice),
                "upload walls");
    }
    // outputs
    TRY(dalloc(s->O.contact, 6 * n, "hipMalloc(contact forces)"));
    TRY(dalloc(s->O.wall, 6 * n, "hipMalloc(wall forces)"));
    TRY(dalloc(s->O.body, 6 * n, "hipMalloc(body forces)"));
    TRY(dalloc(s->O.field, 3 * n, "hipMalloc(field forces)"));
    TRY(dalloc(s->dl, (size_t)19 * n, "hipMalloc(download staging)"));
    HIP_TRY(hipHostMalloc((void**)&s->hScalars, sizeof(Scalars), hipHostMallocDefault), "hipHostMalloc");
    std::memset(s->hScalars, 0, sizeof(Scalars));
    // bodies
    s->nb = B ? B->count : 0;
    long long ntl = 0;
    for (int b = 0; b < s->nb; ++b) {
        if (B->tri_count[b] < 0) return set_error(DEM_E_INVALID, "body triangle count cannot be negative");
        ntl += B->tri_count[b];
        if (ntl > INT_MAX) return set_error(DEM_E_CAPACITY, "total triangle count exceeds CUDA 32-bit index capacity");
    }
    s->nt = (int)ntl;
    if (s->nb > 0) {
        const int nb = s->nb;
        TRY(dalloc(s->H.ptOcc, nr, "hipMalloc(body history)"));
        TRY(dalloc(s->H.ptSeen, 2 * nr, "hipMalloc(body history)"));
        TRY(dalloc(s->H.ptKey, (size_t)PT_SLOTS * nr, "hipMalloc(body history keys)"));
        TRY(dalloc(s->H.ptVal, (size_t)PT_SLOTS * 9 * nr, "hipMalloc(body history values)"));
        HIP_TRY(hipMemset(s->H.ptOcc, 0, sizeof(unsigned long long) * nr), "clear history");
        HIP_TRY(hipMemset(s->H.ptSeen, 0, sizeof(unsigned long long) * 2 * nr), "clear history");
        HIP_TRY(hipMemset(s->H.ptKey, 0, sizeof(uint32_t) * PT_SLOTS * nr), "clear history");
        if (s->H2.ppOcc) {
            TRY(dalloc(s->H2.ptOcc, nr, "hipMalloc(body history)"));
            TRY(dalloc(s->H2.ptSeen, 2 * nr, "hipMalloc(body history)"));
            TRY(dalloc(s->H2.ptKey, (size_t)PT_SLOTS * nr, "hipMalloc(body history keys)"));
            TRY(dalloc(s->H2.ptVal, (size_t)PT_SLOTS * 9 * nr, "hipMalloc(body history values)"));
        }
        TRY(dalloc(s->pose, (size_t)13 * nb, "hipMalloc(body poses)"));
        HIP_TRY(hipMemcpy(s->pose, B->pose13, sizeof(double) * 13 * nb, hipMemcpyHostToDevice), "upload body poses");
        std::vector<int> tb(s->nt);
        for (int b = 0, t = 0; b < nb; ++b)
            for (int k = 0; k < B->tri_count[b]; ++k, ++t) tb[t] = b;
        TRY(dalloc(s->tBody, s->nt, "hipMalloc(triangles)"));
        TRY(dalloc(s->tLocal, (size_t)15 * s->nt, "hipMalloc(triangles)"));
        TRY(dalloc(s->tWorld, (size_t)15 * s->nt, "hipMalloc(triangles)"));
        TRY(dalloc(s->tAabb, (size_t)6 * s->nt, "hipMalloc(triangles)"));
        TRY(dalloc(s->tCounts, s->nt + 1, "hipMalloc(triangle grid)"));
        TRY(dalloc(s->tOffsets, s->nt + 1, "hipMalloc(triangle grid)"));
        if (s->nt > 0) {
            HIP_TRY(hipMemcpy(s->tBody, tb.data(), sizeof(int) * s->nt, hipMemcpyHostToDevice), "upload triangles");
            HIP_TRY(hipMemcpy(s->tLocal, B->tri_local, sizeof(double) * 15 * s->nt, hipMemcpyHostToDevice),
                    "upload triangles");
        }
        TRY(dalloc(s->tBits, 1, "hipMalloc"));
        HIP_TRY(hipHostMalloc((void**)&s->hTri, sizeof(int) * (s->nt + 1), hipHostMallocDefault), "hipHostMalloc");
        HIP_TRY(hipHostMalloc((void**)&s->hTriOff, sizeof(long long) * (s->nt + 1), hipHostMallocDefault),
                "hipHostMalloc");
        TRY(dalloc(s->d64, 1, "hipMalloc"));
        const int sixtyfour = 64;
        HIP_TRY(hipMemcpy(s->d64, &sixtyfour, sizeof(int), hipMemcpyHostToDevice), "upload");
        TRY(dalloc(s->regimeInstant, (size_t)7 * nb, "hipMalloc(regime counters)"));
        TRY(dalloc(s->regimeInterval, (size_t)7 * nb, "hipMalloc(regime counters)"));
        TRY(dalloc(s->microInterval, (size_t)22 * nb, "hipMalloc(microstate)"));
        HIP_TRY(hipMemset(s->regimeInterval, 0, sizeof(unsigned long long) * 7 * nb), "clear regime");
        HIP_TRY(hipMemset(s->microInterval, 0, sizeof(double) * 22 * nb), "clear microstate");
        // loads, Fz and microstate sums side by side (one reduction over 29 nb rows)
        TRY(dalloc(s->bodyResult, (size_t)29 * nb, "hipMalloc(body resultant)"));
        s->microInstant = s->bodyResult + 7 * nb;
        TRY(dalloc(s->O.bodyRows, (size_t)BODY_ROW * nb * nr, "hipMalloc(body load partials)"));
        TRY(dalloc(s->O.bodyTouched, (size_t)nb * nr, "hipMalloc(body load partials)"));
        TRY(dalloc(s->O.regime, (size_t)nb * nr, "hipMalloc(body regime state)"));
        TRY(dalloc(s->triWork, nr, "hipMalloc(triangle work list)"));
        TRY(dalloc(s->triHitMask, nr, "hipMalloc(triangle work list)"));
        TRY(dalloc(s->triHitN, nr, "hipMalloc(triangle work list)"));
        TRY(dalloc(s->triHitVal, (size_t)TRI_HITS * nr, "hipMalloc(triangle work list)"));
        s->redBlocks = (int)((nr + RED_TILE - 1) / RED_TILE);
        TRY(dalloc(s->redPartial, (size_t)29 * nb * s->redBlocks, "hipMalloc(reduction partials)"));
        if (s->keepBodyParticle) TRY(dalloc(s->O.bodyParticle, (size_t)nb * n * 6, "hipMalloc(body-particle forces)"));
        if (s->keepBodyDetail) TRY(dalloc(s->O.bodyDetail, (size_t)nb * n * 19, "hipMalloc(body contact details)"));
        HIP_TRY(hipHostMalloc((void**)&s->hPose, sizeof(double) * 13 * nb, hipHostMallocDefault), "hipHostMalloc");
        HIP_TRY(hipHostMalloc((void**)&s->hLoads, sizeof(double) * 7 * nb, hipHostMallocDefault), "hipHostMalloc");
        std::vector<double> mass(B->mass, B->mass + nb);
        TRY(dalloc(s->bodyMass, nb, "hipMalloc(body mass)"));
        HIP_TRY(hipMemcpy(s->bodyMass, mass.data(), sizeof(double) * nb, hipMemcpyHostToDevice), "upload");
        s->hBodyMass = mass;
        s->cRegInst.assign(7 * nb, 0); s->cRegInt.assign(7 * nb, 0);
        s->cMicInst.assign(22 * nb, 0.0); s->cMicInt.assign(22 * nb, 0.0);
        HIP_TRY(hipDeviceSynchronize(), "initialize");
        if (s->nt > 0) {
            TriSet T{};
            T.nt = s->nt; T.nb = nb; T.body = s->tBody; T.local = s->tLocal; T.world = s->tWorld;
            T.aabb = s->tAabb; T.pose = s->pose;
            hipLaunchKernelGGL(k_transform, dim3(blocks_for(s->nt)), dim3(BLOCK), 0, s->stream, T);
            TRY(launch_check("transform triangles"));
        }
    }
    HIP_TRY(hipDeviceSynchronize(), "initialize");
    const char* env = getenv("DEM_REORDER_INTERVAL");
    if (env) s->reorderInterval = std::max(0, atoi(env));
    if (clump) s->reorderInterval = 0;  // owners keep their rows (k_law_clump indexes them directly)
    if (const char* pb = getenv("DEM_PAIR_BLOCKS")) s->pairBlocks = std::max(1, atoi(pb));
    s->initialized = true;
    return DEM_OK;
}

}  // namespace

int dem_initialize(dem_solver* s, const dem_particles_view* P, const dem_components_view* C,
                   const dem_walls_view* W, const dem_bodies_view* B, const dem_mech* mpp,
                   const dem_mech* mpt, const dem_mech* mpw, const dem_control* ctl)
{
    if (!s || !P || !C || !valid_mech(mpp) || !valid_mech(mpt) || !valid_mech(mpw) || !ctl)
        return set_error(DEM_E_INVALID, "null argument");
    const int64_t rows[2] = {0, P->count};
    const int32_t walls[2] = {0, W ? W->count : 0};
    const SceneSpec spec{1, rows, walls, mpp, mpw};
    return initialize_impl(s, P, C, W, B, mpt, ctl, spec);
}

int dem_initialize_batch(dem_solver* s, const dem_particles_view* P, const dem_components_view* C,
                         const dem_walls_view* W, const dem_batch_view* batch, const dem_control* ctl)
{
    if (!s || !P || !C || !batch || !ctl || !batch->particle_start || !batch->wall_start || !batch->pp ||
        !batch->pw || batch->scene_count < 1)
        return set_error(DEM_E_INVALID, "null argument");
    const SceneSpec spec{batch->scene_count, batch->particle_start, batch->wall_start, batch->pp, batch->pw};
    return initialize_impl(s, P, C, W, nullptr, batch->pp, ctl, spec);
}

namespace {

int flush_finish(const dem_solver* cs);
int history_rows(const dem_solver* s, std::vector<int>& hrow);

TriSet triset(const dem_solver* s)
{
    TriSet T{};
    T.nt = s->nt; T.nb = s->nb; T.body = s->tBody; T.local = s->tLocal; T.world = s->tWorld;
    T.aabb = s->tAabb; T.pose = s->pose; T.entries = s->tEntries; T.tkey = s->tUse.keyA; T.ttri = s->tUse.triA;
    for (int a = 0; a < 3; ++a) {
        T.lo[a] = s->tLo[a];
        T.dims[a] = s->tDims[a];
    }
    T.start = s->tStartCap > 0 ? s->tUse.start : nullptr;
    return T;
}

// countTriangleCellsKernel for this step's triangles (box and total into the
// scalars).
int tri_grid_count(dem_solver* s)
{
    s->tEntries = 0;
    s->tStartCap = 0;
    hipLaunchKernelGGL(k_tri_count, dim3(blocks_for(s->nt)), dim3(BLOCK), 0, s->stream, triset(s), s->inv,
                       s->maxRadius, s->tCounts, s->sc);
    return launch_check("countTriangleCells");
}

// The solver's own triangle-grid buffers, grown to `total` entries and, for
// the dense-table path (cells > 0), a table of cells + 1 entries.
int tri_bufs_own(dem_solver* s, long long total, long long cells)
{
    TriBufs& b = s->tOwn;
    if (total > b.cap) {
        for (void* p : {(void*)b.keyA, (void*)b.keyB, (void*)b.triA, (void*)b.triB, (void*)b.counts})
            if (p) hipFree(p);
        b.keyA = b.keyB = nullptr;
        b.triA = b.triB = nullptr;
        b.counts = nullptr;
        b.cap = 0;
        const long long cap = total + total / 4 + 64;
        TRY(dalloc(b.keyA, cap, "hipMalloc(dynamic triangle grid keys)"));
        TRY(dalloc(b.keyB, cap, "hipMalloc(dynamic triangle grid keys)"));
        TRY(dalloc(b.triA, cap, "hipMalloc(dynamic triangle grid ids)"));
        TRY(dalloc(b.triB, cap, "hipMalloc(dynamic triangle grid ids)"));
        b.countBlocks = (int)((cap + RS_TILE - 1) / RS_TILE);
        TRY(dalloc(b.counts, (size_t)256 * b.countBlocks, "hipMalloc(triangle sort counts)"));
        b.cap = cap;
    }
    if (cells > 0 && cells + 1 > b.startCap) {
        for (void* p : {(void*)b.start, (void*)b.cur, (void*)b.scan})
            if (p) hipFree(p);
        b.start = b.cur = b.scan = nullptr;
        b.startCap = 0;
        const long long cap = cells + 1 + cells / 4;
        TRY(dalloc(b.start, (size_t)cap, "hipMalloc(triangle cell table)"));
        TRY(dalloc(b.cur, (size_t)cap, "hipMalloc(triangle cell table)"));
        TRY(dalloc(b.scan, (size_t)((cap + SCAN_TILE - 1) / SCAN_TILE), "hipMalloc(triangle cell table)"));
        b.startCap = cap;
    }
    s->tUse = b;
    s->tInScratch = false;
    return DEM_OK;
}

// The same buffers carved out of the device's scratch array (no allocation):
// the dense-table path needs the two triangle-id arrays, the table, its
// cursors (one slot more, for the count of large cells) and the scan
// partials; the sorted-pairs path the two key and two id arrays and the
// radix counts.  Steps whose grid does not fit fail with DEM_E_CAPACITY.
int tri_bufs_scratch(dem_solver* s, long long total, long long cells)
{
    void* base = nullptr;
    TRY(scratch_on_current_device(&base));
    auto al = [](unsigned long long bytes) { return (bytes + 255ull) & ~255ull; };
    const unsigned long long e = (unsigned long long)total;
    const long long blocks = (total + RS_TILE - 1) / RS_TILE;
    const long long tiles = (cells + 1 + SCAN_TILE - 1) / SCAN_TILE;
    const unsigned long long need =
        2 * al(4 * e) + (cells > 0 ? 2 * al(4ull * (unsigned long long)(cells + 2)) + al(4ull * (unsigned long long)tiles)
                                   : 2 * al(8 * e) + al(4ull * 256ull * (unsigned long long)blocks));
    if (need > (unsigned long long)DEM_NOALLOC_SCRATCH_BYTES)
        return set_error(DEM_E_CAPACITY,
                         "no-allocation scratch too small for this step's triangle grid: %llu entries over %lld "
                         "table cells need %llu bytes, the limit is %llu bytes (DEM_NOALLOC_SCRATCH_BYTES)",
                         e, cells, need, (unsigned long long)DEM_NOALLOC_SCRATCH_BYTES);
    unsigned char* p = (unsigned char*)base;
    auto take = [&](unsigned long long bytes) {
        void* q = p;
        p += al(bytes);
        return q;
    };
    TriBufs b{};
    b.triA = (int*)take(4 * e);
    b.triB = (int*)take(4 * e);
    b.cap = total;
    if (cells > 0) {
        b.start = (unsigned*)take(4ull * (unsigned long long)(cells + 2));
        b.cur = (unsigned*)take(4ull * (unsigned long long)(cells + 2));
        b.scan = (unsigned*)take(4ull * (unsigned long long)tiles);
        b.startCap = cells + 2;
    } else {
        b.keyA = (unsigned long long*)take(8 * e);
        b.keyB = (unsigned long long*)take(8 * e);
        b.counts = (unsigned*)take(4ull * 256ull * (unsigned long long)blocks);
        b.countBlocks = (int)blocks;
    }
    s->tUse = b;
    s->tInScratch = true;
    return DEM_OK;
}

// The triangle grid of this step.  counted: tri_grid_count ran and the host
// scalars already hold its box and total (read back with another value).
// noalloc: the per-step buffers come from the device's scratch array.
int build_triangle_grid(dem_solver* s, bool counted = false, bool noalloc = false)
{
    hipStream_t st = s->stream;
    if (!counted) {
        TRY(tri_grid_count(s));
        HIP_TRY(hipMemcpyAsync(s->hScalars, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st), "copy scalars");
        HIP_TRY(hipStreamSynchronize(st), "triangle grid counts");
    }
    TriSet T = triset(s);
    const unsigned long long total = s->hScalars->tri_total;
    if (s->hScalars->tri_overflow != 0 || total > (unsigned long long)INT_MAX)
        return set_error(DEM_E_CAPACITY, "triangle grid entry count exceeds the supported 32-bit index range");
    if (total == 0) return DEM_OK;
    // cell box of all entries: compact x-major keys of ceil(log2 cells) bits
    unsigned long long cells = 1;
    for (int a = 0; a < 3; ++a) {
        s->tLo[a] = s->hScalars->tri_lo[a];
        s->tDims[a] = (long long)s->hScalars->tri_hi[a] - s->tLo[a] + 1;
        cells *= (unsigned long long)s->tDims[a];  // each dim < 2^21: no overflow
    }
    constexpr unsigned long long kMaxTableCells = 1ull << 24;
    const long long tableCells = cells <= kMaxTableCells ? (long long)cells : 0;
    TRY(noalloc ? tri_bufs_scratch(s, (long long)total, tableCells) : tri_bufs_own(s, (long long)total, tableCells));
    s->tEntries = (int)total;
    if (tableCells > 0) {
        // counting sort over the box; its scan is the dense cell table
        const long tableN = (long)cells + 1;
        const int fillBlocks = (int)std::min<long>((tableN + BLOCK - 1) / BLOCK, 4096L);
        const unsigned triBlocks = (unsigned)((s->nt + BLOCK / 64 - 1) / (BLOCK / 64));
        T = triset(s);
        hipLaunchKernelGGL(k_fill_u32, dim3(fillBlocks), dim3(BLOCK), 0, st, s->tUse.cur, tableN);
        hipLaunchKernelGGL(k_tri_cells<true>, dim3(triBlocks), dim3(BLOCK), 0, st, T, s->inv, s->maxRadius,
                           (const int*)s->tCounts, s->tUse.cur, (int*)nullptr);
        const long tiles = (tableN + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(BLOCK), 0, st, (const unsigned*)s->tUse.cur, tableN,
                           s->tUse.start, s->tUse.scan);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BLOCK), 0, st, s->tUse.scan, tiles);
        hipLaunchKernelGGL(k_scan_add, dim3(fillBlocks), dim3(BLOCK), 0, st, s->tUse.start, tableN,
                           (const unsigned*)s->tUse.scan, s->tUse.cur);
        hipLaunchKernelGGL(k_tri_cells<false>, dim3(triBlocks), dim3(BLOCK), 0, st, T, s->inv, s->maxRadius,
                           (const int*)s->tCounts, s->tUse.cur, s->tUse.triA);
        // the cursors are spent: tCur holds the list of large cells, its
        // count sits in the last slot (cells <= 2^24 leaves it unused)
        int* bigCount = (int*)s->tUse.cur + s->tUse.startCap - 1;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (unsigned*)bigCount, 1L);
        hipLaunchKernelGGL(k_tri_cell_order, dim3((unsigned)((cells + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st,
                           (const unsigned*)s->tUse.start, (long long)cells, s->tUse.triA, bigCount, (int*)s->tUse.cur);
        hipLaunchKernelGGL(k_tri_cell_order_big, dim3(1024), dim3(64), 0, st, (const unsigned*)s->tUse.start,
                           (const int*)bigCount, (const int*)s->tUse.cur, s->tUse.triA, s->tUse.triB);
        TRY(launch_check("fillTriangleCells"));
        s->tStartCap = tableN;
        return DEM_OK;
    }
    // a box too large for a table: sorted (key, triangle) pairs, searched;
    // each triangle writes its entries from its offset in triangle order
    int* counts = s->hTri;
    HIP_TRY(hipMemcpyAsync(counts, s->tCounts, sizeof(int) * s->nt, hipMemcpyDeviceToHost, st),
            "copy triangle grid counts");
    HIP_TRY(hipStreamSynchronize(st), "triangle grid counts");
    long long* off = s->hTriOff;
    for (long long i = 0, acc = 0; i < s->nt; ++i) { off[i] = acc; acc += counts[i]; }
    int bits = 1;
    while (bits < 63 && (1ull << bits) < cells) ++bits;
    int* hbits = s->hTri + s->nt;
    *hbits = bits;
    HIP_TRY(hipMemcpyAsync(s->tBits, hbits, sizeof(int), hipMemcpyHostToDevice, st), "upload sort bits");
    HIP_TRY(hipMemcpyAsync(s->tOffsets, off, sizeof(long long) * s->nt, hipMemcpyHostToDevice, st),
            "upload triangle offsets");
    T = triset(s);
    hipLaunchKernelGGL(k_tri_fill, dim3(blocks_for(s->nt)), dim3(BLOCK), 0, st, T, s->inv, s->maxRadius,
                       s->tCounts, s->tOffsets, s->tUse.keyA, s->tUse.triA);
    TRY(launch_check("fillTriangleCells"));
    const int nblk = (int)((total + RS_TILE - 1) / RS_TILE);
    TRY(radix_sort(st, s->tUse.keyA, s->tUse.triA, s->tUse.keyB, s->tUse.triB, (long)total, s->tUse.counts, nblk, s->tBits,
                   s->rsDigits, (bits + 7) / 8));
    return DEM_OK;
}

}  // namespace

// Waits until k_vscan has written sequence number `seq` into the mirror (a
// few microseconds after k_cells ends); falls back to a stream
// synchronisation after a second.
static int await_mirror(dem_solver* s, int seq)
{
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(&s->hMirror[0], __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
            HIP_TRY(hipStreamSynchronize(s->stream), "read displacement flag");
            if (__atomic_load_n(&s->hMirror[0], __ATOMIC_ACQUIRE) != seq)
                return set_error(DEM_E_DEVICE, "displacement flag not received");
            break;
        }
    }
    return DEM_OK;
}

// The range count of k_cells stays set on the device until the host reads it
// here: every copy of the counters still in flight lands first (so none can
// bring the old count back), then the count is cleared on both sides.
static int range_error(dem_solver* s)
{
    HIP_TRY(hipStreamSynchronize(s->stream), "range check");
    HIP_TRY(hipMemsetAsync(&s->sc->grid_overflow, 0, sizeof(int), s->stream), "clear range count");
    HIP_TRY(hipStreamSynchronize(s->stream), "clear range count");
    __atomic_store_n(&s->hScalars->grid_overflow, 0, __ATOMIC_RELAXED);
    return set_error(DEM_E_CAPACITY, "component grid coordinate exceeds signed 21-bit hash range");
}

// Whether the last copy of the counters that reached the host carries a range
// count (a copy still in flight shows an older, equally valid count).
static bool range_seen(const dem_solver* s)
{
    return !s->capturing && __atomic_load_n(&s->hScalars->grid_overflow, __ATOMIC_RELAXED) != 0;
}

// dem_compute_forces; noalloc: the triangle grid's per-step buffers come from
// the device's scratch array instead of the solver's own (grown on demand).
static int compute_forces_impl(dem_solver* s, const double* pose13, double* loads6, int collect, bool noalloc)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    TRY(flush_finish(s));
    hipStream_t st = s->stream;
    const long n = s->n;
    const int nb = s->nb;
    // a range error of an earlier step whose counters reached the host after
    // that step returned aborts here (the reference throws at the step)
    if (range_seen(s)) return range_error(s);
    if (pose13 && nb > 0) {
        std::memcpy(s->hPose, pose13, sizeof(double) * 13 * nb);
        HIP_TRY(hipMemcpyAsync(s->pose, s->hPose, sizeof(double) * 13 * nb, hipMemcpyHostToDevice, st),
                "upload body poses");
        if (s->nt > 0) {
            hipLaunchKernelGGL(k_transform, dim3(blocks_for(s->nt)), dim3(BLOCK), 0, st, triset(s));
            TRY(launch_check("transform triangles"));
        }
    }
    // stage events only when asked for: each adds a gap between kernels
    const bool timed = s->timing && !s->capturing;
    if (timed) HIP_TRY(hipEventRecord(s->ev[0], st), "event");
    if (!s->scalarsReset || s->capturing)
        hipLaunchKernelGGL(k_step_init, dim3(1), dim3(BLOCK), 0, st, s->sc, s->sceneAcc, s->nscenes);
    s->scalarsReset = false;
    Soa S = s->soa();   // owners
    Soa R = s->rows();  // contact rows (the owners, or their components with clumps)
    const long nr = R.n;
    // candidate lists (single-sphere owners; body loads are reduced by input row,
    // so coupled steps need no per-step sort either)
    const bool verlet = s->verlet;
    bool shrunk = false;  // this step rebuilds the lists with a thinner skin
    if (verlet && s->vlistValid && s->skinAdapt && s->sinceRebuild >= s->nextSkinCheck && !s->capturing) {
        // the largest displacement since the rebuild (k_cells of the last
        // step): if the motion would let a much thinner skin last
        // kSkinSteps steps, rebuild with it now
        hipLaunchKernelGGL(k_disp_max, dim3(1), dim3(BLOCK), 0, st, (const double*)s->dispMax, blocks_for(nr),
                           s->dDisp);
        HIP_TRY(hipMemcpyAsync(s->hDisp, s->dDisp, sizeof(double), hipMemcpyDeviceToHost, st), "read displacement");
        HIP_TRY(hipStreamSynchronize(st), "read displacement");
        const double rate = std::sqrt(*s->hDisp) / (double)s->sinceRebuild;
        const double want = std::min(s->skinMax, std::max(s->skinMin, 2.0 * rate * (double)kSkinSteps));
        if (want <= 0.5 * s->skin) {
            s->skin = want;
            s->vlistValid = false;
            shrunk = true;
        }
        s->nextSkinCheck *= 2;
    }
    const bool lists = verlet && s->vlistValid;
    if (s->clump) hipLaunchKernelGGL(k_expand, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, S, R, s->cOwner);
    // a list step on current records gets k_posr's outputs from k_cells
    const bool fusedPosr = lists && s->recFresh;
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, R, s->dSp,
                       s->nscenes == 1 ? s->hsp[0].inv : 0.0, s->offs, s->sc,
                       s->boxPart, s->partScene, s->sceneAcc, lists ? s->refC : nullptr,
                       0.25 * s->skin * s->skin, fusedPosr ? s->posr : nullptr, s->cidS, s->meta,
                       lists ? s->dispMax : nullptr);
    bool rebuild = true;
    bool triCounted = false;
    bool scanned = false;  // k_cells wrote k_posr's outputs and the list scan is queued
    if (lists) {
        if (s->capturing) return set_error(DEM_E_INVALID, "candidate-list steps cannot be captured");
        if (s->nt > 0) {
            // the triangle grid's box and total come back with the displacement
            // flag; the particle scan (doing nothing if the flag asks for a
            // rebuild) is queued before they are read, so the device works
            // while the host sizes the triangle grid
            TRY(tri_grid_count(s));
            HIP_TRY(hipMemcpyAsync(s->hScalars, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st),
                    "read displacement flag");
            if (fusedPosr) {
                HIP_TRY(hipEventRecord(s->flagEv, st), "event");
                if (timed) HIP_TRY(hipEventRecord(s->ev[1], st), "event");
                hipLaunchKernelGGL(k_vscan, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, S, s->posr,
                                   s->meta, s->cidS, s->cand, s->ncand, s->sc, s->H, s->stepIndex, s->nbr,
                                   s->ncount, s->tape.ppOcc, collect || s->countAlways, s->listBlk,
                                   (const int*)&s->sc->need_rebuild, (int*)nullptr, 0);
                TRY(launch_check("contact scan"));
                HIP_TRY(hipEventSynchronize(s->flagEv), "read displacement flag");
            } else {
                HIP_TRY(hipStreamSynchronize(st), "read displacement flag");
            }
            rebuild = s->hScalars->need_rebuild != 0;
            scanned = fusedPosr && !rebuild;
            triCounted = true;
        } else if (fusedPosr) {
            // the step's scan is queued before the flag is read (doing nothing
            // if the flag asks for a rebuild), so the device does not wait for
            // the host's decision; its first block hands the flag and the range
            // count of k_cells to host memory under a sequence number
            const int* skip = &s->sc->need_rebuild;
            const int seq = ++s->mirrorSeq;
            if (timed) HIP_TRY(hipEventRecord(s->ev[1], st), "event");
            hipLaunchKernelGGL(k_vscan, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, S, s->posr, s->meta,
                               s->cidS, s->cand, s->ncand, s->sc, s->H, s->stepIndex, s->nbr, s->ncount,
                               s->tape.ppOcc, collect || s->countAlways, s->listBlk, skip, s->dMirror, seq);
            TRY(launch_check("contact scan"));
            TRY(await_mirror(s, seq));
            rebuild = s->hMirror[1] != 0;
            s->hScalars->grid_overflow = s->hMirror[2];
            scanned = !rebuild;
        } else {
            // the flag with the range count of k_cells
            HIP_TRY(hipMemcpyAsync(s->hScalars, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st),
                    "read displacement flag");
            HIP_TRY(hipStreamSynchronize(st), "read displacement flag");
            rebuild = s->hScalars->need_rebuild != 0;
        }
    }
    if (lists && !rebuild) ++s->sinceRebuild;
    bool triDone = false;
    if (rebuild) {
        hipLaunchKernelGGL(k_grid_setup, dim3(s->nscenes), dim3(BLOCK), 0, st, s->sc, s->dSp, s->boxPart,
                           s->partScene, s->sceneAcc, s->dGrid, s->nscenes, s->cellCap, s->gridDone);
        // with triangles the step reads counts back anyway: doing that now
        // also brings this step's key bits, so the sort launches only its passes
        int passes = 8;
        if (s->nt > 0) {
            if (triCounted) {
                HIP_TRY(hipMemcpyAsync(s->hScalars, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st),
                        "copy scalars");
                HIP_TRY(hipStreamSynchronize(st), "sort key bits");
            }
            TRY(build_triangle_grid(s, triCounted, noalloc));
            triDone = true;
            passes = (s->hScalars->nbits + 7) / 8;
        }
        hipLaunchKernelGGL(k_keys, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, R, s->dGrid, s->keyA, s->valA);
        TRY(launch_check("grid keys"));
        TRY(radix_sort(st, s->keyA, s->valA, s->keyB, s->valB, nr, s->rsCounts, s->rsBlocks, &s->sc->nbits,
                       s->rsDigits, passes));
        hipLaunchKernelGGL(k_fixup, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, s->keyA, s->valA, nr,
                           R.i + (long)I_ORIG * nr);
        // with candidate lists the rows are put in cell order at every rebuild
        // (unless they must stay in place)
        const bool reorder = !s->clump && ((verlet && !s->keepRows) ||
                                           (s->reorderInterval > 0 && (s->stepIndex % s->reorderInterval) == 0));
        if (reorder) {
            Soa D{s->dState[1 - s->cur], s->iState[1 - s->cur], n};
            if (verlet) {
                // rows in Morton order of their cells; valA becomes sorted
                // position -> new row
                int sbits = 0;
                while (sbits < 16 && (1 << sbits) < s->nscenes) ++sbits;
                hipLaunchKernelGGL(k_morton_keys, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->valA, s->dGrid,
                                   s->mKey, s->mVal, s->sc, 48 + sbits);
                TRY(radix_sort(st, s->mKey, s->mVal, s->keyB, s->valB, n, s->rsCounts, s->rsBlocks, &s->sc->mbits,
                               s->rsDigits, 8));
                hipLaunchKernelGGL(k_morton_perm, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->valA, s->mVal,
                                   s->valB, n);
                hipLaunchKernelGGL(k_morton_rank, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->mVal, s->valA, n);
                hipLaunchKernelGGL(k_permute, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, D, s->valB);
            } else {
                hipLaunchKernelGGL(k_permute, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, D, s->valA);
                hipLaunchKernelGGL(k_identity, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->valA, n);
            }
            s->cur = 1 - s->cur;
            S = s->soa();
            R = S;
            s->recFresh = false;  // the records are in the old row order
            if (s->H2.ppOcc) {
                hipLaunchKernelGGL(k_hist_move, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->H, s->H2);
                std::swap(s->H, s->H2);
            }
        }
        const int cellBlocks = (int)std::min<unsigned long long>((s->cellCap + BLOCK - 1) / BLOCK, 65535ull);
        hipLaunchKernelGGL(k_cellstart, dim3(cellBlocks), dim3(BLOCK), 0, st, s->keyA, nr, s->sc, s->cellStart,
                           s->cellCap);
    }
    ++s->nEvaluations;
    if (rebuild) ++s->nSorts;
    s->lastSorted = rebuild;
    // with candidate lists the packed arrays are indexed by row, else by
    // sorted position
    if (!scanned)
        hipLaunchKernelGGL(k_posr, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, R, verlet ? s->ident : s->valA, s->offs,
                           s->posr, s->cidS, verlet ? s->meta : nullptr, s->cOwner, s->ownS, s->cRank);
    if (!s->recFresh) hipLaunchKernelGGL(k_pack, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, R, s->offs, s->recs());
    TRY(launch_check("grid build"));
    bool useLists = verlet;
    if (verlet && rebuild) {
        if (s->skinAdapt) {
            // lists the motion outran in under a quarter of kSkinSteps: widen
            if (lists && !shrunk && s->sinceRebuild < kSkinSteps / 4) s->skin = std::min(s->skinMax, 2.0 * s->skin);
            s->sinceRebuild = 0;
            s->nextSkinCheck = 64;
        }
        hipLaunchKernelGGL(k_vbuild, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->posr, s->keyA, s->valA,
                           s->cellStart, s->dGrid, (const SceneParams*)s->dSp, s->sc, s->skin, s->cand, s->ncand,
                           s->refC);
        TRY(launch_check("candidate lists"));
        if (s->capturing) return set_error(DEM_E_INVALID, "candidate-list steps cannot be captured");
        HIP_TRY(hipMemcpyAsync(s->hFlag + 1, &s->sc->cand_overflow, sizeof(int), hipMemcpyDeviceToHost, st),
                "read candidate overflow");
        HIP_TRY(hipStreamSynchronize(st), "read candidate overflow");
        ++s->rebuilds;
        if (s->hFlag[1] != 0) {
            // more partners than a list holds: fall back to the full scan for good
            s->verlet = false;
            s->verletBroken = true;
            s->vlistValid = false;
            useLists = false;
            // the full scan reads the packed arrays by sorted position
            hipLaunchKernelGGL(k_posr, dim3(blocks_for(nr)), dim3(BLOCK), 0, st, R, s->valA, s->offs, s->posr,
                               s->cidS, (int4*)nullptr, s->cOwner, s->ownS, s->cRank);
        } else {
            s->vlistValid = true;
        }
    }
    if (s->nt > 0 && !triDone) TRY(build_triangle_grid(s, triCounted, noalloc));
    if (timed && !scanned) HIP_TRY(hipEventRecord(s->ev[1], st), "event");
    const int step = s->stepIndex++;
    const int nblk = blocks_for(nr);
    // list steps index rows directly (rows in Morton order, see k_morton_keys)
    const int* ord = useLists ? s->ident : s->valA;
    // the reference reads its per-step counters only on diagnostics steps
    const bool counters = collect || s->countAlways;
    if (scanned)
        ;
    else if (useLists)
        hipLaunchKernelGGL(k_vscan, dim3(nblk), dim3(BLOCK), 0, st, S, s->posr, s->meta, s->cidS,
                           s->cand, s->ncand, s->sc, s->H, step, s->nbr, s->ncount, s->tape.ppOcc,
                           counters, s->listBlk, (const int*)nullptr, (int*)nullptr, 0);
    else if (s->clump)
        hipLaunchKernelGGL(k_scan<true>, dim3(nblk), dim3(BLOCK), 0, st, R, s->valA, s->posr, s->cidS, s->keyA,
                           s->cellStart, s->dGrid, s->sc, s->H, step, s->nbr, s->ncount, s->tape.ppOcc,
                           counters, (const int*)s->ownS, s->listBlk);
    else
        hipLaunchKernelGGL(k_scan<false>, dim3(nblk), dim3(BLOCK), 0, st, R, s->valA, s->posr, s->cidS, s->keyA,
                           s->cellStart, s->dGrid, s->sc, s->H, step, s->nbr, s->ncount, s->tape.ppOcc,
                           counters, (const int*)nullptr, s->listBlk);
    hipLaunchKernelGGL(k_off_scan, dim3(1), dim3(BLOCK), 0, st, s->listBlk, nblk, s->sc);
    hipLaunchKernelGGL(k_off_apply, dim3(nblk), dim3(BLOCK), 0, st, s->ncount, nr, s->listBlk, s->nbr,
                       s->listOff, s->pairs);
    TRY(launch_check("contact scan"));
    if (timed) HIP_TRY(hipEventRecord(s->ev[2], st), "event");
    {
        auto* kp = useLists ? (s->nscenes == 1 ? k_pairs<true, true> : k_pairs<true, false>)
                            : (s->nscenes == 1 ? k_pairs<false, true> : k_pairs<false, false>);
        hipLaunchKernelGGL(kp, dim3(s->pairBlocks), dim3(BLOCK), 0, st, R, ord, s->recs(), s->pairs,
                           (const unsigned*)s->listOff, (const int*)s->ncount, s->H, s->dMatPP, s->hMatPP0,
                           step == 0,
                           s->offs, s->sc, s->pairF, s->rowF, s->tape.pair, s->tape.pairPrev);
    }
    TRY(launch_check("particle contact law"));
    if (timed) HIP_TRY(hipEventRecord(s->ev[3], st), "event");
    LawArgs A{};
    A.S = S; A.R = R; A.rank = s->cRank; A.ownerStart = s->ownerStart; A.posr = s->posr; A.rec = s->recs(); A.sorted = ord; A.ncount = s->ncount; A.off = s->listOff;
    A.pairF = s->pairF; A.rowF = s->rowF; A.H = s->H;
    A.W = Walls{s->nw, s->wallBuf, s->wallBuf + 3 * s->nw};
    A.T = triset(s); A.O = s->O; A.sc = s->sc;
    A.pt = s->pt; A.matPW = s->dMatPW; A.sp = s->dSp; A.first = step == 0;
    if (step == 0) A.pt.dt = 0.0;
    A.g[0] = s->g[0]; A.g[1] = s->g[1]; A.g[2] = s->g[2];
    A.gravity = s->gravity; A.universal = s->universal; A.byorig = s->byorig;
    A.step = step; A.offs = s->offs; A.nb = nb; A.tape = s->tape; A.counters = counters;
    A.identRows = ord == s->ident; A.oneScene = s->nscenes == 1;
    if (s->universal) hipLaunchKernelGGL(k_byorig, dim3(blocks_for(n)), dim3(BLOCK), 0, st, S, s->byorig);
    if (nb > 0) {
        const int fb = 1024;
        hipLaunchKernelGGL(k_fill_u32, dim3(fb), dim3(BLOCK), 0, st, (unsigned*)s->O.regime, (long)nb * nr);
        if (s->O.bodyParticle) hipLaunchKernelGGL(k_fill, dim3(fb), dim3(BLOCK), 0, st, s->O.bodyParticle, (long)nb * n * 6, 0.0);
        if (s->O.bodyDetail) hipLaunchKernelGGL(k_fill, dim3(fb), dim3(BLOCK), 0, st, s->O.bodyDetail, (long)nb * n * 19, 0.0);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(BLOCK), 0, st, s->regimeInstant, (long)7 * nb);
    }
    if (s->clump) {
        hipLaunchKernelGGL(k_law_clump, dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
    } else {
        if (nb > 0) {
            hipLaunchKernelGGL(k_tri_sift, dim3(blocks_for(n)), dim3(BLOCK), 0, st, A, s->triWork);
            hipLaunchKernelGGL(k_tri_area, dim3(TRI_AREA_BLOCKS), dim3(BLOCK), 0, st, A, (const int*)s->triWork,
                               s->triHitMask, s->triHitN, s->triHitVal);
            hipLaunchKernelGGL(k_law_pt, dim3((unsigned)((n + TRI_BLOCK - 1) / TRI_BLOCK)), dim3(TRI_BLOCK), 0, st, A,
                               (const int*)s->triWork, (const unsigned long long*)s->triHitMask,
                               (const int*)s->triHitN, (const double2*)s->triHitVal);
            TRY(launch_check("particle-triangle contact law"));
        }
        // a step of a multi-step run that another step follows finishes and
        // advances its rows in the law kernel (zero-offset single spheres,
        // no bodies, no diagnostics or tape, no universal gravitation)
        // (the differentiable forward run, which counts every evaluation and
        // records the wall tape, takes the counted variant: it keeps the
        // counters for the host and writes the tape)
        const bool finishHere = s->nextDt > 0.0 && !collect && nb == 0 && !s->offs && !s->universal &&
                                !s->tape.pair && !s->tape.pwOcc;
        const bool counted = s->countAlways || s->tape.wallCount || s->tape.plateAcc;
        if (s->universal) {
            hipLaunchKernelGGL((k_law<true, false>), dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
        } else if (finishHere) {
            A.finishDt = s->nextDt;
            A.nextDt = s->nextDt;
            A.sceneAcc = s->sceneAcc;
            A.nscenes = s->nscenes;
            if (counted)
                hipLaunchKernelGGL((k_law<false, true, false, true>), dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
            else if (s->nscenes == 1)
                hipLaunchKernelGGL((k_law<false, true, true>), dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
            else
                hipLaunchKernelGGL((k_law<false, true>), dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
            s->advancedAhead = true;
            s->scalarsReset = !counted;
            s->recFresh = true;
        } else {
            hipLaunchKernelGGL((k_law<false, false>), dim3(blocks_for(n)), dim3(BLOCK), 0, st, A);
        }
    }
    TRY(launch_check("contact law"));
    if (timed) HIP_TRY(hipEventRecord(s->ev[4], st), "event");
    const bool interval = A.pt.dt > 0.0;
    if (nb > 0) {
        // single-sphere owners: only the rows k_tri_sift listed; clumps: every row
        const int* work = s->clump ? nullptr : s->triWork;
        hipLaunchKernelGGL(k_regime, dim3(REGIME_BLOCKS, (unsigned)nb), dim3(BLOCK), 0, st, s->O.regime, ord, nr, work,
                           (const Scalars*)s->sc, s->regimeInstant, s->regimeInterval, interval ? 1 : 0);
        // body loads [6 nb], Fz [nb] and microstate [22 nb] sums in one
        // reduction; the microstate sums also feed the interval
        hipLaunchKernelGGL(k_reduce_body, dim3(s->redBlocks, nb), dim3(BLOCK), 0, st, (const double*)s->O.bodyRows,
                           (const unsigned char*)s->O.bodyTouched, nr, nb, s->redPartial, s->redBlocks);
        hipLaunchKernelGGL(k_reduce_final, dim3(29 * nb), dim3(64), 0, st, s->redPartial, 29 * nb,
                           s->redBlocks, s->bodyResult, interval ? s->microInterval : nullptr, 7 * nb);
        TRY(launch_check("body reductions"));
        ++s->intervalSteps;
        if (!interval) --s->intervalSteps;
    }
    if (timed) HIP_TRY(hipEventRecord(s->ev[5], st), "event");
    s->evReady = timed;
    // a list step without rebuild has brought the scalars back with its
    // displacement flag (the only later writer of the range flag is the rebuild)
    if (!scanned || collect)
        HIP_TRY(hipMemcpyAsync(s->hScalars, s->sc, sizeof(Scalars), hipMemcpyDeviceToHost, st), "copy scalars");
    const bool needLoads = loads6 && nb > 0;
    if (needLoads)
        HIP_TRY(hipMemcpyAsync(s->hLoads, s->bodyResult, sizeof(double) * 7 * nb, hipMemcpyDeviceToHost, st),
                "download body resultant forces");
    if (collect) {
        HIP_TRY(hipStreamSynchronize(st), "collect diagnostics");
        const Scalars& h = *s->hScalars;
        if (h.grid_overflow != 0) return range_error(s);
        dem_step_stats stt{};
        stt.particle_contacts = h.particle_contacts;
        stt.body_contacts = h.body_contacts;
        double md;
        std::memcpy(&md, &h.body_max_depth_bits, sizeof md);
        stt.body_max_depth = md;
        stt.particle_history_overflows = h.pp_overflow;
        stt.body_history_overflows = h.pw_overflow + h.pt_overflow;
        stt.particle_history_high_water = h.pp_high;
        stt.body_history_high_water = h.body_high;
        stt.triangle_grid_overflows = h.tri_overflow;
        if (nb > 0) {
            double* fz = s->hLoads + 6 * nb;  // pinned, sized at initialisation
            HIP_TRY(hipMemcpy(fz, s->bodyResult + 6 * nb, sizeof(double) * nb, hipMemcpyDeviceToHost), "download fz");
            double sum = 0.0;
            for (int b = 0; b < nb; ++b) sum += fz[b];
            stt.body_contact_fz = sum;
            HIP_TRY(hipMemcpy(s->cRegInst.data(), s->regimeInstant, sizeof(uint64_t) * 7 * nb, hipMemcpyDeviceToHost), "download regime");
            HIP_TRY(hipMemcpy(s->cRegInt.data(), s->regimeInterval, sizeof(uint64_t) * 7 * nb, hipMemcpyDeviceToHost), "download regime");
            HIP_TRY(hipMemcpy(s->cMicInst.data(), s->microInstant, sizeof(double) * 22 * nb, hipMemcpyDeviceToHost), "download microstate");
            HIP_TRY(hipMemcpy(s->cMicInt.data(), s->microInterval, sizeof(double) * 22 * nb, hipMemcpyDeviceToHost), "download microstate");
            hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(BLOCK), 0, st, s->regimeInterval, (long)7 * nb);
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(BLOCK), 0, st, s->microInterval, (long)22 * nb, 0.0);
            TRY(launch_check("reset body regime interval counters"));
            stt.body_regime_interval_native_steps = s->intervalSteps;
            s->intervalSteps = 0;
        }
        s->cached = stt;
        if (stt.particle_history_overflows != 0)
            return set_error(DEM_E_CAPACITY,
                             "CUDA particle contact-history capacity exceeded; increase kGpuParticleHistorySlots");
        if (stt.body_history_overflows != 0)
            return set_error(DEM_E_CAPACITY,
                             "CUDA wall/body contact-history capacity exceeded; increase the configured history slots");
        if (stt.triangle_grid_overflows != 0)
            return set_error(DEM_E_CAPACITY, "CUDA triangle-grid coordinate/count overflow");
    }
    if (needLoads) {
        HIP_TRY(hipStreamSynchronize(st), "download body resultant forces");
        if (range_seen(s)) return range_error(s);
        const std::vector<double>& mass = s->hBodyMass;
        for (int b = 0; b < nb; ++b) {
            for (int k = 0; k < 6; ++k) loads6[6 * b + k] = s->hLoads[6 * b + k];
            if (s->gravity) {
                loads6[6 * b] += s->g[0] * mass[b];
                loads6[6 * b + 1] += s->g[1] * mass[b];
                loads6[6 * b + 2] += s->g[2] * mass[b];
            }
        }
    }
    // list steps have the count of this step already (it came with the
    // displacement flag); other steps report it once their copy has landed
    if (range_seen(s)) return range_error(s);
    return DEM_OK;
}

int dem_compute_forces(dem_solver* s, const double* pose13, double* loads6, int collect)
{
    return compute_forces_impl(s, pose13, loads6, collect, false);
}

int dem_compute_forces_noalloc(dem_solver* s, const double* pose13, double* loads6, int collect)
{
    const int rc = compute_forces_impl(s, pose13, loads6, collect, true);
    // the step's kernels are done with the scratch array when the call returns
    if (rc == DEM_OK && s->tInScratch) HIP_TRY(hipStreamSynchronize(s->stream), "computeForces (no allocation)");
    return rc;
}

namespace {
// finishParticleStep is deferred so it can run fused with the next
// advanceParticles; any other call that reads or writes particle state
// completes it first.
// History-table row of every input row (the tables follow the particles
// when candidate-list rebuilds reorder them).
int history_rows(const dem_solver* s, std::vector<int>& hrow)
{
    if (s->clump) {
        // component rows never move: the row is the component index
        hrow.resize(s->nc);
        std::iota(hrow.begin(), hrow.end(), 0);
        return DEM_OK;
    }
    const long n = s->n;
    std::vector<int> orig(n), hist(n);
    HIP_TRY(hipMemcpy(orig.data(), s->iState[s->cur] + (size_t)I_ORIG * n, sizeof(int) * n, hipMemcpyDeviceToHost),
            "download particle rows");
    HIP_TRY(hipMemcpy(hist.data(), s->iState[s->cur] + (size_t)I_HIST * n, sizeof(int) * n, hipMemcpyDeviceToHost),
            "download history rows");
    hrow.assign(n, 0);
    for (long p = 0; p < n; ++p) hrow[orig[p]] = hist[p];
    return DEM_OK;
}

int flush_finish(const dem_solver* cs)
{
    dem_solver* s = const_cast<dem_solver*>(cs);
    if (!s->pendingFinish) return DEM_OK;
    s->pendingFinish = false;
    s->recFresh = false;
    hipLaunchKernelGGL(k_finish, dim3(blocks_for(s->n)), dim3(BLOCK), 0, s->stream, s->soa(), s->pendingDt);
    return launch_check("advanceHalfVelocity");
}
}  // namespace

int dem_init_half_step(dem_solver* s, double h)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    TRY(flush_finish(s));
    s->recFresh = false;
    hipLaunchKernelGGL(k_init_half, dim3(blocks_for(s->n)), dim3(BLOCK), 0, s->stream, s->soa(), h);
    return launch_check("initializeLeapfrog");
}

int dem_advance_particles(dem_solver* s, double dt, double* vmax)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (s->advancedAhead) {
        // the law kernel of the last evaluation advanced the rows already
        s->advancedAhead = false;
        if (vmax) *vmax = 0.0;
        return DEM_OK;
    }
    // zero-offset single-sphere rows: the advance also writes the contact records
    const Recs R = (!s->offs && !s->clump) ? s->recs() : Recs{nullptr, nullptr};
    if (s->pendingFinish) {
        s->pendingFinish = false;
        // the next force evaluation's counters are reset here (the host has
        // taken this step's already)
        hipLaunchKernelGGL(k_finish_advance, dim3(blocks_for(s->n)), dim3(BLOCK), 0, s->stream, s->soa(),
                           s->pendingDt, dt, R, s->sc, s->sceneAcc, s->nscenes);
        s->scalarsReset = true;
    } else {
        hipLaunchKernelGGL(k_advance, dim3(blocks_for(s->n)), dim3(BLOCK), 0, s->stream, s->soa(), dt, R);
    }
    s->recFresh = R.main != nullptr;
    if (vmax) *vmax = 0.0;
    return launch_check("advancePosition");
}

int dem_finish_particle_step(dem_solver* s, double dt, double)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (s->advancedAhead) return DEM_OK;  // finished by the law kernel
    TRY(flush_finish(s));
    s->pendingFinish = true;
    s->pendingDt = dt;
    return DEM_OK;
}

int dem_download_particle_state(const dem_solver* s, double* pos, double* vel, double* ang, double* quat)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    TRY(flush_finish(s));
    const long n = s->n;
    hipLaunchKernelGGL(k_gather_state, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->soa(), s->dl);
    TRY(launch_check("gather state"));
    std::vector<double> h((size_t)13 * n);
    HIP_TRY(hipMemcpyAsync(h.data(), s->dl, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
            "download particle state");
    HIP_TRY(hipStreamSynchronize(s->stream), "download particle state");
    for (long i = 0; i < n; ++i) {
        if (pos) for (int a = 0; a < 3; ++a) pos[3 * i + a] = h[(size_t)a * n + i];
        if (vel) for (int a = 0; a < 3; ++a) vel[3 * i + a] = h[(size_t)(3 + a) * n + i];
        if (ang) for (int a = 0; a < 3; ++a) ang[3 * i + a] = h[(size_t)(6 + a) * n + i];
        if (quat) for (int a = 0; a < 4; ++a) quat[4 * i + a] = h[(size_t)(9 + a) * n + i];
    }
    return DEM_OK;
}

int dem_download_force_output(const dem_solver* s, double* contact, double* field, double* wall,
                              double* total, double* bp, double* bd)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    const long n = s->n;
    hipStream_t st = s->stream;
    std::vector<double> c(6 * n), w(6 * n), b(6 * n), f(3 * n);
    struct Item { const double* src; int rows; double* dst; };
    Item items[] = {{s->O.contact, 6, c.data()}, {s->O.wall, 6, w.data()}, {s->O.body, 6, b.data()},
                    {s->O.field, 3, f.data()}};
    for (auto& it : items) {
        if (it.src == s->O.body && s->nb == 0) continue;  // no bodies: body force is zero
        hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(n)), dim3(BLOCK), 0, st, s->soa(), it.src, it.rows, s->dl);
        TRY(launch_check("gather forces"));
        HIP_TRY(hipMemcpyAsync(it.dst, s->dl, sizeof(double) * it.rows * n, hipMemcpyDeviceToHost, st),
                "download force output");
        HIP_TRY(hipStreamSynchronize(st), "download force output");
    }
    for (long i = 0; i < n; ++i) {
        double cf[6], wf[6], bf[6], ff[6];
        for (int k = 0; k < 6; ++k) { cf[k] = c[(size_t)k * n + i]; wf[k] = w[(size_t)k * n + i]; bf[k] = b[(size_t)k * n + i]; }
        for (int k = 0; k < 3; ++k) ff[k] = f[(size_t)k * n + i];
        ff[3] = ff[4] = ff[5] = 0.0;
        // total = (contact + field) + (wall + body), torque = contact + (wall + body) (contact:1215-1217)
        double tf[6];
        for (int k = 0; k < 3; ++k) tf[k] = (cf[k] + ff[k]) + (wf[k] + bf[k]);
        for (int k = 3; k < 6; ++k) tf[k] = cf[k] + (wf[k] + bf[k]);
        for (int k = 0; k < 6; ++k) {
            if (contact) contact[6 * i + k] = cf[k];
            if (field) field[6 * i + k] = 0.0 + ff[k];
            if (wall) wall[6 * i + k] = wf[k];
            if (total) total[6 * i + k] = tf[k];
        }
    }
    if (bp && s->nb > 0) {
        if (!s->O.bodyParticle) return set_error(DEM_E_INVALID, "body-particle force output was not kept");
        HIP_TRY(hipMemcpy(bp, s->O.bodyParticle, sizeof(double) * 6 * s->nb * n, hipMemcpyDeviceToHost),
                "download body-particle force output");
    }
    if (bd && s->nb > 0) {
        if (!s->O.bodyDetail) return set_error(DEM_E_INVALID, "body detail output was not kept");
        HIP_TRY(hipMemcpy(bd, s->O.bodyDetail, sizeof(double) * 19 * s->nb * n, hipMemcpyDeviceToHost),
                "download body detail output");
    }
    return DEM_OK;
}

int dem_get_step_stats(const dem_solver* s, dem_step_stats* stats, uint64_t* ri, uint64_t* rv, double* mi,
                       double* mv)
{
    if (!s) return set_error(DEM_E_INVALID, "null solver");
    if (stats) *stats = s->cached;
    const size_t nb = (size_t)s->nb;
    if (ri && nb) std::memcpy(ri, s->cRegInst.data(), sizeof(uint64_t) * 7 * nb);
    if (rv && nb) std::memcpy(rv, s->cRegInt.data(), sizeof(uint64_t) * 7 * nb);
    if (mi && nb) std::memcpy(mi, s->cMicInst.data(), sizeof(double) * 22 * nb);
    if (mv && nb) std::memcpy(mv, s->cMicInt.data(), sizeof(double) * 22 * nb);
    return DEM_OK;
}

int dem_particle_grid_entry_count(const dem_solver* s, size_t* out)
{
    if (!s || !out) return set_error(DEM_E_INVALID, "null argument");
    *out = (size_t)s->nc;
    return DEM_OK;
}

int dem_triangle_grid_entry_count(const dem_solver* s, size_t* out)
{
    if (!s || !out) return set_error(DEM_E_INVALID, "null argument");
    *out = (size_t)s->tEntries;
    return DEM_OK;
}

int dem_device_memory_bytes(const dem_solver* s, size_t* out)
{
    if (!s || !out) return set_error(DEM_E_INVALID, "null argument");
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot), "query HIP memory usage");
    *out = s->freeBefore > fr ? s->freeBefore - fr : 0;
    return DEM_OK;
}

int dem_debug_run_counters(const dem_solver* s, long long* out, int capacity, int* count)
{
    if (!s) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    unsigned long long passes[2] = {0ull, 0ull};
    if (s->initialized && capacity > 6) {
        HIP_TRY(hipMemcpyAsync(passes, &s->sc->scan_lds, sizeof passes, hipMemcpyDeviceToHost, s->stream),
                "read scan counters");
        HIP_TRY(hipStreamSynchronize(s->stream), "read scan counters");
    }
    const long long v[9] = {s->nEvaluations, s->nSorts, (long long)s->rebuilds, s->nGraphSteps,
                            s->nEagerSteps, s->lastSorted ? 1 : 0, (long long)passes[0], (long long)passes[1],
                            s->verlet ? std::llround(s->skin * 1e12) : 0LL};
    const int k = std::max(0, std::min(capacity, 9));
    for (int i = 0; i < k; ++i) out[i] = v[i];
    if (count) *count = k;
    return DEM_OK;
}

int dem_debug_grid(const dem_solver* s, uint64_t* keys, int32_t* ids)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    const long n = s->nc;
    unsigned long long* dk = nullptr;
    int* di = nullptr;
    TRY(dalloc(dk, n, "hipMalloc"));
    TRY(dalloc(di, n, "hipMalloc"));
    hipLaunchKernelGGL(k_debug_grid, dim3(blocks_for(n)), dim3(BLOCK), 0, s->stream, s->rows(), s->valA, dk, di);
    HIP_TRY(hipStreamSynchronize(s->stream), "debug grid");
    std::vector<uint64_t> hk(n);
    std::vector<int32_t> hi(n);
    HIP_TRY(hipMemcpy(hk.data(), dk, sizeof(uint64_t) * n, hipMemcpyDeviceToHost), "download");
    HIP_TRY(hipMemcpy(hi.data(), di, sizeof(int) * n, hipMemcpyDeviceToHost), "download");
    hipFree(dk);
    hipFree(di);
    // After an evaluation that sorted the grid the device order is returned
    // as it is (the radix sort and its tie fix-up are what is checked).  A
    // candidate-list step between rebuilds keeps the last rebuild's order
    // while the cells move, so its current cells are put in (cell key, input
    // row) order here.
    if (s->lastSorted) {
        if (keys) std::memcpy(keys, hk.data(), sizeof(uint64_t) * n);
        if (ids) std::memcpy(ids, hi.data(), sizeof(int32_t) * n);
        return DEM_OK;
    }
    std::vector<long> idx(n);
    std::iota(idx.begin(), idx.end(), 0L);
    std::sort(idx.begin(), idx.end(), [&](long a, long b) {
        return hk[a] != hk[b] ? hk[a] < hk[b] : hi[a] < hi[b];
    });
    for (long i = 0; i < n; ++i) {
        if (keys) keys[i] = hk[idx[i]];
        if (ids) ids[i] = hi[idx[i]];
    }
    return DEM_OK;
}

// The particle-pair value table is slot-major on the device ([slot][row][9]);
// the host tables are row-major ([row][slot][9]).
static void pp_to_row_major(std::vector<double>& v, size_t rows)
{
    std::vector<double> r(v.size());
    for (size_t b = 0; b < (size_t)PP_SLOTS; ++b)
        for (size_t h = 0; h < rows; ++h)
            std::memcpy(&r[(h * PP_SLOTS + b) * 9], &v[(b * rows + h) * 9], 9 * sizeof(double));
    v.swap(r);
}
static void pp_to_slot_major(std::vector<double>& v, size_t rows)
{
    std::vector<double> r(v.size());
    for (size_t h = 0; h < rows; ++h)
        for (size_t b = 0; b < (size_t)PP_SLOTS; ++b)
            std::memcpy(&r[(b * rows + h) * 9], &v[(h * PP_SLOTS + b) * 9], 9 * sizeof(double));
    v.swap(r);
}

int dem_debug_history(const dem_solver* s, uint64_t* pk, double* pv, uint64_t* wk, double* wv, uint64_t* bk,
                      double* bv)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    HIP_TRY(hipStreamSynchronize(s->stream), "debug history");
    const long n = s->nc;
    std::vector<int> hrow;
    TRY(history_rows(s, hrow));
    struct Tab { const void* occ; bool wide; const uint32_t* key; const double* val; int slots; uint64_t* ok; double* ov; };
    Tab tabs[] = {{s->H.ppOcc, true, s->H.ppKey, s->H.ppVal, PP_SLOTS, pk, pv},
                  {s->H.pwOcc, false, s->H.pwKey, s->H.pwVal, PW_SLOTS, wk, wv},
                  {s->H.ptOcc, true, s->H.ptKey, s->H.ptVal, PT_SLOTS, bk, bv}};
    for (auto& T : tabs) {
        if (!T.ok && !T.ov) continue;
        if (!T.occ) {
            if (T.ok) std::memset(T.ok, 0, sizeof(uint64_t) * T.slots * n);
            if (T.ov) std::memset(T.ov, 0, sizeof(double) * T.slots * 9 * n);
            continue;
        }
        std::vector<unsigned long long> occ(n);
        if (T.wide) {
            HIP_TRY(hipMemcpy(occ.data(), T.occ, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost), "download");
        } else {
            std::vector<unsigned> o32(n);
            HIP_TRY(hipMemcpy(o32.data(), T.occ, sizeof(unsigned) * n, hipMemcpyDeviceToHost), "download");
            for (long i = 0; i < n; ++i) occ[i] = o32[i];
        }
        std::vector<uint32_t> key((size_t)T.slots * n);
        HIP_TRY(hipMemcpy(key.data(), T.key, sizeof(uint32_t) * key.size(), hipMemcpyDeviceToHost), "download");
        std::vector<double> val;
        if (T.ov) {
            val.resize((size_t)T.slots * 9 * n);
            HIP_TRY(hipMemcpy(val.data(), T.val, sizeof(double) * val.size(), hipMemcpyDeviceToHost), "download");
            if (T.val == s->H.ppVal) pp_to_row_major(val, (size_t)n);
        }
        for (long i = 0; i < n; ++i) {
            const long h = hrow[i];
            for (int b = 0; b < T.slots; ++b) {
                const bool on = (occ[h] >> b) & 1ull;
                const size_t idx = (size_t)i * T.slots + b, src = (size_t)h * T.slots + b;
                if (T.ok) T.ok[idx] = on ? key[src] : 0;
                if (T.ov)
                    for (int q = 0; q < 9; ++q) T.ov[idx * 9 + q] = on ? val[src * 9 + q] : 0.0;
            }
        }
    }
    return DEM_OK;
}

int dem_debug_triangle_grid(const dem_solver* s, uint64_t* keys, int32_t* tris)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    HIP_TRY(hipStreamSynchronize(s->stream), "debug triangle grid");
    if (s->tEntries == 0) return DEM_OK;
    if (keys) {
        // compact box index -> the reference's packed 21-bit cell key
        if (s->tStartCap > 0) {
            // counting-sort layout: cell c holds entries [start[c], start[c + 1])
            std::vector<unsigned> start((size_t)s->tStartCap);
            HIP_TRY(hipMemcpy(start.data(), s->tUse.start, sizeof(unsigned) * start.size(), hipMemcpyDeviceToHost),
                    "download");
            for (size_t c = 0; c + 1 < start.size(); ++c)
                for (unsigned k = start[c]; k < start[c + 1]; ++k) keys[k] = c;
        } else {
            HIP_TRY(hipMemcpy(keys, s->tUse.keyA, sizeof(uint64_t) * s->tEntries, hipMemcpyDeviceToHost), "download");
        }
        const unsigned long long Y = (unsigned long long)s->tDims[1], Z = (unsigned long long)s->tDims[2];
        const unsigned long long mask = (1ull << 21) - 1ull;
        for (int k = 0; k < s->tEntries; ++k) {
            const unsigned long long c = keys[k];
            const long long x = (long long)(c / (Y * Z)) + s->tLo[0], y = (long long)((c / Z) % Y) + s->tLo[1],
                            z = (long long)(c % Z) + s->tLo[2];
            keys[k] = (((unsigned long long)(x + (1 << 20)) & mask) << 42) |
                      (((unsigned long long)(y + (1 << 20)) & mask) << 21) | ((unsigned long long)(z + (1 << 20)) & mask);
        }
    }
    if (tris) HIP_TRY(hipMemcpy(tris, s->tUse.triA, sizeof(int) * s->tEntries, hipMemcpyDeviceToHost), "download");
    return DEM_OK;
}

int dem_synchronize(const dem_solver* s)
{
    if (!s || !s->stream) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    TRY(flush_finish(s));
    HIP_TRY(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    return DEM_OK;
}

int dem_debug_enable_timings(dem_solver* s, int enabled)
{
    if (!s) return set_error(DEM_E_INVALID, "null argument");
    s->timing = enabled != 0;
    if (!s->timing) s->evReady = false;
    return DEM_OK;
}

int dem_debug_last_timings(const dem_solver* s, double* ms, int capacity, int* count)
{
    if (!s || !count) return set_error(DEM_E_INVALID, "null argument");
    *count = 0;
    if (!s->evReady) return DEM_OK;
    HIP_TRY(hipStreamSynchronize(s->stream), "timings");
    for (int i = 0; i + 1 < NEV && i < capacity; ++i) {
        float v = 0.0f;
        HIP_TRY(hipEventElapsedTime(&v, s->ev[i], s->ev[i + 1]), "hipEventElapsedTime");
        ms[i] = v;
        *count = i + 1;
    }
    return DEM_OK;
}

// CudaDemSolver::loadContactHistory / saveContactHistory
// (cuda_backend_cuda.cu:545-770): the DMHIST2 layout, byte for byte (native
// byte order, checked by the endian marker):
//   "DMHIST2\0", u32 version 2, u32 0x01020304, u32 owners, u32 components,
//   u32 slots[3] = {64, 16, 64}, u32 topology record size 40, u32 owner-state
//   record size 188; per component (i32 id, i32 owner id, f64 radius, body
//   x, y, z); per owner (i32 id, 23 f64: x y z vx vy vz qw qx qy qz wx wy wz
//   vhx vhy vhz whx why whz hdqw hdqx hdqy hdqz); the particle, wall and body
//   history keys (u64, component-major, 0 = empty) and their values (9 f64
//   per slot).  DMHIST1 ("DMHIST1\0", version 1, components, slots[3], keys,
//   values) is accepted for the legacy sphere topology and restores no
//   leapfrog state.  Keys and values of slots that the occupancy masks mark
//   free are written as 0, as the reference's expiry leaves them.
int dem_load_contact_history(dem_solver* s, const char* path)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "cannot load contact history before CUDA initialization");
    if (!path) return set_error(DEM_E_INVALID, "null argument");
    TRY(flush_finish(s));
    HIP_TRY(hipStreamSynchronize(s->stream), "load contact history");
    std::ifstream in(path, std::ios::binary);
    if (!in) return set_error(DEM_E_IO, "cannot open CUDA contact-history checkpoint: %s", path);
    auto rd = [&](void* p, size_t bytes) {
        in.read(reinterpret_cast<char*>(p), (std::streamsize)bytes);
        return (bool)in;
    };
    const long n = s->n;
    char magic[8];
    uint32_t version = 0, comps = 0, slots[3] = {0, 0, 0};
    if (!rd(magic, 8) || !rd(&version, 4))
        return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
    s->restored = false;
    bool v2 = false;
    std::vector<double> state;
    if (std::memcmp(magic, "DMHIST1", 8) == 0 && version == 1) {
        if (!rd(&comps, 4) || !rd(slots, 12))
            return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
        if (!s->legacySphere)
            return set_error(DEM_E_INVALID, "legacy DMHIST1 checkpoint is forbidden for clump topology: %s", path);
    } else if (std::memcmp(magic, "DMHIST2", 8) == 0 && version == 2) {
        uint32_t endian = 0, owners = 0, topo = 0, rec = 0;
        if (!rd(&endian, 4) || !rd(&owners, 4) || !rd(&comps, 4) || !rd(slots, 12) || !rd(&topo, 4) || !rd(&rec, 4))
            return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
        if (endian != 0x01020304u || owners != (uint32_t)n || topo != 40u || rec != 188u)
            return set_error(DEM_E_INVALID, "incompatible DMHIST2 header: %s", path);
        if (comps != (uint32_t)s->nc) return set_error(DEM_E_INVALID, "DMHIST2 component count mismatch: %s", path);
        for (uint32_t i = 0; i < comps; ++i) {
            int32_t ids[2];
            double geo[4];
            if (!rd(ids, 8) || !rd(geo, 32))
                return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
            if (ids[0] != s->compIds[i] || ids[1] != s->ownerIds[i] || geo[0] != s->compR[i] ||
                geo[1] != s->compB[3 * i] || geo[2] != s->compB[3 * i + 1] || geo[3] != s->compB[3 * i + 2])
                return set_error(DEM_E_INVALID, "DMHIST2 component topology mismatch: %s", path);
        }
        state.resize((size_t)23 * n);
        for (long i = 0; i < n; ++i) {
            int32_t id;
            if (!rd(&id, 4) || !rd(&state[(size_t)23 * i], 23 * 8))
                return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
            if (id != s->pids[i]) return set_error(DEM_E_INVALID, "DMHIST2 owner identity mismatch: %s", path);
        }
        v2 = true;
    } else {
        return set_error(DEM_E_INVALID, "unsupported contact-history checkpoint: %s", path);
    }
    if (comps != (uint32_t)s->nc || slots[0] != (uint32_t)PP_SLOTS || slots[1] != (uint32_t)PW_SLOTS ||
        slots[2] != (uint32_t)PT_SLOTS)
        return set_error(DEM_E_INVALID, "incompatible CUDA contact-history checkpoint: %s", path);
    const size_t nc = (size_t)s->nc;
    std::vector<uint64_t> pk(nc * PP_SLOTS), wk(nc * PW_SLOTS), bk(nc * PT_SLOTS);
    std::vector<double> pv(nc * PP_SLOTS * 9), wv(nc * PW_SLOTS * 9), bv(nc * PT_SLOTS * 9);
    if (!rd(pk.data(), 8 * pk.size()) || !rd(wk.data(), 8 * wk.size()) || !rd(bk.data(), 8 * bk.size()) ||
        !rd(pv.data(), 8 * pv.size()) || !rd(wv.data(), 8 * wv.size()) || !rd(bv.data(), 8 * bv.size()))
        return set_error(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path);
    char extra;
    if (in.read(&extra, 1)) return set_error(DEM_E_INVALID, "oversized CUDA contact-history checkpoint: %s", path);
    for (uint64_t k : pk)
        if (k > 0xFFFFFFFFull) return set_error(DEM_E_INVALID, "incompatible CUDA contact-history checkpoint: %s", path);
    for (uint64_t k : wk)
        if (k > 0xFFFFFFFFull) return set_error(DEM_E_INVALID, "incompatible CUDA contact-history checkpoint: %s", path);
    for (uint64_t k : bk)
        if (k > 0xFFFFFFFFull) return set_error(DEM_E_INVALID, "incompatible CUDA contact-history checkpoint: %s", path);
    hipStream_t st = s->stream;
    if (v2) {
        // owner state by input row -> current rows
        std::vector<double> hd((size_t)NDF * n);
        std::vector<int> orig(n);
        HIP_TRY(hipMemcpy(hd.data(), s->dState[s->cur], sizeof(double) * hd.size(), hipMemcpyDeviceToHost),
                "download particle state");
        HIP_TRY(hipMemcpy(orig.data(), s->iState[s->cur] + (size_t)I_ORIG * n, sizeof(int) * n,
                          hipMemcpyDeviceToHost), "download particle rows");
        const int fields[23] = {F_X,   F_Y,   F_Z,   F_VX,  F_VY,  F_VZ,  F_QW,  F_QX,
                                F_QY,  F_QZ,  F_WX,  F_WY,  F_WZ,  F_VHX, F_VHY, F_VHZ,
                                F_WHX, F_WHY, F_WHZ, F_HQW, F_HQX, F_HQY, F_HQZ};
        for (long p = 0; p < n; ++p) {
            const double* q = &state[(size_t)23 * orig[p]];
            for (int k = 0; k < 23; ++k) hd[(size_t)fields[k] * n + p] = q[k];
        }
        HIP_TRY(hipMemcpy(s->dState[s->cur], hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice),
                "upload particle state");
    }
    // keys -> occupancy masks; every restored slot counts as seen on the step
    // before the next one (the reference resets lastSeen to -1 and the step
    // counter to 0, so nothing expires on the next evaluation)
    std::vector<int> hrow;
    TRY(history_rows(s, hrow));
    // file rows are input rows; the tables are indexed by history row
    auto masks = [&](const std::vector<uint64_t>& keys, int slots_, std::vector<uint32_t>& k32,
                     std::vector<unsigned long long>& occ) {
        k32.resize(keys.size());
        occ.assign(nc, 0ull);
        for (size_t i = 0; i < nc; ++i) {
            const size_t h = (size_t)hrow[i];
            for (int b = 0; b < slots_; ++b) {
                const uint64_t k = keys[i * slots_ + b];
                k32[h * slots_ + b] = (uint32_t)k;
                if (k) occ[h] |= 1ull << b;
            }
        }
    };
    auto rows9 = [&](std::vector<double>& v, int slots_) {
        std::vector<double> r(v.size());
        for (size_t i = 0; i < nc; ++i)
            std::copy(v.begin() + i * slots_ * 9, v.begin() + (i + 1) * slots_ * 9,
                      r.begin() + (size_t)hrow[i] * slots_ * 9);
        v.swap(r);
    };
    rows9(pv, PP_SLOTS);
    rows9(wv, PW_SLOTS);
    rows9(bv, PT_SLOTS);
    std::vector<uint32_t> pk32, wk32, bk32;
    std::vector<unsigned long long> pocc, wocc, bocc;
    masks(pk, PP_SLOTS, pk32, pocc);
    masks(wk, PW_SLOTS, wk32, wocc);
    masks(bk, PT_SLOTS, bk32, bocc);
    std::vector<unsigned long long> pseen(2 * nc, 0ull);
    std::copy(pocc.begin(), pocc.end(), pseen.begin() + nc);
    std::vector<unsigned> wocc32(nc), wseen(2 * nc, 0u);
    for (size_t i = 0; i < nc; ++i) wocc32[i] = wseen[nc + i] = (unsigned)wocc[i];
    HIP_TRY(hipMemcpy(s->H.ppOcc, pocc.data(), 8 * nc, hipMemcpyHostToDevice), "upload particle history");
    HIP_TRY(hipMemcpy(s->H.ppSeen, pseen.data(), 16 * nc, hipMemcpyHostToDevice), "upload particle history");
    HIP_TRY(hipMemcpy(s->H.ppKey, pk32.data(), 4 * pk32.size(), hipMemcpyHostToDevice), "upload particle history keys");
    pp_to_slot_major(pv, nc);
    HIP_TRY(hipMemcpy(s->H.ppVal, pv.data(), 8 * pv.size(), hipMemcpyHostToDevice), "upload particle history values");
    HIP_TRY(hipMemcpy(s->H.pwOcc, wocc32.data(), 4 * nc, hipMemcpyHostToDevice), "upload wall history");
    HIP_TRY(hipMemcpy(s->H.pwSeen, wseen.data(), 8 * nc, hipMemcpyHostToDevice), "upload wall history");
    HIP_TRY(hipMemcpy(s->H.pwKey, wk32.data(), 4 * wk32.size(), hipMemcpyHostToDevice), "upload wall history keys");
    HIP_TRY(hipMemcpy(s->H.pwVal, wv.data(), 8 * wv.size(), hipMemcpyHostToDevice), "upload wall history values");
    if (s->H.ptOcc) {
        std::vector<unsigned long long> bseen(2 * nc, 0ull);
        std::copy(bocc.begin(), bocc.end(), bseen.begin() + nc);
        HIP_TRY(hipMemcpy(s->H.ptOcc, bocc.data(), 8 * nc, hipMemcpyHostToDevice), "upload body history");
        HIP_TRY(hipMemcpy(s->H.ptSeen, bseen.data(), 16 * nc, hipMemcpyHostToDevice), "upload body history");
        HIP_TRY(hipMemcpy(s->H.ptKey, bk32.data(), 4 * bk32.size(), hipMemcpyHostToDevice), "upload body history keys");
        HIP_TRY(hipMemcpy(s->H.ptVal, bv.data(), 8 * bv.size(), hipMemcpyHostToDevice), "upload body history values");
    }
    (void)st;
    s->stepIndex = 0;
    s->restored = v2;
    s->historyLoaded = true;
    s->recFresh = false;
    s->vlistValid = false;
    return DEM_OK;
}

int dem_save_contact_history(const dem_solver* cs, const char* path)
{
    if (!cs || !cs->initialized) return set_error(DEM_E_INVALID, "cannot save contact history before CUDA initialization");
    if (!path) return set_error(DEM_E_INVALID, "null argument");
    TRY(flush_finish(cs));
    HIP_TRY(hipStreamSynchronize(cs->stream), "save contact history");
    const long n = cs->n;
    const size_t nc = (size_t)cs->nc;
    std::vector<double> hd((size_t)NDF * n);
    std::vector<int> orig(n);
    HIP_TRY(hipMemcpy(hd.data(), cs->dState[cs->cur], sizeof(double) * hd.size(), hipMemcpyDeviceToHost),
            "download particle state");
    HIP_TRY(hipMemcpy(orig.data(), cs->iState[cs->cur] + (size_t)I_ORIG * n, sizeof(int) * n, hipMemcpyDeviceToHost),
            "download particle rows");
    std::vector<double> state((size_t)23 * n);
    const int fields[23] = {F_X,   F_Y,   F_Z,   F_VX,  F_VY,  F_VZ,  F_QW,  F_QX,  F_QY,  F_QZ,  F_WX, F_WY,
                            F_WZ,  F_VHX, F_VHY, F_VHZ, F_WHX, F_WHY, F_WHZ, F_HQW, F_HQX, F_HQY, F_HQZ};
    for (long p = 0; p < n; ++p)
        for (int k = 0; k < 23; ++k) state[(size_t)23 * orig[p] + k] = hd[(size_t)fields[k] * n + p];
    std::vector<int> hrow;
    TRY(history_rows(cs, hrow));
    // history tables with free slots cleared
    auto table = [&](const void* occDev, bool wide, const uint32_t* keyDev, const double* valDev, int slots_,
                     std::vector<uint64_t>& keys, std::vector<double>& vals) -> int {
        keys.assign(nc * slots_, 0ull);
        vals.assign(nc * slots_ * 9, 0.0);
        if (!occDev) return DEM_OK;
        std::vector<unsigned long long> occ(nc);
        if (wide) {
            HIP_TRY(hipMemcpy(occ.data(), occDev, 8 * nc, hipMemcpyDeviceToHost), "download history");
        } else {
            std::vector<unsigned> o32(nc);
            HIP_TRY(hipMemcpy(o32.data(), occDev, 4 * nc, hipMemcpyDeviceToHost), "download history");
            for (size_t i = 0; i < nc; ++i) occ[i] = o32[i];
        }
        std::vector<uint32_t> k32(nc * slots_);
        std::vector<double> v(nc * slots_ * 9);
        HIP_TRY(hipMemcpy(k32.data(), keyDev, 4 * k32.size(), hipMemcpyDeviceToHost), "download history keys");
        HIP_TRY(hipMemcpy(v.data(), valDev, 8 * v.size(), hipMemcpyDeviceToHost), "download history values");
        if (valDev == cs->H.ppVal) pp_to_row_major(v, nc);
        for (size_t i = 0; i < nc; ++i) {
            const size_t h = (size_t)hrow[i];
            for (int b = 0; b < slots_; ++b)
                if ((occ[h] >> b) & 1ull) {
                    const size_t idx = i * slots_ + b, src = h * slots_ + b;
                    keys[idx] = k32[src];
                    for (int q = 0; q < 9; ++q) vals[idx * 9 + q] = v[src * 9 + q];
                }
        }
        return DEM_OK;
    };
    std::vector<uint64_t> pk, wk, bk;
    std::vector<double> pv, wv, bv;
    TRY(table(cs->H.ppOcc, true, cs->H.ppKey, cs->H.ppVal, PP_SLOTS, pk, pv));
    TRY(table(cs->H.pwOcc, false, cs->H.pwKey, cs->H.pwVal, PW_SLOTS, wk, wv));
    TRY(table(cs->H.ptOcc, true, cs->H.ptKey, cs->H.ptVal, PT_SLOTS, bk, bv));
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return set_error(DEM_E_IO, "cannot create CUDA contact-history checkpoint: %s", path);
    auto wr = [&](const void* p, size_t bytes) { out.write(reinterpret_cast<const char*>(p), (std::streamsize)bytes); };
    const char magic[8] = {'D', 'M', 'H', 'I', 'S', 'T', '2', '\0'};
    const uint32_t head[9] = {2u, 0x01020304u, (uint32_t)n, (uint32_t)nc, (uint32_t)PP_SLOTS, (uint32_t)PW_SLOTS,
                              (uint32_t)PT_SLOTS, 40u, 188u};
    wr(magic, 8);
    wr(head, sizeof head);
    for (size_t i = 0; i < nc; ++i) {
        const int32_t ids[2] = {cs->compIds[i], cs->ownerIds[i]};
        const double geo[4] = {cs->compR[i], cs->compB[3 * i], cs->compB[3 * i + 1], cs->compB[3 * i + 2]};
        wr(ids, 8);
        wr(geo, 32);
    }
    for (long i = 0; i < n; ++i) {
        const int32_t id = cs->pids[i];
        wr(&id, 4);
        wr(&state[(size_t)23 * i], 23 * 8);
    }
    wr(pk.data(), 8 * pk.size());
    wr(wk.data(), 8 * wk.size());
    wr(bk.data(), 8 * bk.size());
    wr(pv.data(), 8 * pv.size());
    wr(wv.data(), 8 * wv.size());
    wr(bv.data(), 8 * bv.size());
    out.close();
    if (!out) return set_error(DEM_E_IO, "failed to write CUDA contact-history checkpoint: %s", path);
    return DEM_OK;
}

int dem_has_restored_leapfrog_state(const dem_solver* s, int* out)
{
    if (!s || !out) return set_error(DEM_E_INVALID, "null argument");
    *out = s->restored ? 1 : 0;
    return DEM_OK;
}

#include "dem_diff_checks.inl"

}  // extern "C"

#include "dem_diff.inl"
#include "dem_graph.inl"

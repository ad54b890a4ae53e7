/*
 * dem_oracle.c -- CPU restatement of the reference's CUDA DEM step.
 *
 * TEST INFRASTRUCTURE ONLY (see dem_oracle.h).  Reference paths below are
 * relative to the reference repository root; "contact" means
 * src/cuda/cuda_contact_backend.cu, "particle" src/cuda/cuda_particle_backend.cu,
 * "grid" src/cuda/cuda_grid_backend.cu, "body" src/cuda/cuda_body_backend.cu,
 * "backend" src/cuda_backend_cuda.cu.
 *
 * Deliberate choices (same in the HIP path):
 *   - Sums that the reference forms with device atomics (body loads, body
 *     contact counters, regime and microstate aggregates) are formed here per
 *     component and then over the components in input order with a fixed
 *     tree (tree_sum); the reference's own order is nondeterministic.
 *
 * Every function is pinned by a check independent of this restatement: the
 * reference's validation cases and known answers, its CPU binary at zero
 * velocity, or closed-form mechanics (tests/test_oracle.py,
 * test_transport_pins.py, test_microstate_pins.py, test_clumps.py).
 *   - thrust::sort_by_key on integer keys is a stable radix sort, so the
 *     particle grid order is (key, component index) and the triangle grid
 *     order is (key, fill position) = (key, triangle index).
 */
#include "dem_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PP_SLOTS DEM_PARTICLE_HISTORY_SLOTS
#define PW_SLOTS DEM_WALL_HISTORY_SLOTS
#define PT_SLOTS DEM_BODY_HISTORY_SLOTS
#define NREG DEM_BODY_REGIME_COUNTERS
#define NMICRO DEM_BODY_MICROSTATE_VALUES

static const double kPi = 3.1415926535897932384626433832795;
static const double kGravityConstant = 6.67430e-11;

static _Thread_local char g_error[512];

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

const char* ora_last_error(void) { return g_error; }

/* ---- vector helpers (contact:49-72) ------------------------------------ */
typedef struct { double x, y, z; } V3;
typedef struct { double x, y; } V2;

static V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 mul(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 cross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static double norm3(V3 a) { return sqrt(dot(a, a)); }
static V3 normalized(V3 a)
{
    const double length = norm3(a);
    return length > 1.0e-14 ? mul(a, 1.0 / length) : v3(0.0, 0.0, 0.0);
}

/* rotateIv: world = R(q) * body (contact:105-112, particle:103-114, body:25-49) */
static V3 rotate_iv(double qw, double qx, double qy, double qz, V3 v)
{
    return v3((1.0 - 2.0 * (qy * qy + qz * qz)) * v.x + 2.0 * (qx * qy - qw * qz) * v.y +
                  2.0 * (qx * qz + qw * qy) * v.z,
              2.0 * (qx * qy + qw * qz) * v.x + (1.0 - 2.0 * (qx * qx + qz * qz)) * v.y +
                  2.0 * (qy * qz - qw * qx) * v.z,
              2.0 * (qx * qz - qw * qy) * v.x + 2.0 * (qy * qz + qw * qx) * v.y +
                  (1.0 - 2.0 * (qx * qx + qy * qy)) * v.z);
}

/* packCell (grid:19-26, contact:134-141) */
uint64_t ora_pack_cell(int x, int y, int z)
{
    const int bias = 1 << 20;
    const uint64_t mask = (1ull << 21) - 1ull;
    return (((uint64_t)(int64_t)(x + bias) & mask) << 42) |
           (((uint64_t)(int64_t)(y + bias) & mask) << 21) |
           ((uint64_t)(int64_t)(z + bias) & mask);
}

/* lowerBound / upperBound (contact:144-176) */
static int lower_bound_u64(const uint64_t* keys, int count, uint64_t key)
{
    int first = 0, length = count;
    while (length > 0) {
        const int half = length >> 1, middle = first + half;
        if (keys[middle] < key) { first = middle + 1; length -= half + 1; }
        else length = half;
    }
    return first;
}
static int upper_bound_u64(const uint64_t* keys, int count, uint64_t key)
{
    int first = 0, length = count;
    while (length > 0) {
        const int half = length >> 1, middle = first + half;
        if (key < keys[middle]) length = half;
        else { first = middle + 1; length -= half + 1; }
    }
    return first;
}

/* ---- solver state ------------------------------------------------------- */
typedef struct {
    double epsS, mu, epsN, cohesion, beta, muT, muR, kN, dt, gx, gy, gz;
    int gravityEnabled, universalGravityEnabled;
} Mech; /* GpuMechanicalParams (contact.cuh:24-39) */

typedef struct { uint64_t key; int32_t id; } KeyId;

struct ora_solver {
    int initialized;
    /* owners (particle.cuh:5-54) */
    int n;
    int32_t *pid, *status;
    double *mass, *radius, *inx, *iny, *inz;
    double *x, *y, *z, *vx, *vy, *vz, *wx, *wy, *wz, *qw, *qx, *qy, *qz;
    double *fx, *fy, *fz, *tx, *ty, *tz, *ax, *ay, *az, *awx, *awy, *awz;
    double *vhx, *vhy, *vhz, *whx, *why, *whz, *hdqw, *hdqx, *hdqy, *hdqz;
    /* components sorted by (owner, id) (particle.cuh:56-78, backend:121-190) */
    int nc;
    int32_t *cid, *cowner, *ownerStart;
    double *cr, *cbx, *cby, *cbz, *cox, *coy, *coz, *cx, *cy, *cz;
    double *cvx, *cvy, *cvz, *cvhx, *cvhy, *cvhz;
    double maxRadius, meshSize, invMesh;
    /* walls */
    int nw;
    double *wox, *woy, *woz, *wnx, *wny, *wnz;
    /* bodies and triangles (body.cuh:5-64) */
    int nb, nt;
    double* bmass;
    double* bpose; /* [nb][13] */
    int32_t* tbody;
    double* tlocal; /* [nt][15] */
    double* tworld; /* [nt][15] v1..v4, n */
    double* taabb;  /* [nt][6] min xyz, max xyz */
    /* particle grid */
    uint64_t* gkey;
    int32_t* gid;
    /* triangle grid (contact.cuh:80-91) */
    int tentries;
    uint64_t* tkey;
    int32_t* ttri;
    /* history (contact.cuh:41-52) */
    uint64_t *hpk, *hwk, *hbk;
    int32_t *hpl, *hwl, *hbl;
    double *hpv, *hwv, *hbv;
    /* forces (contact.cuh:54-78) */
    int keepBodyParticle, keepBodyDetail;
    double *fcontact, *ffield, *fwall, *ftotal, *fbp, *fbd, *fbr;
    double *bodyRow, *fzRow; /* [nb][6][nc], [nb][nc]: per-component body load and Fz */
    int pcount, bcount, pover, bover, phigh, bhigh;
    double bmaxdepth, bfz;
    int32_t* regimeState;
    uint64_t *regimeInstant, *regimeInterval;
    double *microState, *microInstant, *microInterval;
    Mech pp, pt, pw;
    int stepIndex;
    int restored;     /* restoredParticleLeapfrogState (backend:658-770) */
    int legacySphere; /* one zero-offset component per owner (backend:172-177) */
    uint64_t intervalSteps;
    dem_step_stats cached;
    uint64_t *cachedRegimeInstant, *cachedRegimeInterval;
    double *cachedMicroInstant, *cachedMicroInterval;
    int threads;      /* owner-loop threads (ora_set_threads) */
};

#define ALLOC(p, count) do { const size_t n_ = (size_t)(count); (p) = calloc(n_ > 0 ? n_ : 1, sizeof(*(p))); } while (0)

int ora_create(ora_solver** out)
{
    if (!out) return fail(DEM_E_INVALID, "null output pointer");
    *out = calloc(1, sizeof(ora_solver));
    if (!*out) return fail(DEM_E_INVALID, "out of memory");
#ifdef _OPENMP
    (*out)->threads = omp_get_max_threads();
#else
    (*out)->threads = 1;
#endif
    return DEM_OK;
}

int ora_set_threads(ora_solver* s, int threads)
{
    if (!s || threads < 1) return fail(DEM_E_INVALID, "thread count must be positive");
#ifdef _OPENMP
    s->threads = threads;
#else
    s->threads = 1;
#endif
    return DEM_OK;
}

void ora_destroy(ora_solver* s)
{
    if (!s) return;
    void* ptrs[] = {s->pid, s->status, s->mass, s->radius, s->inx, s->iny, s->inz,
                    s->x, s->y, s->z, s->vx, s->vy, s->vz, s->wx, s->wy, s->wz,
                    s->qw, s->qx, s->qy, s->qz, s->fx, s->fy, s->fz, s->tx, s->ty, s->tz,
                    s->ax, s->ay, s->az, s->awx, s->awy, s->awz, s->vhx, s->vhy, s->vhz,
                    s->whx, s->why, s->whz, s->hdqw, s->hdqx, s->hdqy, s->hdqz,
                    s->cid, s->cowner, s->ownerStart, s->cr, s->cbx, s->cby, s->cbz,
                    s->cox, s->coy, s->coz, s->cx, s->cy, s->cz, s->cvx, s->cvy, s->cvz,
                    s->cvhx, s->cvhy, s->cvhz, s->wox, s->woy, s->woz, s->wnx, s->wny, s->wnz,
                    s->bmass, s->bpose, s->tbody, s->tlocal, s->tworld, s->taabb,
                    s->gkey, s->gid, s->tkey, s->ttri, s->hpk, s->hwk, s->hbk,
                    s->hpl, s->hwl, s->hbl, s->hpv, s->hwv, s->hbv,
                    s->fcontact, s->ffield, s->fwall, s->ftotal, s->fbp, s->fbd, s->fbr,
                    s->bodyRow, s->fzRow,
                    s->regimeState, s->regimeInstant, s->regimeInterval, s->microState,
                    s->microInstant, s->microInterval, s->cachedRegimeInstant,
                    s->cachedRegimeInterval, s->cachedMicroInstant, s->cachedMicroInterval};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
    free(s);
}

static Mech make_mech(const dem_mech* m, const dem_control* c)
{
    Mech r;
    memset(&r, 0, sizeof r);
    r.epsS = m->epsS; r.mu = m->mu; r.epsN = m->epsN; r.cohesion = m->cohesion;
    r.beta = m->beta; r.muT = m->muT; r.muR = m->muR; r.kN = m->kN;
    r.dt = c->dt;
    return r;
}

typedef struct { int owner, id, input; } CompOrder;
static int cmp_comp(const void* a, const void* b)
{
    const CompOrder* p = a; const CompOrder* q = b;
    if (p->owner != q->owner) return p->owner < q->owner ? -1 : 1;
    if (p->id != q->id) return p->id < q->id ? -1 : 1;
    return p->input < q->input ? -1 : (p->input > q->input);
}
static int cmp_int(const void* a, const void* b)
{
    const int p = *(const int*)a, q = *(const int*)b;
    return p < q ? -1 : (p > q);
}

static void transform_triangles(ora_solver* s);
static void update_components(ora_solver* s);

/* CudaDemSolver::initialize (backend:302-381), uploadComponents (backend:121-190) */
int ora_initialize(ora_solver* s, const dem_particles_view* P,
                   const dem_components_view* C, const dem_walls_view* W,
                   const dem_bodies_view* B, const dem_mech* mpp,
                   const dem_mech* mpt, const dem_mech* mpw, const dem_control* ctl)
{
    if (!s || !P || !C || !mpp || !mpt || !mpw || !ctl)
        return fail(DEM_E_INVALID, "null argument");
    if (s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver already initialized");
    if (P->count <= 0) return fail(DEM_E_INVALID, "CUDA DEM solver requires particles");
    const int n = P->count;
    { /* validateStableParticleIds (backend:67-78) */
        int* ids = malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) ids[i] = P->id[i];
        qsort(ids, (size_t)n, sizeof(int), cmp_int);
        for (int i = 1; i < n; ++i) {
            if (ids[i] == ids[i - 1]) {
                const int dup = ids[i];
                free(ids);
                return fail(DEM_E_INVALID,
                            "duplicate particle ID %d; CUDA contact history requires stable unique particle IDs",
                            dup);
            }
        }
        free(ids);
    }
    if (C->count <= 0) return fail(DEM_E_INVALID, "CUDA clump topology requires owners and components");
    s->n = n;
    ALLOC(s->pid, n); ALLOC(s->status, n); ALLOC(s->mass, n); ALLOC(s->radius, n);
    ALLOC(s->inx, n); ALLOC(s->iny, n); ALLOC(s->inz, n);
    double** fields[] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->wx, &s->wy, &s->wz,
                         &s->qw, &s->qx, &s->qy, &s->qz, &s->fx, &s->fy, &s->fz, &s->tx, &s->ty,
                         &s->tz, &s->ax, &s->ay, &s->az, &s->awx, &s->awy, &s->awz, &s->vhx,
                         &s->vhy, &s->vhz, &s->whx, &s->why, &s->whz, &s->hdqw, &s->hdqx,
                         &s->hdqy, &s->hdqz};
    for (size_t k = 0; k < sizeof fields / sizeof fields[0]; ++k) ALLOC(*fields[k], n);
    for (int i = 0; i < n; ++i) {
        s->pid[i] = P->id[i]; s->status[i] = P->status[i];
        s->mass[i] = P->mass[i]; s->radius[i] = P->radius[i];
        s->inx[i] = P->inertia[3 * i]; s->iny[i] = P->inertia[3 * i + 1];
        s->inz[i] = P->inertia[3 * i + 2];
        s->x[i] = P->pos[3 * i]; s->y[i] = P->pos[3 * i + 1]; s->z[i] = P->pos[3 * i + 2];
        s->vx[i] = P->vel[3 * i]; s->vy[i] = P->vel[3 * i + 1]; s->vz[i] = P->vel[3 * i + 2];
        s->wx[i] = P->ang_vel[3 * i]; s->wy[i] = P->ang_vel[3 * i + 1];
        s->wz[i] = P->ang_vel[3 * i + 2];
        s->qw[i] = P->quat[4 * i]; s->qx[i] = P->quat[4 * i + 1];
        s->qy[i] = P->quat[4 * i + 2]; s->qz[i] = P->quat[4 * i + 3];
    }
    /* components */
    const int nc = C->count;
    CompOrder* order = malloc(sizeof(CompOrder) * (size_t)nc);
    for (int i = 0; i < nc; ++i) {
        order[i].owner = C->owner[i]; order[i].id = C->id[i]; order[i].input = i;
    }
    /* std::sort with a (owner, id) comparator; ties cannot matter for
       valid input because (owner, id) pairs are unique per owner */
    qsort(order, (size_t)nc, sizeof(CompOrder), cmp_comp);
    s->nc = nc;
    ALLOC(s->cid, nc); ALLOC(s->cowner, nc); ALLOC(s->ownerStart, n + 1);
    double** cf[] = {&s->cr, &s->cbx, &s->cby, &s->cbz, &s->cox, &s->coy, &s->coz, &s->cx,
                     &s->cy, &s->cz, &s->cvx, &s->cvy, &s->cvz, &s->cvhx, &s->cvhy, &s->cvhz};
    for (size_t k = 0; k < sizeof cf / sizeof cf[0]; ++k) ALLOC(*cf[k], nc);
    double maxr = 0.0;
    for (int o = 0; o < nc; ++o) {
        const int in = order[o].input;
        const int ownerRow = C->owner[in];
        if (ownerRow < 0 || ownerRow >= n || C->id[in] < 0 || !isfinite(C->radius[in]) ||
            C->radius[in] <= 0.0) {
            free(order);
            return fail(DEM_E_INVALID, "invalid native clump component topology");
        }
        s->cid[o] = C->id[in]; s->cowner[o] = ownerRow; s->cr[o] = C->radius[in];
        s->cbx[o] = C->pos_body[3 * in]; s->cby[o] = C->pos_body[3 * in + 1];
        s->cbz[o] = C->pos_body[3 * in + 2];
        ++s->ownerStart[ownerRow + 1];
        if (s->cr[o] > maxr) maxr = s->cr[o];
    }
    free(order);
    for (int i = 1; i <= n; ++i) {
        s->ownerStart[i] += s->ownerStart[i - 1];
        const int count = s->ownerStart[i] - s->ownerStart[i - 1];
        if (count < 1 || count > 8) return fail(DEM_E_INVALID, "each CUDA owner requires 1 to 8 components");
    }
    s->maxRadius = maxr;
    s->meshSize = (ctl->mesh_scale > 0.0 ? ctl->mesh_scale : 2.5) * maxr;
    /* legacy sphere topology (backend:172-177) */
    s->legacySphere = nc == n;
    for (int i = 0; i < nc && s->legacySphere; ++i)
        s->legacySphere = s->cowner[i] == i && s->cid[i] == P->id[i] && s->cr[i] == P->radius[i] &&
                          s->cbx[i] == 0.0 && s->cby[i] == 0.0 && s->cbz[i] == 0.0;
    /* walls */
    s->nw = W ? W->count : 0;
    ALLOC(s->wox, s->nw); ALLOC(s->woy, s->nw); ALLOC(s->woz, s->nw);
    ALLOC(s->wnx, s->nw); ALLOC(s->wny, s->nw); ALLOC(s->wnz, s->nw);
    for (int w = 0; w < s->nw; ++w) {
        s->wox[w] = W->origin[3 * w]; s->woy[w] = W->origin[3 * w + 1]; s->woz[w] = W->origin[3 * w + 2];
        s->wnx[w] = W->normal[3 * w]; s->wny[w] = W->normal[3 * w + 1]; s->wnz[w] = W->normal[3 * w + 2];
    }
    /* bodies */
    s->nb = B ? B->count : 0;
    s->nt = 0;
    for (int b = 0; b < s->nb; ++b) {
        if (B->tri_count[b] < 0) return fail(DEM_E_INVALID, "body triangle count cannot be negative");
        if ((long long)s->nt + B->tri_count[b] > INT_MAX)
            return fail(DEM_E_CAPACITY, "total triangle count exceeds CUDA 32-bit index capacity");
        s->nt += B->tri_count[b];
    }
    ALLOC(s->bmass, s->nb); ALLOC(s->bpose, 13 * s->nb);
    ALLOC(s->tbody, s->nt); ALLOC(s->tlocal, 15 * (size_t)s->nt);
    ALLOC(s->tworld, 15 * (size_t)s->nt); ALLOC(s->taabb, 6 * (size_t)s->nt);
    {
        int t = 0;
        for (int b = 0; b < s->nb; ++b) {
            s->bmass[b] = B->mass[b];
            memcpy(s->bpose + 13 * b, B->pose13 + 13 * b, 13 * sizeof(double));
            for (int k = 0; k < B->tri_count[b]; ++k, ++t) {
                s->tbody[t] = b;
                memcpy(s->tlocal + 15 * (size_t)t, B->tri_local + 15 * (size_t)t, 15 * sizeof(double));
            }
        }
    }
    /* history (contact:1400-1446): keys 0, lastSeen -1, values 0 */
    ALLOC(s->hpk, (size_t)nc * PP_SLOTS); ALLOC(s->hwk, (size_t)nc * PW_SLOTS);
    ALLOC(s->hbk, (size_t)nc * PT_SLOTS);
    ALLOC(s->hpl, (size_t)nc * PP_SLOTS); ALLOC(s->hwl, (size_t)nc * PW_SLOTS);
    ALLOC(s->hbl, (size_t)nc * PT_SLOTS);
    for (size_t i = 0; i < (size_t)nc * PP_SLOTS; ++i) s->hpl[i] = -1;
    for (size_t i = 0; i < (size_t)nc * PW_SLOTS; ++i) s->hwl[i] = -1;
    for (size_t i = 0; i < (size_t)nc * PT_SLOTS; ++i) s->hbl[i] = -1;
    ALLOC(s->hpv, (size_t)nc * PP_SLOTS * 9); ALLOC(s->hwv, (size_t)nc * PW_SLOTS * 9);
    ALLOC(s->hbv, (size_t)nc * PT_SLOTS * 9);
    /* forces (contact:1456-1555) */
    s->keepBodyParticle = ctl->keep_body_particle;
    s->keepBodyDetail = ctl->keep_body_detail;
    ALLOC(s->fcontact, 6 * n); ALLOC(s->ffield, 6 * n); ALLOC(s->fwall, 6 * n); ALLOC(s->ftotal, 6 * n);
    if (s->keepBodyParticle) ALLOC(s->fbp, (size_t)s->nb * n * 6);
    if (s->keepBodyDetail) ALLOC(s->fbd, (size_t)s->nb * n * 19);
    ALLOC(s->fbr, 6 * s->nb);
    ALLOC(s->bodyRow, (size_t)6 * s->nb * nc);
    ALLOC(s->fzRow, (size_t)s->nb * nc);
    ALLOC(s->regimeState, (size_t)s->nb * nc);
    ALLOC(s->regimeInstant, NREG * s->nb); ALLOC(s->regimeInterval, NREG * s->nb);
    ALLOC(s->microState, (size_t)s->nb * nc * NMICRO);
    ALLOC(s->microInstant, NMICRO * s->nb); ALLOC(s->microInterval, NMICRO * s->nb);
    ALLOC(s->cachedRegimeInstant, NREG * s->nb); ALLOC(s->cachedRegimeInterval, NREG * s->nb);
    ALLOC(s->cachedMicroInstant, NMICRO * s->nb); ALLOC(s->cachedMicroInterval, NMICRO * s->nb);
    ALLOC(s->gkey, nc); ALLOC(s->gid, nc);
    /* mechanical blocks (backend:344-376) */
    s->pp = make_mech(mpp, ctl);
    s->pp.gx = ctl->gravity[0]; s->pp.gy = ctl->gravity[1]; s->pp.gz = ctl->gravity[2];
    s->pp.gravityEnabled = ctl->gravity_enabled ? 1 : 0;
    s->pp.universalGravityEnabled = ctl->universal_gravity ? 1 : 0;
    s->pt = make_mech(mpt, ctl);
    s->pw = make_mech(mpw, ctl);
    update_components(s);
    if (s->nb > 0) transform_triangles(s);
    s->initialized = 1;
    return DEM_OK;
}

/* updateComponentsKernel (particle:116-147) */
static void update_components(ora_solver* s)
{
    for (int c = 0; c < s->nc; ++c) {
        const int o = s->cowner[c];
        const V3 off = rotate_iv(s->qw[o], s->qx[o], s->qy[o], s->qz[o], v3(s->cbx[c], s->cby[c], s->cbz[c]));
        s->cox[c] = off.x; s->coy[c] = off.y; s->coz[c] = off.z;
        s->cx[c] = s->x[o] + off.x; s->cy[c] = s->y[o] + off.y; s->cz[c] = s->z[o] + off.z;
        const V3 ow = rotate_iv(s->qw[o], s->qx[o], s->qy[o], s->qz[o], v3(s->wx[o], s->wy[o], s->wz[o]));
        s->cvx[c] = s->vx[o] + ow.y * off.z - ow.z * off.y;
        s->cvy[c] = s->vy[o] + ow.z * off.x - ow.x * off.z;
        s->cvz[c] = s->vz[o] + ow.x * off.y - ow.y * off.x;
        const V3 hw = rotate_iv(s->qw[o], s->qx[o], s->qy[o], s->qz[o], v3(s->whx[o], s->why[o], s->whz[o]));
        s->cvhx[c] = s->vhx[o] + hw.y * off.z - hw.z * off.y;
        s->cvhy[c] = s->vhy[o] + hw.z * off.x - hw.x * off.z;
        s->cvhz[c] = s->vhz[o] + hw.x * off.y - hw.y * off.x;
    }
}

/* transformTrianglesKernel (body:71-112) */
static void transform_triangles(ora_solver* s)
{
    for (int i = 0; i < s->nt; ++i) {
        const double* b = s->bpose + 13 * s->tbody[i];
        const double* l = s->tlocal + 15 * (size_t)i;
        double* w = s->tworld + 15 * (size_t)i;
        const double qw = b[3], qx = b[4], qy = b[5], qz = b[6];
        for (int v = 0; v < 5; ++v) {
            /* rotateByQuatIv forms the nine matrix entries first (body:25-49) */
            const double r00 = 1.0 - 2.0 * (qy * qy + qz * qz), r01 = 2.0 * (qx * qy - qw * qz),
                         r02 = 2.0 * (qx * qz + qw * qy), r10 = 2.0 * (qx * qy + qw * qz),
                         r11 = 1.0 - 2.0 * (qx * qx + qz * qz), r12 = 2.0 * (qy * qz - qw * qx),
                         r20 = 2.0 * (qx * qz - qw * qy), r21 = 2.0 * (qy * qz + qw * qx),
                         r22 = 1.0 - 2.0 * (qx * qx + qy * qy);
            const double X = l[3 * v], Y = l[3 * v + 1], Z = l[3 * v + 2];
            double ox = r00 * X + r01 * Y + r02 * Z, oy = r10 * X + r11 * Y + r12 * Z,
                   oz = r20 * X + r21 * Y + r22 * Z;
            if (v < 4) { ox += b[0]; oy += b[1]; oz += b[2]; }
            w[3 * v] = ox; w[3 * v + 1] = oy; w[3 * v + 2] = oz;
        }
        double* bb = s->taabb + 6 * (size_t)i;
        bb[0] = w[0]; bb[1] = w[1]; bb[2] = w[2]; bb[3] = w[0]; bb[4] = w[1]; bb[5] = w[2];
        for (int v = 1; v < 4; ++v) {
            bb[0] = fmin(bb[0], w[3 * v]); bb[1] = fmin(bb[1], w[3 * v + 1]); bb[2] = fmin(bb[2], w[3 * v + 2]);
            bb[3] = fmax(bb[3], w[3 * v]); bb[4] = fmax(bb[4], w[3 * v + 1]); bb[5] = fmax(bb[5], w[3 * v + 2]);
        }
    }
}

static int cmp_keyid(const void* a, const void* b)
{
    const KeyId* p = a; const KeyId* q = b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->id < q->id ? -1 : (p->id > q->id);
}

/* computeComponentKeysKernel + gpuBuildComponentGrid (grid:51-71,121-145) */
static int build_component_grid(ora_solver* s)
{
    if (s->nc <= 0 || s->meshSize <= 0.0)
        return fail(DEM_E_INVALID, "invalid component grid shape or mesh size");
    s->invMesh = 1.0 / s->meshSize;
    const double minCell = -(double)(1 << 20), maxCell = (double)((1 << 20) - 1);
    KeyId* kv = malloc(sizeof(KeyId) * (size_t)s->nc);
    int overflow = 0;
    for (int i = 0; i < s->nc; ++i) {
        const double cxd = floor(s->cx[i] * s->invMesh), cyd = floor(s->cy[i] * s->invMesh),
                     czd = floor(s->cz[i] * s->invMesh);
        kv[i].id = i;
        if (!isfinite(cxd) || !isfinite(cyd) || !isfinite(czd) || cxd < minCell || cxd > maxCell ||
            cyd < minCell || cyd > maxCell || czd < minCell || czd > maxCell) {
            ++overflow;
            kv[i].key = 0;
            continue;
        }
        kv[i].key = ora_pack_cell((int)cxd, (int)cyd, (int)czd);
    }
    if (overflow) {
        free(kv);
        return fail(DEM_E_CAPACITY, "component grid coordinate exceeds signed 21-bit hash range");
    }
    qsort(kv, (size_t)s->nc, sizeof(KeyId), cmp_keyid);
    for (int i = 0; i < s->nc; ++i) { s->gkey[i] = kv[i].key; s->gid[i] = kv[i].id; }
    free(kv);
    return DEM_OK;
}

/* countTriangleCellsKernel, fillTriangleCellsKernel, gpuBuildTriangleGrid
   (contact:409-476, 1567-1610) */
static int build_triangle_grid(ora_solver* s)
{
    if (s->nt <= 0) { s->tentries = 0; return DEM_OK; }
    const double inv = 1.0 / s->meshSize, pad = s->maxRadius;
    const double minCell = -(double)(1 << 20), maxCell = (double)((1 << 20) - 1);
    int* counts = calloc((size_t)s->nt, sizeof(int));
    int overflow = 0;
    unsigned long long total = 0;
    for (int i = 0; i < s->nt; ++i) {
        const double* bb = s->taabb + 6 * (size_t)i;
        const double mnx = floor((bb[0] - pad) * inv), mny = floor((bb[1] - pad) * inv),
                     mnz = floor((bb[2] - pad) * inv), mxx = floor((bb[3] + pad) * inv),
                     mxy = floor((bb[4] + pad) * inv), mxz = floor((bb[5] + pad) * inv);
        if (!isfinite(mnx) || !isfinite(mny) || !isfinite(mnz) || !isfinite(mxx) || !isfinite(mxy) ||
            !isfinite(mxz) || mnx < minCell || mny < minCell || mnz < minCell || mxx > maxCell ||
            mxy > maxCell || mxz > maxCell) {
            ++overflow; counts[i] = 0; continue;
        }
        const long long nx = (long long)(int)mxx - (int)mnx + 1, ny = (long long)(int)mxy - (int)mny + 1,
                        nz = (long long)(int)mxz - (int)mnz + 1;
        if (nx <= 0 || ny <= 0 || nz <= 0 || nx > LLONG_MAX / ny || nx * ny > LLONG_MAX / nz) {
            ++overflow; counts[i] = 0; continue;
        }
        const long long c64 = nx * ny * nz;
        if (c64 > INT_MAX) { ++overflow; counts[i] = 0; continue; }
        counts[i] = (int)c64;
        total += (unsigned long long)c64;
    }
    if (overflow != 0 || total > (unsigned long long)INT_MAX) {
        free(counts);
        return fail(DEM_E_CAPACITY, "triangle grid entry count exceeds the supported 32-bit index range");
    }
    s->tentries = (int)total;
    KeyId* kv = malloc(sizeof(KeyId) * (total ? total : 1));
    size_t out = 0;
    for (int i = 0; i < s->nt; ++i) {
        const double* bb = s->taabb + 6 * (size_t)i;
        const int mnx = (int)floor((bb[0] - pad) * inv), mny = (int)floor((bb[1] - pad) * inv),
                  mnz = (int)floor((bb[2] - pad) * inv), mxx = (int)floor((bb[3] + pad) * inv),
                  mxy = (int)floor((bb[4] + pad) * inv), mxz = (int)floor((bb[5] + pad) * inv);
        int local = 0;
        for (int X = mnx; X <= mxx && local < counts[i]; ++X)
            for (int Y = mny; Y <= mxy && local < counts[i]; ++Y)
                for (int Z = mnz; Z <= mxz && local < counts[i]; ++Z) {
                    kv[out].key = ora_pack_cell(X, Y, Z);
                    kv[out].id = i; /* ties resolve by fill position = triangle order */
                    ++out; ++local;
                }
    }
    qsort(kv, out, sizeof(KeyId), cmp_keyid);
    free(s->tkey); free(s->ttri);
    ALLOC(s->tkey, out); ALLOC(s->ttri, out);
    for (size_t k = 0; k < out; ++k) { s->tkey[k] = kv[k].key; s->ttri[k] = kv[k].id; }
    free(kv); free(counts);
    return DEM_OK;
}

/* clearForcesKernel (contact:478-528) */
static void clear_forces(ora_solver* s)
{
    memset(s->fcontact, 0, sizeof(double) * 6 * s->n);
    memset(s->ffield, 0, sizeof(double) * 6 * s->n);
    memset(s->fwall, 0, sizeof(double) * 6 * s->n);
    memset(s->ftotal, 0, sizeof(double) * 6 * s->n);
    if (s->fbp) memset(s->fbp, 0, sizeof(double) * (size_t)s->nb * s->n * 6);
    if (s->fbd) memset(s->fbd, 0, sizeof(double) * (size_t)s->nb * s->n * 19);
    memset(s->regimeInstant, 0, sizeof(uint64_t) * NREG * s->nb);
    memset(s->regimeState, 0, sizeof(int32_t) * (size_t)s->nb * s->nc);
    memset(s->microState, 0, sizeof(double) * (size_t)s->nb * s->nc * NMICRO);
    memset(s->microInstant, 0, sizeof(double) * NMICRO * s->nb);
    memset(s->fbr, 0, sizeof(double) * 6 * s->nb);
    memset(s->bodyRow, 0, sizeof(double) * 6 * (size_t)s->nb * s->nc);
    memset(s->fzRow, 0, sizeof(double) * (size_t)s->nb * s->nc);
    s->pcount = s->bcount = s->phigh = s->bhigh = 0;
    s->bmaxdepth = 0.0; s->bfz = 0.0;
}

/* ---- history (contact:190-243) ------------------------------------------ */
static int acquire_slot(uint64_t* keys, int32_t* lastSeen, double* values, int comp,
                        uint64_t key, int step, int slots, int* overflow, int* highWater)
{
    const int begin = comp * slots;
    int empty = -1;
    for (int slot = 0; slot < slots; ++slot) {
        const int index = begin + slot;
        if (keys[index] == key) {
            lastSeen[index] = step;
            if (slot + 1 > *highWater) *highWater = slot + 1;
            return index;
        }
        if (keys[index] == 0 && empty < 0) empty = index;
    }
    if (empty < 0) { ++*overflow; return -1; }
    keys[empty] = key;
    lastSeen[empty] = step;
    for (int k = 0; k < 9; ++k) values[(size_t)empty * 9 + k] = 0.0;
    if (empty - begin + 1 > *highWater) *highWater = empty - begin + 1;
    return empty;
}

static void expire_history(uint64_t* keys, int32_t* lastSeen, double* values, int comp,
                           int step, int slots)
{
    const int begin = comp * slots;
    for (int slot = 0; slot < slots; ++slot) {
        const int index = begin + slot;
        if (keys[index] != 0 && lastSeen[index] < step - 1) {
            keys[index] = 0;
            lastSeen[index] = -1;
            for (int k = 0; k < 9; ++k) values[(size_t)index * 9 + k] = 0.0;
        }
    }
}

static V3 hload(const double* values, int slot, int cat)
{
    const size_t b = (size_t)slot * 9 + (size_t)cat * 3;
    return v3(values[b], values[b + 1], values[b + 2]);
}
static void hstore(double* values, int slot, int cat, V3 v)
{
    const size_t b = (size_t)slot * 9 + (size_t)cat * 3;
    values[b] = v.x; values[b + 1] = v.y; values[b + 2] = v.z;
}

/* transportTangentHistory / transportAxialHistory (contact:245-265) */
static V3 transport_tangent(V3 previous, V3 normal)
{
    const double pn = norm3(previous);
    const V3 projected = sub(previous, mul(normal, dot(previous, normal)));
    const double qn = norm3(projected);
    if (pn <= 1.0e-30 || qn <= 1.0e-30) return v3(0.0, 0.0, 0.0);
    return mul(projected, pn / qn);
}
static V3 transport_axial(V3 previous, V3 normal)
{
    const double pn = norm3(previous);
    if (pn <= 1.0e-30) return v3(0.0, 0.0, 0.0);
    const double alignment = dot(previous, normal);
    if (fabs(alignment) <= 1.0e-30) return v3(0.0, 0.0, 0.0);
    return mul(normal, alignment < 0.0 ? -pn : pn);
}

/* limitedHistoryResponseRates (contact:267-299), yield test
   gpuElasticTrialBranchYielded (contact.cuh:18-22) */
static V3 limited_rates(V3 historyRate, V3 responseRate, double stiffness, double damping,
                        double dt, V3 previous, double limit, double* values, int slot,
                        int cat, int store, int* yielded)
{
    const V3 delta = mul(historyRate, dt);
    const V3 inc = mul(delta, stiffness);
    const V3 elasticTrial = add(inc, previous);
    const V3 trial = add(mul(responseRate, damping), elasticTrial);
    const double trialNorm = norm3(elasticTrial);
    V3 response;
    const int y = trialNorm > limit;
    if (yielded) *yielded = y;
    if (y) {
        const double deltaNorm = norm3(delta);
        if (deltaNorm > 1.0e-14) {
            response = mul(elasticTrial, limit / trialNorm);
        } else {
            const double pn = norm3(previous);
            response = pn > 1.0e-14 ? mul(previous, limit / pn) : v3(0.0, 0.0, 0.0);
        }
        if (store) hstore(values, slot, cat, response);
    } else {
        response = trial;
        if (store) hstore(values, slot, cat, add(previous, inc));
    }
    return response;
}

/* wallHistoryResponse (contact:951-973) */
static V3 wall_response(V3 rate, double stiffness, double damping, double dt, V3 previous,
                        double limit, int checkDamped, int zeroAtZero, double* values,
                        int slot, int cat)
{
    const V3 delta = mul(rate, dt);
    const V3 elasticTrial = add(previous, mul(delta, stiffness));
    const V3 staticTrial = add(elasticTrial, mul(rate, damping));
    const double testNorm = norm3(checkDamped ? staticTrial : elasticTrial);
    if (testNorm > limit) {
        const double deltaNorm = norm3(delta);
        const V3 lim = checkDamped ? staticTrial : elasticTrial;
        const V3 response = deltaNorm > 1.0e-14 && testNorm > 1.0e-14
                                ? mul(lim, limit / testNorm)
                                : (zeroAtZero ? v3(0.0, 0.0, 0.0) : previous);
        hstore(values, slot, cat, response);
        return response;
    }
    hstore(values, slot, cat, elasticTrial);
    return staticTrial;
}

/* cN / cS damping coefficients, as computed inline at contact:619-626 */
static double damping_coeff(double eps, double k, double m)
{
    const double l = log(eps);
    return -2.0 * l * sqrt(k * m / (kPi * kPi + l * l));
}

static V3 comp_pos(const ora_solver* s, int c) { return v3(s->cx[c], s->cy[c], s->cz[c]); }
static V3 comp_vel(const ora_solver* s, int c) { return v3(s->cvx[c], s->cvy[c], s->cvz[c]); }
static V3 comp_vh(const ora_solver* s, int c) { return v3(s->cvhx[c], s->cvhy[c], s->cvhz[c]); }
static V3 comp_off(const ora_solver* s, int c) { return v3(s->cox[c], s->coy[c], s->coz[c]); }
static V3 omega_world(const ora_solver* s, int i)
{
    return rotate_iv(s->qw[i], s->qx[i], s->qy[i], s->qz[i], v3(s->wx[i], s->wy[i], s->wz[i]));
}
static V3 omega_half_world(const ora_solver* s, int i)
{
    return rotate_iv(s->qw[i], s->qx[i], s->qy[i], s->qz[i], v3(s->whx[i], s->why[i], s->whz[i]));
}
static V3 body_omega_world(const ora_solver* s, int b)
{
    const double* p = s->bpose + 13 * b;
    return rotate_iv(p[3], p[4], p[5], p[6], v3(p[10], p[11], p[12]));
}

/* Per-thread contact counters of one evaluation (merged after the owner
   loop; counts add, high-water marks and depths take the maximum). */
typedef struct {
    int pcount, pover, phigh, bcount, bover, bhigh;
    double bmaxdepth;
} Tally;

/* particleContact (contact:579-706) */
static void particle_contact(ora_solver* s, Tally* t, const Mech* m, int ci, int cj, int step,
                             V3* forceSum, V3* torqueSum)
{
    const int i = s->cowner[ci], j = s->cowner[cj];
    if (i == j) return;
    const V3 pi = comp_pos(s, ci), pj = comp_pos(s, cj);
    const V3 delta = sub(pj, pi);
    const double d2 = dot(delta, delta);
    const double ri = s->cr[ci], rj = s->cr[cj], rs = ri + rj;
    if (d2 >= rs * rs || d2 <= 1.0e-28) return;
    const double dist = sqrt(d2);
    const double depth = rs - dist;
    const V3 dir = mul(delta, 1.0 / dist);
    const V3 evi = comp_vel(s, ci), evj = comp_vel(s, cj);
    const V3 erv = sub(evj, evi);
    const V3 mvi = m->dt > 0.0 && s->status[i] != -1 ? comp_vh(s, ci) : evi;
    const V3 mvj = m->dt > 0.0 && s->status[j] != -1 ? comp_vh(s, cj) : evj;
    const V3 mrv = sub(mvj, mvi);
    const V3 ewi = omega_world(s, i), ewj = omega_world(s, j);
    const V3 mwi = m->dt > 0.0 && s->status[i] != -1 ? omega_half_world(s, i) : ewi;
    const V3 mwj = m->dt > 0.0 && s->status[j] != -1 ? omega_half_world(s, j) : ewj;
    const double reduced = s->mass[i] * s->mass[j] / (s->mass[i] + s->mass[j]);
    const double kS = 2.0 * m->kN / 7.0;
    const double cN = damping_coeff(m->epsN, m->kN, reduced);
    const double cS = damping_coeff(m->epsS, kS, reduced);
    const double R = ri * rj / rs;
    const double brs = m->beta * R * m->beta * R;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m->kN * brs, cR = cN * brs;
    V3 fn = add(mul(dir, -m->kN * depth), mul(dir, cN * dot(erv, dir)));
    fn = add(fn, mul(dir, m->cohesion * 4.0 * brs));
    const uint64_t key = (uint64_t)(uint32_t)s->cid[cj] + 1ull;
    const int slot = acquire_slot(s->hpk, s->hpl, s->hpv, ci, key, step, PP_SLOTS, &t->pover, &t->phigh);
    if (slot < 0) return;
    const double li = (ri * ri - rj * rj + d2) / (2.0 * dist);
    const double lj = dist - li;
    V3 etv = sub(erv, mul(dir, dot(erv, dir)));
    etv = add(etv, mul(cross(dir, ewi), li));
    etv = add(etv, mul(cross(dir, ewj), lj));
    V3 mtv = sub(mrv, mul(dir, dot(mrv, dir)));
    mtv = add(mtv, mul(cross(dir, mwi), li));
    mtv = add(mtv, mul(cross(dir, mwj), lj));
    const V3 oldTangent = transport_tangent(hload(s->hpv, slot, 1), dir);
    const V3 friction = limited_rates(mtv, etv, kS, cS, m->dt, oldTangent, m->mu * norm3(fn),
                                      s->hpv, slot, 1, 1, NULL);
    const V3 erw = sub(ewj, ewi), mrw = sub(mwj, mwi);
    const V3 etw = mul(dir, dot(erw, dir)), mtw = mul(dir, dot(mrw, dir));
    const V3 erl = sub(erw, etw), mrl = sub(mrw, mtw);
    const double ms = m->beta * R * norm3(fn);
    const V3 twist = limited_rates(mtw, etw, kT, cT, m->dt, transport_axial(hload(s->hpv, slot, 2), dir),
                                   m->mu * m->muT * ms, s->hpv, slot, 2, 1, NULL);
    const V3 roll = limited_rates(mrl, erl, kR, cR, m->dt, transport_tangent(hload(s->hpv, slot, 0), dir),
                                  m->muR * ms, s->hpv, slot, 0, 1, NULL);
    const V3 cf = add(fn, friction);
    const V3 couple = add(add(twist, roll), mul(cross(dir, friction), li));
    *forceSum = add(*forceSum, cf);
    *torqueSum = add(*torqueSum, add(couple, cross(comp_off(s, ci), cf)));
    ++t->pcount;
}

/* wallContact (contact:975-1063) */
static void wall_contact(ora_solver* s, Tally* t, const Mech* m, int comp, int w, int step, V3* forceSum,
                         V3* torqueSum)
{
    const int particle = s->cowner[comp];
    const V3 outward = v3(s->wnx[w], s->wny[w], s->wnz[w]);
    const V3 normal = mul(outward, -1.0);
    const V3 origin = v3(s->wox[w], s->woy[w], s->woz[w]);
    const V3 position = comp_pos(s, comp);
    const double distance = dot(sub(position, origin), outward);
    const double depth = s->cr[comp] - distance;
    if (depth <= 0.0) return;
    const uint64_t key = (uint64_t)w + 1ull;
    const int slot = acquire_slot(s->hwk, s->hwl, s->hwv, comp, key, step, PW_SLOTS, &t->bover, &t->bhigh);
    if (slot < 0) return;
    const double radius = s->cr[comp];
    const double lever = radius - depth;
    const V3 velocity = comp_vel(s, comp);
    const V3 omega = omega_world(s, particle);
    const V3 rv = mul(velocity, -1.0);
    V3 tv = sub(rv, mul(normal, dot(rv, normal)));
    tv = add(tv, mul(cross(normal, omega), lever));
    const V3 rw = mul(omega, -1.0);
    const V3 twistRate = mul(normal, dot(rw, normal));
    V3 rollRate = sub(rw, twistRate);
    if (fabs(lever) > 1.0e-30) rollRate = add(rollRate, mul(cross(normal, tv), 0.5 / lever));
    const double kS = 2.0 * m->kN / 7.0;
    const double cN = damping_coeff(m->epsN, m->kN, s->mass[particle]);
    const double cS = damping_coeff(m->epsS, kS, s->mass[particle]);
    const double brs = m->beta * radius * m->beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m->kN * brs, cR = cN * brs;
    const V3 en = mul(normal, -m->kN * depth);
    const V3 dn = mul(normal, cN * dot(rv, normal));
    const V3 fn = add(en, dn);
    const double flim = m->mu * norm3(fn);
    const double ms = m->beta * radius * norm3(fn);
    const V3 friction = wall_response(tv, kS, cS, m->dt, transport_tangent(hload(s->hwv, slot, 1), normal),
                                      flim, 0, 0, s->hwv, slot, 1);
    const V3 twist = wall_response(twistRate, kT, cT, m->dt, transport_axial(hload(s->hwv, slot, 2), normal),
                                   m->mu * m->muT * ms, 0, 1, s->hwv, slot, 2);
    const V3 roll = wall_response(rollRate, kR, cR, m->dt, transport_tangent(hload(s->hwv, slot, 0), normal),
                                  m->muR * ms, 1, 1, s->hwv, slot, 0);
    const V3 cf = add(fn, friction);
    const V3 couple = add(add(twist, roll), mul(cross(normal, friction), lever));
    *forceSum = add(*forceSum, cf);
    *torqueSum = add(*torqueSum, add(couple, cross(comp_off(s, comp), cf)));
}

/* ---- arc tangent ---------------------------------------------------------
 * The area routine needs atan2.  Library atan2 differs between the CPU and
 * the GPU in the last bits, which can flip the yield branch of a
 * particle-triangle contact (the manifold stores a tangential value of norm
 * exactly mu|Fn|, which the per-triangle pass compares against mu|Fn|).  Both
 * sides therefore use this one evaluation, built only from +, -, *, /, sqrt:
 * one half-angle reduction atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))) brings
 * t <= 1 below tan(pi/8) ~ 0.414, then the Taylor series
 * atan(z) = z - z^3/3 + ... - z^43/43 (truncation < 1e-17 relative).  The
 * result is within 3 ulp of the correctly rounded value. */
static double series_atan(double t)
{
    t = t / (1.0 + sqrt(1.0 + t * t));
    const double z = t * t;
    double p = -1.0 / 43.0;
    p = 1.0 / 41.0 + z * p;
    p = -1.0 / 39.0 + z * p;
    p = 1.0 / 37.0 + z * p;
    p = -1.0 / 35.0 + z * p;
    p = 1.0 / 33.0 + z * p;
    p = -1.0 / 31.0 + z * p;
    p = 1.0 / 29.0 + z * p;
    p = -1.0 / 27.0 + z * p;
    p = 1.0 / 25.0 + z * p;
    p = -1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = -1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = -1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = -1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = -1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = -1.0 / 3.0 + z * p;
    p = 1.0 + z * p;
    return (t * p) * 2.0;
}

double ora_atan2(double y, double x)
{
    static const double kHalfPi = 1.5707963267948966192313216916398;
    if (isnan(x) || isnan(y)) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    double a;
    if (ax == 0.0 && ay == 0.0) a = 0.0;
    else if (isinf(ax) && isinf(ay)) a = 0.5 * kHalfPi;
    else if (isinf(ax)) a = 0.0;
    else if (isinf(ay)) a = kHalfPi;
    else if (ay <= ax) a = series_atan(ay / ax);
    else a = kHalfPi - series_atan(ax / ay);
    if (signbit(x)) a = kPi - a;
    return signbit(y) ? -a : a;
}

/* ---- sphere section vs triangle area (contact:313-407) ------------------ */
static double cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
static double dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

static double circle_edge_area(V2 a, V2 b, double radius)
{
    const V2 d = {b.x - a.x, b.y - a.y};
    const double aa = dot2(d, d);
    double cuts[4] = {0.0, 1.0, 0.0, 0.0};
    int n = 2;
    if (aa > 1.0e-30) {
        const double bb = 2.0 * dot2(a, d);
        const double cc = dot2(a, a) - radius * radius;
        const double disc = bb * bb - 4.0 * aa * cc;
        if (disc > 0.0) {
            const double root = sqrt(disc);
            const double t0 = (-bb - root) / (2.0 * aa);
            const double t1 = (-bb + root) / (2.0 * aa);
            if (t0 > 0.0 && t0 < 1.0) cuts[n++] = t0;
            if (t1 > 0.0 && t1 < 1.0) cuts[n++] = t1;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (cuts[j] < cuts[i]) { const double t = cuts[i]; cuts[i] = cuts[j]; cuts[j] = t; }
    double area = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        const double ta = cuts[i], tb = cuts[i + 1];
        const V2 u = {a.x + d.x * ta, a.y + d.y * ta};
        const V2 v = {a.x + d.x * tb, a.y + d.y * tb};
        const double h = 0.5 * (ta + tb);
        const V2 mid = {a.x + d.x * h, a.y + d.y * h};
        if (dot2(mid, mid) <= radius * radius) area += 0.5 * cross2(u, v);
        else area += 0.5 * radius * radius * ora_atan2(cross2(u, v), dot2(u, v));
    }
    return area;
}

static double circle_triangle_area(V3 center, V3 normal, V3 a, V3 b, V3 c, double radius)
{
    const V3 axis = fabs(normal.x) < 0.8 ? v3(1.0, 0.0, 0.0) : v3(0.0, 1.0, 0.0);
    const V3 u = normalized(cross(normal, axis));
    const V3 v = cross(normal, u);
    const V3 ra = sub(a, center), rb = sub(b, center), rc = sub(c, center);
    const V2 pa = {dot(ra, u), dot(ra, v)}, pb = {dot(rb, u), dot(rb, v)}, pc = {dot(rc, u), dot(rc, v)};
    double area = circle_edge_area(pa, pb, radius) + circle_edge_area(pb, pc, radius) +
                  circle_edge_area(pc, pa, radius);
    area = fabs(area);
    return fmin(area, kPi * radius * radius);
}

double ora_circle_triangle_area(const double center[3], const double normal[3], const double a[3],
                                const double b[3], const double c[3], double radius)
{
    return circle_triangle_area(v3(center[0], center[1], center[2]), v3(normal[0], normal[1], normal[2]),
                                v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2]), radius);
}

static V3 tri_vertex(const ora_solver* s, int t, int v)
{
    const double* w = s->tworld + 15 * (size_t)t;
    return v3(w[3 * v], w[3 * v + 1], w[3 * v + 2]);
}

/* bodyTriangleAreaScale (contact:383-407) */
static double tri_area_scale(const ora_solver* s, int comp, int t)
{
    const V3 normal = tri_vertex(s, t, 4);
    const V3 a = tri_vertex(s, t, 1), b = tri_vertex(s, t, 2), c = tri_vertex(s, t, 3);
    const V3 position = comp_pos(s, comp);
    const double distance = dot(normal, sub(position, a));
    if (distance <= 0.0) return 0.0;
    const double crs = s->cr[comp] * s->cr[comp] - distance * distance;
    if (crs <= 1.0e-30) return 0.0;
    const double cr = sqrt(crs);
    const V3 cc = sub(position, mul(normal, distance));
    const double area = circle_triangle_area(cc, normal, a, b, c, cr);
    if (area <= 0.0) return 0.0;
    const double scale = fmin(1.0, area / (kPi * crs));
    return scale > 1.0e-12 ? scale : 0.0;
}

/* bodyContact (contact:708-853) with the manifold history read-only */
static void body_contact(ora_solver* s, Tally* ty, const Mech* m, int comp, int t, int slot, V3* forceSum,
                         V3* torqueSum)
{
    const int particle = s->cowner[comp];
    const int body = s->tbody[t];
    V3 normal = tri_vertex(s, t, 4);
    const V3 a = tri_vertex(s, t, 1), b = tri_vertex(s, t, 2), c = tri_vertex(s, t, 3);
    const V3 position = comp_pos(s, comp);
    const double distance = dot(normal, sub(position, a));
    if (distance <= 0.0) return;
    const double depth = s->cr[comp] - distance;
    if (depth <= 0.0) return;
    const double crs = s->cr[comp] * s->cr[comp] - distance * distance;
    if (crs <= 1.0e-30) return;
    const double cr = sqrt(crs);
    const V3 cc = sub(position, mul(normal, distance));
    const double area = circle_triangle_area(cc, normal, a, b, c, cr);
    if (area <= 0.0) return;
    const double scale = fmin(1.0, area / (kPi * crs));
    if (scale <= 1.0e-12) return;
    normal = mul(normal, -1.0);
    const double* bp = s->bpose + 13 * body;
    const V3 center = v3(bp[0], bp[1], bp[2]);
    const V3 arm = sub(cc, center);
    const V3 wp = omega_world(s, particle);
    const V3 wb = body_omega_world(s, body);
    const V3 bv = add(v3(bp[7], bp[8], bp[9]), cross(wb, arm));
    const V3 rv = sub(bv, comp_vel(s, comp));
    const double kS = 2.0 * m->kN / 7.0;
    const double reduced = s->mass[particle];
    const double cN = damping_coeff(m->epsN, m->kN, reduced);
    const double cS = damping_coeff(m->epsS, kS, reduced);
    const double radius = s->cr[comp];
    const double brs = m->beta * radius * m->beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m->kN * brs, cR = cN * brs;
    V3 fn = add(mul(normal, -m->kN * depth), mul(normal, cN * dot(rv, normal)));
    fn = add(fn, mul(normal, m->cohesion * 4.0 * brs));
    V3 tv = sub(rv, mul(normal, dot(rv, normal)));
    tv = add(tv, mul(cross(normal, wp), distance));
    const V3 friction = limited_rates(tv, tv, kS, cS, m->dt, transport_tangent(hload(s->hbv, slot, 1), normal),
                                      m->mu * norm3(fn), s->hbv, slot, 1, 0, NULL);
    const V3 rw = sub(wb, wp);
    const V3 twistRate = mul(normal, dot(rw, normal));
    const V3 rollRate = add(sub(rw, twistRate), mul(cross(normal, tv), 0.5 / fmax(distance, 1.0e-14)));
    const double ms = m->beta * radius * norm3(fn);
    const V3 twist = limited_rates(twistRate, twistRate, kT, cT, m->dt,
                                   transport_axial(hload(s->hbv, slot, 2), normal), m->mu * m->muT * ms,
                                   s->hbv, slot, 2, 0, NULL);
    const V3 roll = limited_rates(rollRate, rollRate, kR, cR, m->dt,
                                  transport_tangent(hload(s->hbv, slot, 0), normal), m->muR * ms, s->hbv,
                                  slot, 0, 0, NULL);
    const V3 pf = mul(add(fn, friction), scale);
    const V3 moment = add(twist, roll);
    const V3 couple = mul(add(moment, mul(cross(normal, friction), distance)), scale);
    const V3 ptq = add(couple, cross(comp_off(s, comp), pf));
    const V3 bnf = mul(fn, -scale), bfr = mul(friction, -scale);
    const V3 bnt = mul(cross(arm, fn), -scale), bft = mul(cross(arm, friction), -scale);
    const V3 btw = mul(twist, -scale), brl = mul(roll, -scale);
    const V3 btq = add(add(bnt, bft), add(btw, brl));
    *forceSum = add(*forceSum, pf);
    *torqueSum = add(*torqueSum, ptq);
    if (s->fbp) {
        double* d = s->fbp + ((size_t)body * s->n + particle) * 6;
        d[0] += pf.x; d[1] += pf.y; d[2] += pf.z; d[3] += ptq.x; d[4] += ptq.y; d[5] += ptq.z;
    }
    if (s->fbd) {
        double* d = s->fbd + ((size_t)body * s->n + particle) * 19;
        const V3 g[6] = {bnf, bfr, bnt, bft, btw, brl};
        for (int k = 0; k < 6; ++k) { d[3 * k] += g[k].x; d[3 * k + 1] += g[k].y; d[3 * k + 2] += g[k].z; }
        if (d[18] < depth) d[18] = depth;
    }
    /* the reference adds these with atomics (contact:843-852); here each
       component's contributions are summed in triangle order and the
       components are reduced afterwards in a fixed tree (tree_sum) */
    const size_t N = (size_t)s->nc;
    double* r = s->bodyRow + (size_t)body * 6 * N + comp;
    r[0] += bnf.x + bfr.x; r[N] += bnf.y + bfr.y; r[2 * N] += bnf.z + bfr.z;
    r[3 * N] += btq.x; r[4 * N] += btq.y; r[5 * N] += btq.z;
    s->fzRow[(size_t)body * N + comp] += bnf.z + bfr.z;
    ++ty->bcount;
    if (ty->bmaxdepth < depth) ty->bmaxdepth = depth;
}

/* advanceBodyManifoldHistory (contact:855-949) */
static int advance_manifold(ora_solver* s, Tally* ty, const Mech* m, int comp, int body, int step, V3 normal,
                            V3 cc, double distance, int* slide, int* twisty)
{
    *slide = 0; *twisty = 0;
    const int particle = s->cowner[comp];
    const uint64_t key = (uint64_t)(body + 1);
    const int slot = acquire_slot(s->hbk, s->hbl, s->hbv, comp, key, step, PT_SLOTS, &ty->bover, &ty->bhigh);
    if (slot < 0) return -1;
    const double* bp = s->bpose + 13 * body;
    const V3 center = v3(bp[0], bp[1], bp[2]);
    const V3 arm = sub(cc, center);
    const V3 ewp = omega_world(s, particle);
    const int dyn = m->dt > 0.0 && s->status[particle] != -1;
    const V3 mwp = dyn ? omega_half_world(s, particle) : ewp;
    const V3 wb = body_omega_world(s, body);
    const V3 bv = add(v3(bp[7], bp[8], bp[9]), cross(wb, arm));
    const V3 epv = comp_vel(s, comp);
    const V3 mpv = dyn ? comp_vh(s, comp) : epv;
    const V3 erv = sub(bv, epv), mrv = sub(bv, mpv);
    const double kS = 2.0 * m->kN / 7.0;
    const double reduced = s->mass[particle];
    const double cN = damping_coeff(m->epsN, m->kN, reduced);
    const double cS = damping_coeff(m->epsS, kS, reduced);
    const double radius = s->cr[comp];
    const double brs = m->beta * radius * m->beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m->kN * brs, cR = cN * brs;
    const double depth = radius - distance;
    V3 fn = add(mul(normal, -m->kN * depth), mul(normal, cN * dot(erv, normal)));
    fn = add(fn, mul(normal, m->cohesion * 4.0 * brs));
    V3 tv = sub(mrv, mul(normal, dot(mrv, normal)));
    tv = add(tv, mul(cross(normal, mwp), distance));
    limited_rates(tv, tv, kS, cS, m->dt, transport_tangent(hload(s->hbv, slot, 1), normal),
                  m->mu * norm3(fn), s->hbv, slot, 1, 1, slide);
    const V3 rw = sub(wb, mwp);
    const V3 twistRate = mul(normal, dot(rw, normal));
    const V3 rollRate = add(sub(rw, twistRate), mul(cross(normal, tv), 0.5 / fmax(distance, 1.0e-14)));
    const double ms = m->beta * radius * norm3(fn);
    limited_rates(twistRate, twistRate, kT, cT, m->dt, transport_axial(hload(s->hbv, slot, 2), normal),
                  m->mu * m->muT * ms, s->hbv, slot, 2, 1, twisty);
    limited_rates(rollRate, rollRate, kR, cR, m->dt, transport_tangent(hload(s->hbv, slot, 0), normal),
                  m->muR * ms, s->hbv, slot, 0, 1, NULL);
    return slot;
}

/* computeContactsKernel (contact:1065-1221), owners in row order */
static void merge_tally(ora_solver* s, const Tally* t)
{
    s->pcount += t->pcount; s->pover += t->pover; s->bcount += t->bcount; s->bover += t->bover;
    if (t->phigh > s->phigh) s->phigh = t->phigh;
    if (t->bhigh > s->bhigh) s->bhigh = t->bhigh;
    if (t->bmaxdepth > s->bmaxdepth) s->bmaxdepth = t->bmaxdepth;
}

/* The owner loop runs on ora_set_threads() threads (OpenMP).  Every write of
   an owner's iteration goes to that owner's rows (its forces, its
   components' history slots, body rows and regime flags); only the counters
   are shared, and they are merged after the loop, so results do not depend
   on the thread count. */
static void compute_contacts(ora_solver* s, const Mech* pm, const Mech* bm, const Mech* wm, int step)
{
#pragma omp parallel num_threads(s->threads)
    {
    Tally ty;
    memset(&ty, 0, sizeof ty);
#pragma omp for schedule(dynamic, 512)
    for (int i = 0; i < s->n; ++i) {
        V3 cF = v3(0, 0, 0), cT = v3(0, 0, 0), fF = v3(0, 0, 0), wF = v3(0, 0, 0), wT = v3(0, 0, 0);
        V3 bF = v3(0, 0, 0), bT = v3(0, 0, 0);
        if (pm->gravityEnabled) fF = v3(s->mass[i] * pm->gx, s->mass[i] * pm->gy, s->mass[i] * pm->gz);
        if (pm->universalGravityEnabled) {
            const V3 pi = v3(s->x[i], s->y[i], s->z[i]);
            for (int j = 0; j < s->n; ++j) {
                if (j == i) continue;
                const V3 d = sub(pi, v3(s->x[j], s->y[j], s->z[j]));
                const double d2 = dot(d, d);
                if (d2 > 1.0e-28)
                    fF = add(fF, mul(d, kGravityConstant * s->mass[i] * s->mass[j] / (d2 * sqrt(d2))));
            }
        }
        for (int comp = s->ownerStart[i]; comp < s->ownerStart[i + 1]; ++comp) {
            expire_history(s->hpk, s->hpl, s->hpv, comp, step, PP_SLOTS);
            expire_history(s->hwk, s->hwl, s->hwv, comp, step, PW_SLOTS);
            expire_history(s->hbk, s->hbl, s->hbv, comp, step, PT_SLOTS);
            for (int w = 0; w < s->nw; ++w) wall_contact(s, &ty, wm, comp, w, step, &wF, &wT);
            const V3 position = comp_pos(s, comp);
            const int ix = (int)floor(position.x * s->invMesh);
            const int iy = (int)floor(position.y * s->invMesh);
            const int iz = (int)floor(position.z * s->invMesh);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        const uint64_t key = ora_pack_cell(ix + dx, iy + dy, iz + dz);
                        const int b = lower_bound_u64(s->gkey, s->nc, key);
                        const int e = upper_bound_u64(s->gkey, s->nc, key);
                        for (int k = b; k < e; ++k) {
                            const int other = s->gid[k];
                            if (s->cowner[other] != i) particle_contact(s, &ty, pm, comp, other, step, &cF, &cT);
                        }
                    }
            if (s->nt <= 0 || s->tentries <= 0) continue;
            const uint64_t key = ora_pack_cell(ix, iy, iz);
            const int b = lower_bound_u64(s->tkey, s->tentries, key);
            const int e = upper_bound_u64(s->tkey, s->tentries, key);
            for (int body = 0; body < s->nb; ++body) {
                double total = 0.0, wD = 0.0;
                V3 wN = v3(0, 0, 0), wC = v3(0, 0, 0);
                for (int k = b; k < e; ++k) {
                    const int t = s->ttri[k];
                    if (s->tbody[t] != body) continue;
                    const double sc = tri_area_scale(s, comp, t);
                    if (sc <= 0.0) continue;
                    const V3 outward = tri_vertex(s, t, 4), a = tri_vertex(s, t, 1);
                    const double distance = dot(outward, sub(position, a));
                    const V3 cc = sub(position, mul(outward, distance));
                    total += sc;
                    wN = add(wN, mul(outward, -sc));
                    wC = add(wC, mul(cc, sc));
                    wD += distance * sc;
                }
                if (total <= 0.0) continue;
                const V3 mN = normalized(wN);
                const V3 mC = mul(wC, 1.0 / total);
                const double mD = wD / total;
                int slide = 0, twisty = 0;
                const int slot = advance_manifold(s, &ty, bm, comp, body, step, mN, mC, mD, &slide, &twisty);
                if (slot < 0) continue;
                s->regimeState[(size_t)body * s->nc + comp] = 1 | (slide ? 2 : 0) | (twisty ? 4 : 0);
                Mech ro = *bm;
                ro.dt = 0.0;
                for (int k = b; k < e; ++k) {
                    const int t = s->ttri[k];
                    if (s->tbody[t] != body) continue;
                    body_contact(s, &ty, &ro, comp, t, slot, &bF, &bT);
                }
            }
        }
        double* d;
        d = s->fcontact + 6 * i; d[0] += cF.x; d[1] += cF.y; d[2] += cF.z; d[3] += cT.x; d[4] += cT.y; d[5] += cT.z;
        d = s->ffield + 6 * i; d[0] += fF.x; d[1] += fF.y; d[2] += fF.z; d[3] += 0.0; d[4] += 0.0; d[5] += 0.0;
        d = s->fwall + 6 * i; d[0] += wF.x; d[1] += wF.y; d[2] += wF.z; d[3] += wT.x; d[4] += wT.y; d[5] += wT.z;
        const V3 tF = add(add(cF, fF), add(wF, bF));
        const V3 tT = add(cT, add(wT, bT));
        d = s->ftotal + 6 * i; d[0] += tF.x; d[1] += tF.y; d[2] += tF.z; d[3] += tT.x; d[4] += tT.y; d[5] += tT.z;
        s->fx[i] = tF.x; s->fy[i] = tF.y; s->fz[i] = tF.z;
        s->tx[i] = tT.x; s->ty[i] = tT.y; s->tz[i] = tT.z;
    }
#pragma omp critical
    merge_tally(s, &ty);
    }
}

/* observeBodyMicrostateKernel (contact:1226-1356) */
static void observe_microstate(ora_solver* s, const Mech* m, int step)
{
    const double inv = 1.0 / s->meshSize;
    for (int index = 0; index < s->nb * s->nc; ++index) {
        if ((s->regimeState[index] & 1) == 0) continue;
        const int body = index / s->nc;
        const int comp = index - body * s->nc;
        const V3 position = comp_pos(s, comp);
        const int ix = (int)floor(position.x * inv), iy = (int)floor(position.y * inv),
                  iz = (int)floor(position.z * inv);
        const uint64_t gk = ora_pack_cell(ix, iy, iz);
        const int b = lower_bound_u64(s->tkey, s->tentries, gk);
        const int e = upper_bound_u64(s->tkey, s->tentries, gk);
        double total = 0.0, wD = 0.0;
        V3 wN = v3(0, 0, 0), wC = v3(0, 0, 0);
        for (int k = b; k < e; ++k) {
            const int t = s->ttri[k];
            if (s->tbody[t] != body) continue;
            const double sc = tri_area_scale(s, comp, t);
            if (sc <= 0.0) continue;
            const V3 outward = tri_vertex(s, t, 4), a = tri_vertex(s, t, 1);
            const double distance = dot(outward, sub(position, a));
            const V3 cc = sub(position, mul(outward, distance));
            total += sc;
            wN = add(wN, mul(outward, -sc));
            wC = add(wC, mul(cc, sc));
            wD += distance * sc;
        }
        if (total <= 0.0) continue;
        const uint64_t hk = (uint64_t)(body + 1);
        int slot = -1;
        for (int o = 0; o < PT_SLOTS; ++o) {
            const int sl = comp * PT_SLOTS + o;
            if (s->hbk[sl] == hk && s->hbl[sl] == step) { slot = sl; break; }
        }
        if (slot < 0) continue;
        const V3 normal = normalized(wN);
        const V3 cc = mul(wC, 1.0 / total);
        const double distance = wD / total;
        const int particle = s->cowner[comp];
        const double* bp = s->bpose + 13 * body;
        const V3 center = v3(bp[0], bp[1], bp[2]);
        const V3 arm = sub(cc, center);
        const V3 wb = body_omega_world(s, body);
        const V3 bv = add(v3(bp[7], bp[8], bp[9]), cross(wb, arm));
        const V3 rv = sub(bv, comp_vel(s, comp));
        const double cN = damping_coeff(m->epsN, m->kN, s->mass[particle]);
        const double radius = s->cr[comp];
        const double brs = m->beta * radius * m->beta * radius;
        const double depth = radius - distance;
        V3 fn = add(mul(normal, -m->kN * depth), mul(normal, cN * dot(rv, normal)));
        fn = add(fn, mul(normal, m->cohesion * 4.0 * brs));
        const double fnn = norm3(fn);
        const double ms = m->beta * radius * fnn;
        const double tlim = m->mu * fnn;
        const double wlim = m->mu * m->muT * ms;
        const V3 th = hload(s->hbv, slot, 1), wh = hload(s->hbv, slot, 2);
        const V3 axisZ = rotate_iv(bp[3], bp[4], bp[5], bp[6], v3(0.0, 0.0, 1.0));
        const V3 radial = sub(arm, mul(axisZ, dot(arm, axisZ)));
        const double tu = tlim > 1.0e-30 ? fmin(1.0, norm3(th) / tlim) : 0.0;
        const double wu = wlim > 1.0e-30 ? fmin(1.0, norm3(wh) / wlim) : 0.0;
        const V3 btw = mul(wh, -total), bnl = mul(fn, -total), bth = mul(th, -total);
        const V3 btt = cross(arm, bth);
        double* ms22 = s->microState + (size_t)index * NMICRO;
        ms22[0] = fnn * total; ms22[1] = fnn * radius * total; ms22[2] = fnn * norm3(radial) * total;
        ms22[3] = btw.x; ms22[4] = btw.y; ms22[5] = btw.z;
        ms22[6] = norm3(wh) * total; ms22[7] = wlim * total; ms22[8] = norm3(th) * total;
        ms22[9] = tlim * total; ms22[10] = wu * total; ms22[11] = tu * total; ms22[12] = total;
        ms22[13] = bnl.x; ms22[14] = bnl.y; ms22[15] = bnl.z;
        ms22[16] = bth.x; ms22[17] = bth.y; ms22[18] = bth.z;
        ms22[19] = btt.x; ms22[20] = btt.y; ms22[21] = btt.z;
    }
}

/* accumulateBodyRegimeKernel (contact:530-567): integer counters in any
   order; the 22 microstate sums per body through tree_sum */
static void accumulate_regime(ora_solver* s, int interval)
{
    for (int index = 0; index < s->nb * s->nc; ++index) {
        const int st = s->regimeState[index];
        if ((st & 1) == 0) continue;
        const int sl = (st & 2) != 0, tw = (st & 4) != 0;
        const int body = index / s->nc;
        uint64_t* ri = s->regimeInstant + NREG * body;
        uint64_t* rv = s->regimeInterval + NREG * body;
        const int primary = tw ? 3 : (sl ? 2 : 1);
        ri[0]++; ri[primary]++;
        if (sl) ri[4]++;
        if (tw) ri[5]++;
        if (sl && tw) ri[6]++;
        if (!interval) continue;
        rv[0]++; rv[primary]++;
        if (sl) rv[4]++;
        if (tw) rv[5]++;
        if (sl && tw) rv[6]++;
    }
}

static double tree_sum(ora_solver* s, const double* row);

static void reduce_body_rows(ora_solver* s, int interval)
{
    const size_t N = (size_t)s->nc;
    double* row = malloc(sizeof(double) * N);
    s->bfz = 0.0;
    for (int b = 0; b < s->nb; ++b) {
        for (int k = 0; k < 6; ++k) s->fbr[6 * b + k] = tree_sum(s, s->bodyRow + ((size_t)b * 6 + k) * N);
        s->bfz += tree_sum(s, s->fzRow + (size_t)b * N);
        for (int f = 0; f < NMICRO; ++f) {
            for (size_t c = 0; c < N; ++c) row[c] = s->microState[((size_t)b * s->nc + c) * NMICRO + f];
            const double v = tree_sum(s, row);
            s->microInstant[NMICRO * b + f] = v;
            if (interval) s->microInterval[NMICRO * b + f] += v;
        }
    }
    free(row);
}

/* Sum of one per-component row taken in component input order (owner, then
 * component id; independent of the grid) with a fixed tree: tiles of 4096
 * rows; in a tile, 256 partial sums over rows tid, tid + 256, ...; each group
 * of 64 partials folded pairwise with partner lane ^ 32, ^ 16, ..., ^ 1; the
 * 4 group results added in order; the tile results added in order.  Any
 * order is a faithful restatement of the reference's atomic sums; this one
 * is deterministic. */
static double tree_sum(ora_solver* s, const double* row)
{
    const long n = s->nc;
    const double* v = row;
    double total = 0.0;
    for (long base = 0; base < n; base += 4096) {
        double part[256];
        for (int tid = 0; tid < 256; ++tid) {
            double acc = 0.0;
            for (int k = 0; k < 16; ++k) {
                const long e = base + (long)k * 256 + tid;
                if (e < n) acc += v[e];
            }
            part[tid] = acc;
        }
        double blk = 0.0;
        for (int w = 0; w < 4; ++w) {
            double lane[64], nxt[64];
            for (int l = 0; l < 64; ++l) lane[l] = part[w * 64 + l];
            for (int o = 32; o > 0; o >>= 1) {
                for (int l = 0; l < 64; ++l) nxt[l] = lane[l] + lane[l ^ o];
                memcpy(lane, nxt, sizeof lane);
            }
            blk += lane[0];
        }
        total += blk;
    }
    return total;
}

/* accelerationKernel (particle:71-101) */
static void acceleration(ora_solver* s)
{
    for (int i = 0; i < s->n; ++i) {
        s->ax[i] = s->fx[i] / s->mass[i]; s->ay[i] = s->fy[i] / s->mass[i]; s->az[i] = s->fz[i] / s->mass[i];
        const double qw = s->qw[i], qx = s->qx[i], qy = s->qy[i], qz = s->qz[i];
        const double tx = (1.0 - 2.0 * (qy * qy + qz * qz)) * s->tx[i] + 2.0 * (qx * qy + qw * qz) * s->ty[i] +
                          2.0 * (qx * qz - qw * qy) * s->tz[i];
        const double ty = 2.0 * (qx * qy - qw * qz) * s->tx[i] + (1.0 - 2.0 * (qx * qx + qz * qz)) * s->ty[i] +
                          2.0 * (qy * qz + qw * qx) * s->tz[i];
        const double tz = 2.0 * (qx * qz + qw * qy) * s->tx[i] + 2.0 * (qy * qz - qw * qx) * s->ty[i] +
                          (1.0 - 2.0 * (qx * qx + qy * qy)) * s->tz[i];
        const double wx = s->wx[i], wy = s->wy[i], wz = s->wz[i];
        const double ix = s->inx[i], iy = s->iny[i], iz = s->inz[i];
        s->awx[i] = (tx - (iz - iy) * wz * wy) / ix;
        s->awy[i] = (ty - (ix - iz) * wx * wz) / iy;
        s->awz[i] = (tz - (iy - ix) * wy * wx) / iz;
    }
}

/* CudaDemSolver::computeForces (backend:383-524) */
int ora_compute_forces(ora_solver* s, const double* pose13, double* loads6, int collect)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (pose13 && s->nb > 0) {
        memcpy(s->bpose, pose13, sizeof(double) * 13 * s->nb);
        transform_triangles(s);
    }
    update_components(s);
    int rc = build_component_grid(s);
    if (rc) return rc;
    if (s->nt > 0) {
        rc = build_triangle_grid(s);
        if (rc) return rc;
    }
    clear_forces(s);
    const int step = s->stepIndex++;
    Mech pm = s->pp, bm = s->pt, wm = s->pw;
    if (step == 0) { pm.dt = 0.0; bm.dt = 0.0; wm.dt = 0.0; }
    compute_contacts(s, &pm, &bm, &wm, step);
    if (s->nb > 0) {
        observe_microstate(s, &bm, step);
        accumulate_regime(s, bm.dt > 0.0);
        reduce_body_rows(s, bm.dt > 0.0);
    }
    if (bm.dt > 0.0 && s->nb > 0) ++s->intervalSteps;
    acceleration(s);
    if (collect) {
        dem_step_stats st;
        memset(&st, 0, sizeof st);
        st.particle_contacts = s->pcount; st.body_contacts = s->bcount;
        st.body_max_depth = s->bmaxdepth; st.body_contact_fz = s->bfz;
        if (s->nb > 0) {
            memcpy(s->cachedRegimeInstant, s->regimeInstant, sizeof(uint64_t) * NREG * s->nb);
            memcpy(s->cachedRegimeInterval, s->regimeInterval, sizeof(uint64_t) * NREG * s->nb);
            st.body_regime_interval_native_steps = s->intervalSteps;
            memset(s->regimeInterval, 0, sizeof(uint64_t) * NREG * s->nb);
            s->intervalSteps = 0;
            memcpy(s->cachedMicroInstant, s->microInstant, sizeof(double) * NMICRO * s->nb);
            memcpy(s->cachedMicroInterval, s->microInterval, sizeof(double) * NMICRO * s->nb);
            memset(s->microInterval, 0, sizeof(double) * NMICRO * s->nb);
        }
        st.particle_history_overflows = s->pover; st.body_history_overflows = s->bover;
        st.particle_history_high_water = s->phigh; st.body_history_high_water = s->bhigh;
        st.triangle_grid_overflows = 0;
        s->cached = st;
        if (st.particle_history_overflows != 0)
            return fail(DEM_E_CAPACITY, "CUDA particle contact-history capacity exceeded; increase kGpuParticleHistorySlots");
        if (st.body_history_overflows != 0)
            return fail(DEM_E_CAPACITY, "CUDA wall/body contact-history capacity exceeded; increase the configured history slots");
    }
    if (loads6 && s->nb > 0) {
        for (int b = 0; b < s->nb; ++b) {
            for (int k = 0; k < 6; ++k) loads6[6 * b + k] = s->fbr[6 * b + k];
            if (s->pp.gravityEnabled) {
                loads6[6 * b] += s->pp.gx * s->bmass[b];
                loads6[6 * b + 1] += s->pp.gy * s->bmass[b];
                loads6[6 * b + 2] += s->pp.gz * s->bmass[b];
            }
        }
    }
    return DEM_OK;
}

/* quatDerivative / normalizeQuat (particle:149-183) */
static void quat_derivative(double qw, double qx, double qy, double qz, double wx, double wy, double wz,
                            double* dqw, double* dqx, double* dqy, double* dqz)
{
    *dqw = 0.5 * (-qx * wx - qy * wy - qz * wz);
    *dqx = 0.5 * (qw * wx + qy * wz - qz * wy);
    *dqy = 0.5 * (qw * wy - qx * wz + qz * wx);
    *dqz = 0.5 * (qw * wz + qx * wy - qy * wx);
}
static void normalize_quat(double* qw, double* qx, double* qy, double* qz)
{
    const double n2 = *qw * *qw + *qx * *qx + *qy * *qy + *qz * *qz;
    if (n2 <= 1.0e-24) { *qw = 1.0; *qx = 0.0; *qy = 0.0; *qz = 0.0; return; }
    const double inv = 1.0 / sqrt(n2);
    *qw *= inv; *qx *= inv; *qy *= inv; *qz *= inv;
}

static void half_quat(ora_solver* s, int i, double halfDt)
{
    double qhw = s->qw[i], qhx = s->qx[i], qhy = s->qy[i], qhz = s->qz[i];
    double dqw, dqx, dqy, dqz;
    quat_derivative(qhw, qhx, qhy, qhz, s->wx[i], s->wy[i], s->wz[i], &dqw, &dqx, &dqy, &dqz);
    qhw += dqw * halfDt; qhx += dqx * halfDt; qhy += dqy * halfDt; qhz += dqz * halfDt;
    normalize_quat(&qhw, &qhx, &qhy, &qhz);
    quat_derivative(qhw, qhx, qhy, qhz, s->whx[i], s->why[i], s->whz[i], &s->hdqw[i], &s->hdqx[i],
                    &s->hdqy[i], &s->hdqz[i]);
}

/* initializeLeapfrogKernel (particle:185-211) */
int ora_init_half_step(ora_solver* s, double halfDt)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    for (int i = 0; i < s->n; ++i) {
        s->vhx[i] = s->vx[i] + s->ax[i] * halfDt; s->vhy[i] = s->vy[i] + s->ay[i] * halfDt;
        s->vhz[i] = s->vz[i] + s->az[i] * halfDt;
        s->whx[i] = s->wx[i] + s->awx[i] * halfDt; s->why[i] = s->wy[i] + s->awy[i] * halfDt;
        s->whz[i] = s->wz[i] + s->awz[i] * halfDt;
        half_quat(s, i, halfDt);
    }
    return DEM_OK;
}

/* advancePositionKernel (particle:213-238) */
int ora_advance_particles(ora_solver* s, double dt, double* vmax)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    for (int i = 0; i < s->n; ++i) {
        if (s->status[i] == -1) continue;
        s->x[i] += s->vhx[i] * dt; s->y[i] += s->vhy[i] * dt; s->z[i] += s->vhz[i] * dt;
        s->qw[i] += s->hdqw[i] * dt; s->qx[i] += s->hdqx[i] * dt; s->qy[i] += s->hdqy[i] * dt;
        s->qz[i] += s->hdqz[i] * dt;
        normalize_quat(&s->qw[i], &s->qx[i], &s->qy[i], &s->qz[i]);
        s->vx[i] = s->vhx[i] + s->ax[i] * (0.5 * dt); s->vy[i] = s->vhy[i] + s->ay[i] * (0.5 * dt);
        s->vz[i] = s->vhz[i] + s->az[i] * (0.5 * dt);
        s->wx[i] = s->whx[i] + s->awx[i] * (0.5 * dt); s->wy[i] = s->why[i] + s->awy[i] * (0.5 * dt);
        s->wz[i] = s->whz[i] + s->awz[i] * (0.5 * dt);
    }
    if (vmax) *vmax = 0.0;
    return DEM_OK;
}

/* advanceHalfVelocityKernel (particle:240-265) */
int ora_finish_particle_step(ora_solver* s, double dt, double halfDtUnused)
{
    (void)halfDtUnused;
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    for (int i = 0; i < s->n; ++i) {
        s->vhx[i] += s->ax[i] * dt; s->vhy[i] += s->ay[i] * dt; s->vhz[i] += s->az[i] * dt;
        s->whx[i] += s->awx[i] * dt; s->why[i] += s->awy[i] * dt; s->whz[i] += s->awz[i] * dt;
        half_quat(s, i, 0.5 * dt);
    }
    return DEM_OK;
}

int ora_download_particle_state(const ora_solver* s, double* pos, double* vel, double* ang, double* quat)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    for (int i = 0; i < s->n; ++i) {
        if (pos) { pos[3 * i] = s->x[i]; pos[3 * i + 1] = s->y[i]; pos[3 * i + 2] = s->z[i]; }
        if (vel) { vel[3 * i] = s->vx[i]; vel[3 * i + 1] = s->vy[i]; vel[3 * i + 2] = s->vz[i]; }
        if (ang) { ang[3 * i] = s->wx[i]; ang[3 * i + 1] = s->wy[i]; ang[3 * i + 2] = s->wz[i]; }
        if (quat) {
            quat[4 * i] = s->qw[i]; quat[4 * i + 1] = s->qx[i]; quat[4 * i + 2] = s->qy[i];
            quat[4 * i + 3] = s->qz[i];
        }
    }
    return DEM_OK;
}

int ora_download_force_output(const ora_solver* s, double* contact, double* field, double* wall,
                              double* total, double* bp, double* bd)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    const size_t n6 = sizeof(double) * 6 * (size_t)s->n;
    if (contact) memcpy(contact, s->fcontact, n6);
    if (field) memcpy(field, s->ffield, n6);
    if (wall) memcpy(wall, s->fwall, n6);
    if (total) memcpy(total, s->ftotal, n6);
    if (bp) {
        if (!s->fbp && s->nb > 0) return fail(DEM_E_INVALID, "body-particle force output was not kept");
        if (s->nb > 0) memcpy(bp, s->fbp, sizeof(double) * (size_t)s->nb * s->n * 6);
    }
    if (bd) {
        if (!s->fbd && s->nb > 0) return fail(DEM_E_INVALID, "body detail output was not kept");
        if (s->nb > 0) memcpy(bd, s->fbd, sizeof(double) * (size_t)s->nb * s->n * 19);
    }
    return DEM_OK;
}

int ora_get_step_stats(const ora_solver* s, dem_step_stats* st, uint64_t* ri, uint64_t* rv, double* mi,
                   double* mv)
{
    if (!s) return fail(DEM_E_INVALID, "null solver");
    if (st) *st = s->cached;
    if (ri) memcpy(ri, s->cachedRegimeInstant, sizeof(uint64_t) * NREG * s->nb);
    if (rv) memcpy(rv, s->cachedRegimeInterval, sizeof(uint64_t) * NREG * s->nb);
    if (mi) memcpy(mi, s->cachedMicroInstant, sizeof(double) * NMICRO * s->nb);
    if (mv) memcpy(mv, s->cachedMicroInterval, sizeof(double) * NMICRO * s->nb);
    return DEM_OK;
}

int ora_particle_grid_entry_count(const ora_solver* s, size_t* out)
{
    if (!s || !out) return fail(DEM_E_INVALID, "null argument");
    *out = (size_t)s->nc;
    return DEM_OK;
}
int ora_triangle_grid_entry_count(const ora_solver* s, size_t* out)
{
    if (!s || !out) return fail(DEM_E_INVALID, "null argument");
    *out = (size_t)s->tentries;
    return DEM_OK;
}

int ora_debug_grid(const ora_solver* s, uint64_t* keys, int32_t* ids)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (keys) memcpy(keys, s->gkey, sizeof(uint64_t) * s->nc);
    if (ids) memcpy(ids, s->gid, sizeof(int32_t) * s->nc);
    return DEM_OK;
}

int ora_debug_history(const ora_solver* s, uint64_t* pk, double* pv, uint64_t* wk, double* wv,
                      uint64_t* bk, double* bv)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    const size_t nc = (size_t)s->nc;
    if (pk) memcpy(pk, s->hpk, sizeof(uint64_t) * nc * PP_SLOTS);
    if (pv) memcpy(pv, s->hpv, sizeof(double) * nc * PP_SLOTS * 9);
    if (wk) memcpy(wk, s->hwk, sizeof(uint64_t) * nc * PW_SLOTS);
    if (wv) memcpy(wv, s->hwv, sizeof(double) * nc * PW_SLOTS * 9);
    if (bk) memcpy(bk, s->hbk, sizeof(uint64_t) * nc * PT_SLOTS);
    if (bv) memcpy(bv, s->hbv, sizeof(double) * nc * PT_SLOTS * 9);
    return DEM_OK;
}

int ora_debug_triangle_grid(const ora_solver* s, uint64_t* keys, int32_t* tris)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (keys && s->tentries) memcpy(keys, s->tkey, sizeof(uint64_t) * s->tentries);
    if (tris && s->tentries) memcpy(tris, s->ttri, sizeof(int32_t) * s->tentries);
    return DEM_OK;
}

/* ---- contact-history checkpoint (CudaDemSolver::saveContactHistory /
 * loadContactHistory, cuda_backend_cuda.cu:545-770).  Layout (native byte
 * order, checked by the endian marker): "DMHIST2\0", u32 version 2, u32
 * 0x01020304, u32 owners, u32 components, u32 slots[3] = {64, 16, 64}, u32
 * topology record size 40, u32 owner-state record size 188; per component
 * (i32 id, i32 owner id, f64 radius, body x, y, z); per owner (i32 id, 23 f64:
 * x y z vx vy vz qw qx qy qz wx wy wz vhx vhy vhz whx why whz hdqw hdqx hdqy
 * hdqz); then the particle, wall and body history keys (u64, component-major,
 * slot-minor) and values (9 f64 per slot) in that order.  DMHIST1 ("DMHIST1\0",
 * version 1, components, slots[3], keys, values) is accepted for the legacy
 * sphere topology only and restores no leapfrog state. */
static int write_all(FILE* f, const void* p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }
static int read_all(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int ora_save_contact_history(const ora_solver* s, const char* path)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "cannot save contact history before CUDA initialization");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(DEM_E_IO, "cannot create CUDA contact-history checkpoint: %s", path);
    const char magic[8] = {'D', 'M', 'H', 'I', 'S', 'T', '2', '\0'};
    const uint32_t head[10] = {2u, 0x01020304u, (uint32_t)s->n, (uint32_t)s->nc, PP_SLOTS, PW_SLOTS, PT_SLOTS,
                               2u * 4u + 4u * 8u, 4u + 23u * 8u, 0u};
    int ok = write_all(f, magic, 8) && write_all(f, head, 9 * sizeof(uint32_t));
    for (int i = 0; i < s->nc && ok; ++i) {
        const int32_t ids[2] = {s->cid[i], s->pid[s->cowner[i]]};
        const double geo[4] = {s->cr[i], s->cbx[i], s->cby[i], s->cbz[i]};
        ok = write_all(f, ids, sizeof ids) && write_all(f, geo, sizeof geo);
    }
    for (int i = 0; i < s->n && ok; ++i) {
        const int32_t id = s->pid[i];
        const double st[23] = {s->x[i],   s->y[i],   s->z[i],   s->vx[i],  s->vy[i],   s->vz[i],
                               s->qw[i],  s->qx[i],  s->qy[i],  s->qz[i],  s->wx[i],   s->wy[i],
                               s->wz[i],  s->vhx[i], s->vhy[i], s->vhz[i], s->whx[i],  s->why[i],
                               s->whz[i], s->hdqw[i], s->hdqx[i], s->hdqy[i], s->hdqz[i]};
        ok = write_all(f, &id, sizeof id) && write_all(f, st, sizeof st);
    }
    const size_t nc = (size_t)s->nc;
    if (ok)
        ok = write_all(f, s->hpk, 8 * nc * PP_SLOTS) && write_all(f, s->hwk, 8 * nc * PW_SLOTS) &&
             write_all(f, s->hbk, 8 * nc * PT_SLOTS) && write_all(f, s->hpv, 72 * nc * PP_SLOTS) &&
             write_all(f, s->hwv, 72 * nc * PW_SLOTS) && write_all(f, s->hbv, 72 * nc * PT_SLOTS);
    if (fclose(f) != 0) ok = 0;
    return ok ? DEM_OK : fail(DEM_E_IO, "failed to write CUDA contact-history checkpoint: %s", path);
}

int ora_load_contact_history(ora_solver* s, const char* path)
{
    if (!s || !s->initialized) return fail(DEM_E_INVALID, "cannot load contact history before CUDA initialization");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(DEM_E_IO, "cannot open CUDA contact-history checkpoint: %s", path);
    int rc = DEM_OK;
    char magic[8];
    uint32_t version = 0, comps = 0, slots[3] = {0, 0, 0};
    double* state = NULL;
    uint64_t *pk = NULL, *wk = NULL, *bk = NULL;
    double *pv = NULL, *wv = NULL, *bv = NULL;
#define TRUNC() do { rc = fail(DEM_E_IO, "truncated CUDA contact-history checkpoint: %s", path); goto done; } while (0)
    if (!read_all(f, magic, 8) || !read_all(f, &version, 4)) TRUNC();
    s->restored = 0;
    int v2 = 0;
    if (memcmp(magic, "DMHIST1", 8) == 0 && version == 1) {
        if (!read_all(f, &comps, 4) || !read_all(f, slots, 12)) TRUNC();
        if (!s->legacySphere) {
            rc = fail(DEM_E_INVALID, "legacy DMHIST1 checkpoint is forbidden for clump topology: %s", path);
            goto done;
        }
    } else if (memcmp(magic, "DMHIST2", 8) == 0 && version == 2) {
        uint32_t endian = 0, owners = 0, topo = 0, rec = 0;
        if (!read_all(f, &endian, 4) || !read_all(f, &owners, 4) || !read_all(f, &comps, 4) ||
            !read_all(f, slots, 12) || !read_all(f, &topo, 4) || !read_all(f, &rec, 4))
            TRUNC();
        if (endian != 0x01020304u || owners != (uint32_t)s->n || topo != 40u || rec != 188u) {
            rc = fail(DEM_E_INVALID, "incompatible DMHIST2 header: %s", path);
            goto done;
        }
        if (comps != (uint32_t)s->nc) {
            rc = fail(DEM_E_INVALID, "DMHIST2 component count mismatch: %s", path);
            goto done;
        }
        for (uint32_t i = 0; i < comps; ++i) {
            int32_t ids[2];
            double geo[4];
            if (!read_all(f, ids, 8) || !read_all(f, geo, 32)) TRUNC();
            if (ids[0] != s->cid[i] || ids[1] != s->pid[s->cowner[i]] || geo[0] != s->cr[i] ||
                geo[1] != s->cbx[i] || geo[2] != s->cby[i] || geo[3] != s->cbz[i]) {
                rc = fail(DEM_E_INVALID, "DMHIST2 component topology mismatch: %s", path);
                goto done;
            }
        }
        state = calloc((size_t)s->n * 23 + 1, sizeof(double));
        for (int i = 0; i < s->n; ++i) {
            int32_t id;
            if (!read_all(f, &id, 4) || !read_all(f, state + (size_t)i * 23, 23 * 8)) TRUNC();
            if (id != s->pid[i]) {
                rc = fail(DEM_E_INVALID, "DMHIST2 owner identity mismatch: %s", path);
                goto done;
            }
        }
        v2 = 1;
    } else {
        rc = fail(DEM_E_INVALID, "unsupported contact-history checkpoint: %s", path);
        goto done;
    }
    if (comps != (uint32_t)s->nc || slots[0] != PP_SLOTS || slots[1] != PW_SLOTS || slots[2] != PT_SLOTS) {
        rc = fail(DEM_E_INVALID, "incompatible CUDA contact-history checkpoint: %s", path);
        goto done;
    }
    {
        const size_t nc = (size_t)s->nc;
        pk = malloc(8 * nc * PP_SLOTS + 8); wk = malloc(8 * nc * PW_SLOTS + 8); bk = malloc(8 * nc * PT_SLOTS + 8);
        pv = malloc(72 * nc * PP_SLOTS + 8); wv = malloc(72 * nc * PW_SLOTS + 8); bv = malloc(72 * nc * PT_SLOTS + 8);
        if (!read_all(f, pk, 8 * nc * PP_SLOTS) || !read_all(f, wk, 8 * nc * PW_SLOTS) ||
            !read_all(f, bk, 8 * nc * PT_SLOTS) || !read_all(f, pv, 72 * nc * PP_SLOTS) ||
            !read_all(f, wv, 72 * nc * PW_SLOTS) || !read_all(f, bv, 72 * nc * PT_SLOTS))
            TRUNC();
        char extra;
        if (fread(&extra, 1, 1, f) == 1) {
            rc = fail(DEM_E_INVALID, "oversized CUDA contact-history checkpoint: %s", path);
            goto done;
        }
        if (v2) {
            for (int i = 0; i < s->n; ++i) {
                const double* q = state + (size_t)i * 23;
                s->x[i] = q[0]; s->y[i] = q[1]; s->z[i] = q[2];
                s->vx[i] = q[3]; s->vy[i] = q[4]; s->vz[i] = q[5];
                s->qw[i] = q[6]; s->qx[i] = q[7]; s->qy[i] = q[8]; s->qz[i] = q[9];
                s->wx[i] = q[10]; s->wy[i] = q[11]; s->wz[i] = q[12];
                s->vhx[i] = q[13]; s->vhy[i] = q[14]; s->vhz[i] = q[15];
                s->whx[i] = q[16]; s->why[i] = q[17]; s->whz[i] = q[18];
                s->hdqw[i] = q[19]; s->hdqx[i] = q[20]; s->hdqy[i] = q[21]; s->hdqz[i] = q[22];
            }
            s->restored = 1;
        }
        memcpy(s->hpk, pk, 8 * nc * PP_SLOTS); memcpy(s->hwk, wk, 8 * nc * PW_SLOTS);
        memcpy(s->hbk, bk, 8 * nc * PT_SLOTS); memcpy(s->hpv, pv, 72 * nc * PP_SLOTS);
        memcpy(s->hwv, wv, 72 * nc * PW_SLOTS); memcpy(s->hbv, bv, 72 * nc * PT_SLOTS);
        for (size_t i = 0; i < nc * PP_SLOTS; ++i) s->hpl[i] = -1;
        for (size_t i = 0; i < nc * PW_SLOTS; ++i) s->hwl[i] = -1;
        for (size_t i = 0; i < nc * PT_SLOTS; ++i) s->hbl[i] = -1;
        s->stepIndex = 0;
    }
#undef TRUNC
done:
    fclose(f);
    free(state); free(pk); free(wk); free(bk); free(pv); free(wv); free(bv);
    return rc;
}

int ora_has_restored_leapfrog_state(const ora_solver* s, int* out)
{
    if (!s || !out) return fail(DEM_E_INVALID, "null argument");
    *out = s->restored;
    return DEM_OK;
}

/*
 * scene_rsa.c -- seeded random sequential insertion of non-overlapping
 * spheres into an axis-aligned box, with a uniform cell grid for the
 * overlap test.  Builds the synthetic packings the benchmarks and tests run
 * on; deterministic for a given seed on any machine (splitmix64 stream).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static uint64_t next_u64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double next_unit(uint64_t* s) { return (double)(next_u64(s) >> 11) * (1.0 / 9007199254740992.0); }

/*
 * Places spheres of the given radii (in the given order) at uniform random
 * positions inside [lo, hi] (each centre inset by its radius), rejecting
 * overlaps.  Returns the number placed (the first `placed` radii have
 * positions in pos[3*i..]); stops at the first sphere that could not be
 * placed within max_attempts tries.  gap adds a clearance between spheres.
 */
int dem_rsa_box(int n, const double* radius, const double lo[3], const double hi[3],
                uint64_t seed, int max_attempts, double gap, double* pos)
{
    double rmax = 0.0;
    for (int i = 0; i < n; ++i) rmax = radius[i] > rmax ? radius[i] : rmax;
    const double cell = 2.0 * rmax + gap;
    int dims[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = (int)floor((hi[a] - lo[a]) / cell) + 1;
        if (dims[a] < 1) dims[a] = 1;
    }
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    int* head = malloc(sizeof(int) * ncell);
    int* next = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!head || !next) { free(head); free(next); return -1; }
    for (size_t c = 0; c < ncell; ++c) head[c] = -1;
    uint64_t state = seed;
    int placed = 0;
    for (int i = 0; i < n; ++i) {
        const double r = radius[i];
        int ok = 0;
        for (int attempt = 0; attempt < max_attempts && !ok; ++attempt) {
            double p[3];
            int cidx[3];
            for (int a = 0; a < 3; ++a) {
                const double span = (hi[a] - lo[a]) - 2.0 * r;
                p[a] = lo[a] + r + (span > 0.0 ? span * next_unit(&state) : 0.5 * span);
                cidx[a] = (int)floor((p[a] - lo[a]) / cell);
                if (cidx[a] < 0) cidx[a] = 0;
                if (cidx[a] >= dims[a]) cidx[a] = dims[a] - 1;
            }
            int clear = 1;
            for (int dx = -1; dx <= 1 && clear; ++dx)
                for (int dy = -1; dy <= 1 && clear; ++dy)
                    for (int dz = -1; dz <= 1 && clear; ++dz) {
                        const int X = cidx[0] + dx, Y = cidx[1] + dy, Z = cidx[2] + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= dims[0] || Y >= dims[1] || Z >= dims[2]) continue;
                        for (int j = head[((size_t)X * dims[1] + Y) * dims[2] + Z]; j >= 0; j = next[j]) {
                            const double ex = p[0] - pos[3 * j], ey = p[1] - pos[3 * j + 1],
                                         ez = p[2] - pos[3 * j + 2];
                            const double lim = r + radius[j] + gap;
                            if (ex * ex + ey * ey + ez * ez < lim * lim) { clear = 0; break; }
                        }
                    }
            if (!clear) continue;
            pos[3 * i] = p[0]; pos[3 * i + 1] = p[1]; pos[3 * i + 2] = p[2];
            const size_t c = ((size_t)cidx[0] * dims[1] + cidx[1]) * dims[2] + cidx[2];
            next[i] = head[c];
            head[c] = i;
            ok = 1;
        }
        if (!ok) break;
        placed = i + 1;
    }
    free(head);
    free(next);
    return placed;
}

/*
 * Overlap relaxation of a packing (positions in/out): Jacobi sweeps in which
 * every overlapping pair moves each sphere half the overlap apart along the
 * centre line and the box faces (x, y on both sides, z below) push spheres
 * back inside.  Stops when the largest overlap falls below tol * (r_i + r_j)
 * or after max_iter sweeps; returns the sweeps used (-1: allocation failure).
 * Threads split each sweep (OpenMP).
 */
int dem_relax_packing(int n, const double* radius, const double lo[3], const double hi[3],
                      int max_iter, double tol, double* pos)
{
    double rmax = 0.0, top = lo[2];
    for (int i = 0; i < n; ++i) rmax = radius[i] > rmax ? radius[i] : rmax;
    for (int i = 0; i < n; ++i) top = pos[3 * i + 2] + radius[i] > top ? pos[3 * i + 2] + radius[i] : top;
    const double cell = 2.0 * rmax;
    int dims[3];
    dims[0] = (int)floor((hi[0] - lo[0]) / cell) + 1;
    dims[1] = (int)floor((hi[1] - lo[1]) / cell) + 1;
    dims[2] = (int)floor((top - lo[2]) / cell) + 16;
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    int* start = malloc(sizeof(int) * (ncell + 1));
    int* order = malloc(sizeof(int) * (size_t)n);
    int* cellof = malloc(sizeof(int) * (size_t)n);
    double* disp = malloc(sizeof(double) * 3 * (size_t)n);
    if (!start || !order || !cellof || !disp) {
        free(start); free(order); free(cellof); free(disp);
        return -1;
    }
    int it = 0;
    for (; it < max_iter; ++it) {
        for (size_t c = 0; c <= ncell; ++c) start[c] = 0;
        for (int i = 0; i < n; ++i) {
            int c3[3];
            for (int d = 0; d < 3; ++d) {
                c3[d] = (int)floor((pos[3 * i + d] - lo[d]) / cell);
                if (c3[d] < 0) c3[d] = 0;
                if (c3[d] >= dims[d]) c3[d] = dims[d] - 1;
            }
            cellof[i] = (int)(((size_t)c3[0] * dims[1] + c3[1]) * dims[2] + c3[2]);
            ++start[cellof[i] + 1];
        }
        for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
        for (int i = 0; i < n; ++i) order[start[cellof[i]]++] = i;
        for (size_t c = ncell; c > 0; --c) start[c] = start[c - 1];
        start[0] = 0;
        double worst = 0.0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : worst)
        for (int i = 0; i < n; ++i) {
            double dx = 0.0, dy = 0.0, dz = 0.0;
            const int c = cellof[i];
            const int cz = c % dims[2], cy = (c / dims[2]) % dims[1], cx = c / (dims[2] * dims[1]);
            for (int ox = -1; ox <= 1; ++ox)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int oz = -1; oz <= 1; ++oz) {
                        const int X = cx + ox, Y = cy + oy, Z = cz + oz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= dims[0] || Y >= dims[1] || Z >= dims[2]) continue;
                        const size_t cc = ((size_t)X * dims[1] + Y) * dims[2] + Z;
                        for (int k = start[cc]; k < start[cc + 1]; ++k) {
                            const int j = order[k];
                            if (j == i) continue;
                            const double ex = pos[3 * i] - pos[3 * j], ey = pos[3 * i + 1] - pos[3 * j + 1],
                                         ez = pos[3 * i + 2] - pos[3 * j + 2];
                            const double d2 = ex * ex + ey * ey + ez * ez, rs = radius[i] + radius[j];
                            if (d2 >= rs * rs || d2 <= 0.0) continue;
                            const double d = sqrt(d2), ov = rs - d;
                            if (ov / rs > worst) worst = ov / rs;
                            const double k2 = 0.5 * ov / d;
                            dx += k2 * ex; dy += k2 * ey; dz += k2 * ez;
                        }
                    }
            const double r = radius[i];
            const double* pi = pos + 3 * i;
            if (pi[0] - r < lo[0]) dx += lo[0] - (pi[0] - r);
            if (pi[0] + r > hi[0]) dx -= pi[0] + r - hi[0];
            if (pi[1] - r < lo[1]) dy += lo[1] - (pi[1] - r);
            if (pi[1] + r > hi[1]) dy -= pi[1] + r - hi[1];
            if (pi[2] - r < lo[2]) dz += lo[2] - (pi[2] - r);
            disp[3 * i] = dx; disp[3 * i + 1] = dy; disp[3 * i + 2] = dz;
        }
        for (size_t k = 0; k < 3 * (size_t)n; ++k) pos[k] += 0.9 * disp[k];
        if (worst < tol) break;
    }
    free(start); free(order); free(cellof); free(disp);
    return it;
}

/*
 * Sequential deposition under gravity.  Spheres are dropped one at a time,
 * in the given order, at seeded uniform (x, y) over the floor footprint
 * [lo, hi] (centres inset by the radius).  Each falls straight down to its
 * first contact with a deposited sphere or a plane, then descends along its
 * contacts: a step straight down, after which the centre is pushed out of
 * every deposited sphere and plane it penetrates; a step that does not lower
 * the centre is halved, until the steps fall below 1e-4 of the radius.  The
 * sphere then rests in a hollow of its supports (three spheres, or a plane
 * and up to two spheres), as in the rolling-deposition packings of granular
 * beds.  Contact distances are (r_i + r_j)(1 - preload) between spheres and
 * r (1 - preload) to a plane, so resting contacts overlap by that fraction.
 * Planes are (origin, inward unit normal) pairs, the floor among them.
 * Returns 0, or -1 on allocation failure.
 */
typedef struct {
    int* idx;  /* sphere indices, by increasing centre height */
    int len, cap;
} Column;

static int column_insert(Column* c, int i, const double* pos)
{
    if (c->len == c->cap) {
        const int cap = c->cap ? 2 * c->cap : 64;
        int* p = realloc(c->idx, sizeof(int) * (size_t)cap);
        if (!p) return -1;
        c->idx = p;
        c->cap = cap;
    }
    int k = c->len++;
    const double z = pos[3 * i + 2];
    while (k > 0 && pos[3 * c->idx[k - 1] + 2] > z) {
        c->idx[k] = c->idx[k - 1];
        --k;
    }
    c->idx[k] = i;
    return 0;
}

/* first index of a column whose centre height is >= z */
static int column_lower(const Column* c, const double* pos, double z)
{
    int a = 0, b = c->len;
    while (a < b) {
        const int m = (a + b) / 2;
        if (pos[3 * c->idx[m] + 2] < z) a = m + 1;
        else b = m;
    }
    return a;
}

typedef struct {
    const double *pos, *radius, *plane_o, *plane_n;
    int nplane;
    double keep;                     /* 1 - preload */
    const double *lo, *hi;           /* floor footprint */
    const Column* col;
    int dx, dy;                      /* columns */
    double gx0, gy0, cell, rmax;
    uint64_t seed;
} Deposit;

enum { GATHER_CAP = 32 };

typedef struct {
    int *cand, *act;   /* [CAND_CAP] each */
    unsigned planes;   /* bit a: plane a is within reach (the first 32 planes) */
    /* what the result depended on: the spheres within r + rmax of the drop
     * point (horizontally) and within r + rmax + margin of each gather point */
    double drop[2], gp[GATHER_CAP][3];
    int ngp;           /* > GATHER_CAP: more gathers than recorded */
} Scratch;

enum { CAND_CAP = 4096 };

static void clamp_cols(const Deposit* D, const double* p, int lo_, int* gx, int* gy)
{
    *gx = (int)floor((p[0] - D->gx0) / D->cell);
    *gy = (int)floor((p[1] - D->gy0) / D->cell);
    if (*gx < lo_) *gx = lo_;
    if (*gy < lo_) *gy = lo_;
    if (*gx > D->dx - 1 - lo_) *gx = D->dx - 1 - lo_;
    if (*gy > D->dy - 1 - lo_) *gy = D->dy - 1 - lo_;
}

/* whether a sphere deposited at e (radius re) could have changed the result
 * recorded in W for a sphere of radius r */
static int depends_on(const Deposit* D, const Scratch* W, double r, const double* e)
{
    if (W->ngp > GATHER_CAP) return 1;
    const double fx = e[0] - W->drop[0], fy = e[1] - W->drop[1], fr = r + D->rmax;
    if (fx * fx + fy * fy < fr * fr) return 1;
    const double gr = r + 3.0 * D->rmax;  /* r + rmax + margin */
    for (int k = 0; k < W->ngp; ++k) {
        const double ex = e[0] - W->gp[k][0], ey = e[1] - W->gp[k][1], ez = e[2] - W->gp[k][2];
        if (ex * ex + ey * ey + ez * ez < gr * gr) return 1;
    }
    return 0;
}

/* Pushes centre q (radius r) out of the planes within reach and of the
 * spheres in act[0..nact), in sweeps until none is penetrated by more than
 * tol or `sweeps` sweeps are done; returns the deepest penetration of the
 * last sweep. */
static double push_out(const Deposit* D, const Scratch* W, double r, int nact, int sweeps, double tol, double* q)
{
    double worst = 0.0;
    for (int sweep = 0; sweep < sweeps; ++sweep) {
        worst = 0.0;
        for (int a = 0; a < D->nplane; ++a) {
            if (a < 32 && !(W->planes >> a & 1u)) continue;
            const double* o = D->plane_o + 3 * a;
            const double* nm = D->plane_n + 3 * a;
            const double d = (q[0] - o[0]) * nm[0] + (q[1] - o[1]) * nm[1] + (q[2] - o[2]) * nm[2];
            const double lim = r * D->keep;
            if (d < lim) {
                worst = lim - d > worst ? lim - d : worst;
                q[0] += (lim - d) * nm[0];
                q[1] += (lim - d) * nm[1];
                q[2] += (lim - d) * nm[2];
            }
        }
        for (int k = 0; k < nact; ++k) {
            const int j = W->act[k];
            const double* c = D->pos + 3 * j;
            const double R = (r + D->radius[j]) * D->keep;
            const double ex = q[0] - c[0], ey = q[1] - c[1], ez = q[2] - c[2];
            const double d2 = ex * ex + ey * ey + ez * ez;
            if (d2 >= R * R || d2 <= 0.0) continue;
            const double d = sqrt(d2);
            worst = R - d > worst ? R - d : worst;
            q[0] = c[0] + ex * (R / d);
            q[1] = c[1] + ey * (R / d);
            q[2] = c[2] + ez * (R / d);
        }
        if (worst <= tol) break;
    }
    return worst;
}

/* the candidates within `reach` of the spheres' surfaces around p, and the
 * planes within reach of the sphere around p */
static int within(const Deposit* D, Scratch* W, double r, const double* p, int ncand, double reach)
{
    W->planes = 0u;
    for (int a = 0; a < D->nplane && a < 32; ++a) {
        const double* o = D->plane_o + 3 * a;
        const double* nm = D->plane_n + 3 * a;
        const double d = (p[0] - o[0]) * nm[0] + (p[1] - o[1]) * nm[1] + (p[2] - o[2]) * nm[2];
        if (d < r + reach) W->planes |= 1u << a;
    }
    int nact = 0;
    for (int k = 0; k < ncand; ++k) {
        const int j = W->cand[k];
        const double* c = D->pos + 3 * j;
        const double ex = p[0] - c[0], ey = p[1] - c[1], ez = p[2] - c[2];
        const double lim = r + D->radius[j] + reach;
        if (ex * ex + ey * ey + ez * ez < lim * lim) W->act[nact++] = j;
    }
    return nact;
}

/* The resting position of sphere i on the deposited spheres (reads only;
 * W receives the points its result depended on).  Its random numbers come from its
 * own stream (the seed and i), so the result depends on nothing else. */
static void deposit_one(const Deposit* D, Scratch* W, int i, double* out)
{
    const Column* col = D->col;
    const double* pos = D->pos;
    const double* radius = D->radius;
    const double r = radius[i], rmax = D->rmax;
    const int dy = D->dy;
    uint64_t state = D->seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1));
    W->ngp = 0;
    const double sx = (D->hi[0] - D->lo[0]) - 2.0 * r, sy = (D->hi[1] - D->lo[1]) - 2.0 * r;
    double p[3];
    p[0] = D->lo[0] + r + (sx > 0.0 ? sx * next_unit(&state) : 0.5 * sx);
    p[1] = D->lo[1] + r + (sy > 0.0 ? sy * next_unit(&state) : 0.5 * sy);
    int cx, cy;
    clamp_cols(D, p, 1, &cx, &cy);
    /* free fall: the highest first-contact height over the spheres of the
     * 3 x 3 columns (scanned from the top) and the planes */
    double best = -1e300;
    for (int a = 0; a < D->nplane; ++a) {
        const double* o = D->plane_o + 3 * a;
        const double* nm = D->plane_n + 3 * a;
        if (nm[2] <= 1e-12) continue;
        const double z = o[2] + (r * D->keep - (p[0] - o[0]) * nm[0] - (p[1] - o[1]) * nm[1]) / nm[2];
        best = z > best ? z : best;
    }
    W->drop[0] = p[0];
    W->drop[1] = p[1];
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy) {
            const Column* c = &col[(size_t)(cx + ox) * dy + (cy + oy)];
            for (int k = c->len - 1; k >= 0; --k) {
                const int j = c->idx[k];
                const double* q = pos + 3 * j;
                const double R = (r + radius[j]) * D->keep;
                if (q[2] + r + rmax < best) break;
                const double ex = p[0] - q[0], ey = p[1] - q[1], h2 = ex * ex + ey * ey;
                if (h2 >= R * R) continue;
                const double z = q[2] + sqrt(R * R - h2);
                best = z > best ? z : best;
            }
        }
    p[2] = best;
    /* descent along the contacts.  Candidates (spheres of the 5 x 5 columns
     * within reach) are re-gathered once the centre has moved half of
     * `margin` from where they were gathered; a step moves it by at most
     * about 2 eta <= r, so none it can reach is missing.  A step goes down by
     * eta, or sideways by eta where that does not lower the centre, and is
     * taken if pushing the centre out of what it penetrates leaves it lower
     * and penetrating nothing by more than 1e-4 r; a step taken lengthens the
     * next by half (up to r / 2), one refused halves it.  The active set holds the candidates within
     * 4 etaA of the surface around aP: valid for steps of up to eta <= etaA
     * taken from within etaA of aP. */
    const double margin = 2.0 * rmax;
    double g[3] = {1e300, 1e300, 1e300}, aP[3] = {1e300, 1e300, 1e300}, etaA = 0.0;
    int ncand = 0, nact = 0;
    double eta = 0.5 * r;
    const double etaMin = 1e-4 * r;
    for (int it = 0; it < 4000 && eta > etaMin; ++it) {
        const double mx = p[0] - g[0], my = p[1] - g[1], mz = p[2] - g[2];
        if (mx * mx + my * my + mz * mz > 0.25 * margin * margin) {
            g[0] = p[0]; g[1] = p[1]; g[2] = p[2];
            ncand = 0;
            int gx, gy;
            clamp_cols(D, p, 2, &gx, &gy);
            if (W->ngp < GATHER_CAP) {
                W->gp[W->ngp][0] = p[0]; W->gp[W->ngp][1] = p[1]; W->gp[W->ngp][2] = p[2];
            }
            ++W->ngp;
            for (int ox = -2; ox <= 2; ++ox)
                for (int oy = -2; oy <= 2; ++oy) {
                    const Column* c = &col[(size_t)(gx + ox) * dy + (gy + oy)];
                    for (int k = column_lower(c, pos, p[2] - r - rmax - margin); k < c->len; ++k) {
                        const int j = c->idx[k];
                        const double* q = pos + 3 * j;
                        if (q[2] > p[2] + r + rmax + margin) break;
                        const double ex = p[0] - q[0], ey = p[1] - q[1], ez = p[2] - q[2];
                        const double lim = r + radius[j] + margin;
                        if (ex * ex + ey * ey + ez * ez >= lim * lim) continue;
                        if (ncand < CAND_CAP) W->cand[ncand++] = j;
                    }
                }
            etaA = 0.0;
        }
        {
            const double ax = p[0] - aP[0], ay = p[1] - aP[1], az = p[2] - aP[2];
            if (eta > etaA || ax * ax + ay * ay + az * az > etaA * etaA) {
                etaA = eta;
                aP[0] = p[0]; aP[1] = p[1]; aP[2] = p[2];
                nact = within(D, W, r, p, ncand, 4.0 * eta);
            }
        }
        double q[3] = {p[0], p[1], p[2] - eta};
        double worst = push_out(D, W, r, nact, 24, 1e-4 * r, q);
        if (!(worst <= 1e-4 * r && q[2] < p[2] - 1e-3 * eta)) {
            /* no descent straight down: the centre rests in a hollow, or sits
             * on a ridge or top, where gravity's pull along the contacts
             * vanishes; a step sideways in a seeded direction (and down a
             * little) rolls it off a ridge or top, and leads back into a hollow */
            const double ang = 6.283185307179586 * next_unit(&state);
            q[0] = p[0] + eta * cos(ang);
            q[1] = p[1] + eta * sin(ang);
            q[2] = p[2] - 0.25 * eta;
            worst = push_out(D, W, r, nact, 24, 1e-4 * r, q);
        }
        if (worst <= 1e-4 * r && q[2] < p[2] - 1e-3 * eta) {
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
            eta = 1.5 * eta < 0.5 * r ? 1.5 * eta : 0.5 * r;
        } else {
            eta *= 0.5;
        }
    }
    /* the resting position, pushed clear of its contacts to 1e-12 r */
    {
        const int na = within(D, W, r, p, ncand, 1e-3 * r);
        double q[3] = {p[0], p[1], p[2]};
        if (push_out(D, W, r, na, 2000, 1e-12 * r, q) <= 1e-9 * r) {
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        }
    }
    out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
}

/* The spheres are deposited in batches: each sphere of a batch is first
 * computed on the pile as it stood before the batch (in parallel); then, in
 * order, a sphere is kept if no sphere of the batch placed before it landed
 * where it could have changed that result, and recomputed on the current pile
 * otherwise.  Either way
 * it is computed on the pile the one-at-a-time deposition would give it, so
 * the packing is that deposition's (batch 1), whatever the batch size (at
 * most 256) and the number of threads. */
int dem_deposit_packing(int n, const double* radius, const double lo[2], const double hi[2], int nplane,
                        const double* plane_o, const double* plane_n, uint64_t seed, double preload,
                        int batch, double* pos)
{
    double rmax = 0.0;
    for (int i = 0; i < n; ++i) rmax = radius[i] > rmax ? radius[i] : rmax;
    const int BATCH = batch < 1 ? 1 : (batch > 256 ? 256 : batch);
    Deposit D;
    D.pos = pos; D.radius = radius; D.plane_o = plane_o; D.plane_n = plane_n; D.nplane = nplane;
    D.keep = 1.0 - preload; D.lo = lo; D.hi = hi; D.seed = seed; D.rmax = rmax;
    D.cell = 2.0 * rmax;
    /* the planes may lean outwards: the grid spans three footprint widths */
    const double padx = hi[0] - lo[0], pady = hi[1] - lo[1];
    D.gx0 = lo[0] - padx;
    D.gy0 = lo[1] - pady;
    D.dx = (int)floor(3.0 * padx / D.cell) + 1;
    D.dy = (int)floor(3.0 * pady / D.cell) + 1;
    if (D.dx < 5) D.dx = 5;
    if (D.dy < 5) D.dy = 5;
    Column* col = calloc((size_t)D.dx * D.dy, sizeof(Column));
    int* cand = malloc(sizeof(int) * CAND_CAP * 2 * BATCH);
    double* spec = malloc(sizeof(double) * 3 * BATCH);
    Scratch* seen = malloc(sizeof(Scratch) * BATCH);
    if (!col || !cand || !spec || !seen) {
        free(col); free(cand); free(spec); free(seen);
        return -1;
    }
    D.col = col;
    int rc = 0;
    for (int b0 = 0; b0 < n && rc == 0; b0 += BATCH) {
        const int m = n - b0 < BATCH ? n - b0 : BATCH;
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < m; ++k) {
            Scratch* W = &seen[k];
            W->cand = cand + (size_t)CAND_CAP * 2 * k;
            W->act = W->cand + CAND_CAP;
            deposit_one(&D, W, b0 + k, spec + 3 * k);
        }
        for (int k = 0; k < m && rc == 0; ++k) {
            const double rk = radius[b0 + k];
            int clash = 0;
            for (int e = 0; e < k && !clash; ++e) clash = depends_on(&D, &seen[k], rk, pos + 3 * (size_t)(b0 + e));
            double* p = pos + 3 * (size_t)(b0 + k);
            if (clash) {
                Scratch W;
                W.cand = cand;
                W.act = cand + CAND_CAP;
                deposit_one(&D, &W, b0 + k, p);
            } else {
                p[0] = spec[3 * k]; p[1] = spec[3 * k + 1]; p[2] = spec[3 * k + 2];
            }
            int fx, fy;
            clamp_cols(&D, p, 0, &fx, &fy);
            if (column_insert(&col[(size_t)fx * D.dy + fy], b0 + k, pos) != 0) rc = -1;
        }
    }
    for (size_t c = 0; c < (size_t)D.dx * D.dy; ++c) free(col[c].idx);
    free(col); free(cand); free(spec); free(seen);
    return rc;
}

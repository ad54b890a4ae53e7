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

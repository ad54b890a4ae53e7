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

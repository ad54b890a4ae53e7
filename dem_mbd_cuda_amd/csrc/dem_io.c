/*
 * dem_io.c -- libdemio.so: the reference's per-frame output files, written
 * with the same printf conversions so a frame is byte-identical to what the
 * reference host writes for the same values.
 *
 *   ph.%08d.bt  PARTICLE::StateOutput (particle.cpp:320-386): per particle
 *               "%10d %10d" id, status, then 15 x " %16.17e": mass, radius,
 *               pos[3], vel[3], quat[4], angular velocity[3]
 *   F.%08d.bt   FORCE::ParticlesForceOutput (particle.cpp:389-545): per
 *               particle 21 x "%16.12e" separated by one space: contact[6],
 *               wall[6], field[3], total[6]; with bodies, 6 x " %16.12e" per
 *               body of the particle-triangle force FI_pt[b][p][6]
 *   FB.%05d.bt  FORCE::BodysetForceOutput (particle.cpp:548-610): per (body,
 *               particle), body-major, 19 x "%16.12e" separated by one space
 *   At.%08d.bt  BODYSET::StateOutput (multibody.cpp:173-228): per body
 *               "%26.17e" x (time, centre[3], velocity[3], angular
 *               velocity[3], quaternion[4], load[6], impulse[6]), no
 *               separators (iostream setw(26) scientific, precision 17)
 *
 * Every function returns 0 on success, -1 when the file cannot be created
 * and -2 when writing fails.  Arrays are row-major (particle-major).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static int finish(FILE* f, int ok)
{
    if (fclose(f) != 0) ok = 0;
    return ok ? 0 : -2;
}

int dem_write_particle_state(const char* path, int n, const int32_t* id, const int32_t* status,
                             const double* mass, const double* radius, const double* pos, const double* vel,
                             const double* quat, const double* angvel)
{
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    int ok = 1;
    for (int j = 0; j < n && ok; ++j) {
        const double* p = pos + 3 * (size_t)j;
        const double* v = vel + 3 * (size_t)j;
        const double* q = quat + 4 * (size_t)j;
        const double* w = angvel + 3 * (size_t)j;
        ok = fprintf(f,
                     "%*d %*d %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e %*.*e "
                     "%*.*e %*.*e\n",
                     10, id[j], 10, status[j], 16, 17, mass[j], 16, 17, radius[j], 16, 17, p[0], 16, 17, p[1],
                     16, 17, p[2], 16, 17, v[0], 16, 17, v[1], 16, 17, v[2], 16, 17, q[0], 16, 17, q[1], 16, 17,
                     q[2], 16, 17, q[3], 16, 17, w[0], 16, 17, w[1], 16, 17, w[2]) > 0;
    }
    return finish(f, ok);
}

int dem_write_particle_force(const char* path, int n, int nbody, const double* contact, const double* wall,
                             const double* field, const double* total, const double* body_particle)
{
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    int ok = 1;
    for (int j = 0; j < n && ok; ++j) {
        const double* c = contact + 6 * (size_t)j;
        const double* w = wall + 6 * (size_t)j;
        const double* g = field + 3 * (size_t)j;
        const double* t = total + 6 * (size_t)j;
        for (int k = 0; k < 6 && ok; ++k) ok = fprintf(f, k ? " %*.*e" : "%*.*e", 16, 12, c[k]) > 0;
        for (int k = 0; k < 6 && ok; ++k) ok = fprintf(f, " %*.*e", 16, 12, w[k]) > 0;
        for (int k = 0; k < 3 && ok; ++k) ok = fprintf(f, " %*.*e", 16, 12, g[k]) > 0;
        for (int k = 0; k < 6 && ok; ++k) ok = fprintf(f, " %*.*e", 16, 12, t[k]) > 0;
        for (int b = 0; b < nbody && ok; ++b) {
            const double* q = body_particle + ((size_t)b * n + j) * 6;
            for (int k = 0; k < 6 && ok; ++k) ok = fprintf(f, " %*.*e", 16, 12, q[k]) > 0;
        }
        if (ok) ok = fputc('\n', f) != EOF;
    }
    return finish(f, ok);
}

int dem_write_body_force(const char* path, int n, int nbody, const double* detail)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int ok = 1;
    for (long index = 0; index < (long)nbody * n && ok; ++index) {
        const double* d = detail + (size_t)index * 19;
        for (int k = 0; k < 19 && ok; ++k) ok = fprintf(f, k < 18 ? "%16.12e " : "%16.12e", d[k]) > 0;
        if (ok) ok = fputc('\n', f) != EOF;
    }
    return finish(f, ok);
}

int dem_write_body_state(const char* path, int nbody, double time, const double* center, const double* vel,
                         const double* angvel, const double* orien, const double* load, const double* impulse)
{
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    int ok = 1;
    for (int s = 0; s < nbody && ok; ++s) {
        double row[26];
        row[0] = time;
        for (int k = 0; k < 3; ++k) {
            row[1 + k] = center[3 * s + k];
            row[4 + k] = vel[3 * s + k];
            row[7 + k] = angvel[3 * s + k];
        }
        for (int k = 0; k < 4; ++k) row[10 + k] = orien[4 * s + k];
        for (int k = 0; k < 6; ++k) row[14 + k] = load[6 * s + k];
        const int cols = impulse ? 26 : 20;
        if (impulse)
            for (int k = 0; k < 6; ++k) row[20 + k] = impulse[6 * s + k];
        for (int k = 0; k < cols && ok; ++k) ok = fprintf(f, "%26.17e", row[k]) > 0;
        if (ok) ok = fputc('\n', f) != EOF;
    }
    return finish(f, ok);
}

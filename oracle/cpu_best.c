/*
 * Best-effort CPU force loop — TEST AND BASELINE INFRASTRUCTURE ONLY (loaded
 * by bench.py's cpu_baseline leg, never by the library).
 *
 * Restates ph4's own CPU pair loop, src/amuse_ph4/src/idata.cc:195-236
 * (r2i = 1/(r^2 + eps2 + tiny), ri = sqrt(r2i), mass > tiny test, nn on the
 * unsoftened r^2 excluding r^2 <= tiny), on structure-of-arrays j data,
 * vectorised over j (OpenMP simd) and threaded over i (OpenMP for).  The
 * nearest neighbour is a second, vectorised min pass over the same r^2.
 *
 *   cpu_best_run(nthreads, ni, nj, eps2, ix, iv, jm, jx, jv, acc, jerk, pot, nn)
 *   arrays are [n][3] row-major for positions and velocities; returns seconds.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define CB_EXPORT __attribute__((visibility("default")))
#define TINY 0x1p-52  /* _TINY_, src/amuse_ph4/src/stdinc.h:33 */

CB_EXPORT double cpu_best_run(int nthreads, int ni, int nj, double eps2, const double *ix,
                              const double *iv, const double *jm, const double *jx,
                              const double *jv, double *acc, double *jerk, double *pot, int *nn)
{
    double *sx = malloc(sizeof(double) * 7 * (size_t)nj);
    if (!sx) return -1.0;
    double *sy = sx + nj, *sz = sy + nj, *svx = sz + nj, *svy = svx + nj, *svz = svy + nj,
           *sm = svz + nj;
    for (int j = 0; j < nj; j++) {
        sx[j] = jx[3 * j]; sy[j] = jx[3 * j + 1]; sz[j] = jx[3 * j + 2];
        svx[j] = jv[3 * j]; svy[j] = jv[3 * j + 1]; svz[j] = jv[3 * j + 2];
        sm[j] = jm[j] > TINY ? jm[j] : 0.0;
    }
    const double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int i = 0; i < ni; i++) {
        const double xi = ix[3 * i], yi = ix[3 * i + 1], zi = ix[3 * i + 2];
        const double vxi = iv[3 * i], vyi = iv[3 * i + 1], vzi = iv[3 * i + 2];
        double ax = 0, ay = 0, az = 0, jx_ = 0, jy_ = 0, jz_ = 0, p = 0;
#pragma omp simd reduction(+ : ax, ay, az, jx_, jy_, jz_, p)
        for (int j = 0; j < nj; j++) {
            const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
            const double dvx = svx[j] - vxi, dvy = svy[j] - vyi, dvz = svz[j] - vzi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double xv = dx * dvx + dy * dvy + dz * dvz;
            const double r2i = 1.0 / (r2 + eps2 + TINY);
            const double ri = sqrt(r2i);
            const double mri = sm[j] * ri;
            const double mr3i = mri * r2i;
            const double a3 = -3.0 * xv * r2i;
            p -= r2 > TINY ? mri : 0.0;
            ax += mr3i * dx; ay += mr3i * dy; az += mr3i * dz;
            jx_ += mr3i * (dvx + a3 * dx);
            jy_ += mr3i * (dvy + a3 * dy);
            jz_ += mr3i * (dvz + a3 * dz);
        }
        double best = INFINITY;
#pragma omp simd reduction(min : best)
        for (int j = 0; j < nj; j++) {
            const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double key = (r2 > TINY && sm[j] > 0.0) ? r2 : INFINITY;
            best = key < best ? key : best;
        }
        int arg = -1;
        for (int j = 0; j < nj && best < INFINITY; j++) {
            const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
            if (sm[j] > 0.0 && dx * dx + dy * dy + dz * dz == best) { arg = j; break; }
        }
        acc[3 * i] = ax; acc[3 * i + 1] = ay; acc[3 * i + 2] = az;
        jerk[3 * i] = jx_; jerk[3 * i + 1] = jy_; jerk[3 * i + 2] = jz_;
        pot[i] = p;
        nn[i] = arg;
    }
    const double t1 = omp_get_wtime();
    free(sx);
    return t1 - t0;
}

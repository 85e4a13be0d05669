/*
 * CPU oracle for the GRAPE-6 direct-summation force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline; the product library (amuse_amd/lib/libg6.so)
 * never links, loads or calls it.
 *
 * Every function restates the arithmetic of the reference CPU GRAPE-6
 * emulator lib/g6lib/g6lib.c in our own code:
 *
 *   orc_predict        -> g6lib.c:64-92   (j predictor; `compat` selects the
 *                                          reference's velocity polynomial,
 *                                          which uses dt^3/dt^4 where the
 *                                          Taylor series has dt^2/dt^3)
 *   orc_force          -> g6lib.c:123-186 (pair loop, skip rules, nn, pow())
 *   orc_force_ld       -> same sums in x87 80-bit long double, plus the sums
 *                         of |term| per i used to state conditioning
 *   orc_ngb            -> lib/sapporo_light/dev_evaluate_gravity.cu:60-67
 *                         (r^2 <= h2, self excluded) with g6lib's id == -1
 *                         skip (g6lib.c:138-140)
 *   orc_unit_*         -> g6lib.c:190-441 ABI state machine with an AoS
 *                         j-record (g6lib.c:11-25), per-call predictor and
 *                         pow() triple, but dynamic capacity instead of the
 *                         100000-particle cap (g6lib.c:7)
 *   orc_baseline_run   -> P threads, each an independent orc_unit over a
 *                         disjoint i-range (one emulator per process, as a
 *                         G6 emulator scales); used as the timed CPU baseline.
 *
 * Compile with -ffp-contract=off: the predictor and the r^2 of the
 * nearest-neighbour search must round exactly like the reference (one
 * rounding per IEEE add/mul, no fused multiply-add).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Predictor (g6lib.c:64-92).                                               */
/* ------------------------------------------------------------------------ */
static inline void predict_one(double ti, double tj, const double x[3],
                               const double v[3], const double a2[3],
                               const double j6[3], const double k18[3],
                               int compat, double xp[3], double vp[3])
{
    double dt = ti - tj;
    double dt2 = dt * dt;
    double dt3 = dt2 * dt;
    double dt4 = dt2 * dt2;
    for (int k = 0; k < 3; k++) {
        /* position: x + v dt + a/2 dt^2 + j/6 dt^3 + s/24 dt^4, summed in
         * the reference's order (g6lib.c:78-83) */
        double dx = v[k] * dt;
        dx += a2[k] * dt2;
        dx += j6[k] * dt3;
        dx += k18[k] * 3.0 / 4.0 * dt4;
        xp[k] = dx + x[k];
        double dv;
        if (compat) {
            /* g6lib.c:85-87 as written */
            dv = a2[k] * 2.0 * dt;
            dv += j6[k] * 3.0 * dt3;
            dv += k18[k] * 6.0 * dt4;
        } else {
            /* Taylor series: a dt + j/2 dt^2 + s/6 dt^3 */
            dv = a2[k] * 2.0 * dt;
            dv += j6[k] * 3.0 * dt2;
            dv += k18[k] * 3.0 * dt3;
        }
        vp[k] = dv + v[k];
    }
}

ORC_EXPORT void orc_predict(int nj, double ti, const double *tj,
                            const double *x, const double *v,
                            const double *a2, const double *j6,
                            const double *k18, int compat, double *xp,
                            double *vp)
{
    for (int j = 0; j < nj; j++)
        predict_one(ti, tj[j], x + 3 * j, v + 3 * j, a2 + 3 * j, j6 + 3 * j,
                    k18 + 3 * j, compat, xp + 3 * j, vp + 3 * j);
}

/* ------------------------------------------------------------------------ */
/* Force (g6lib.c:123-186).                                                 */
/* ------------------------------------------------------------------------ */
static inline double sq3(const double r[3])
{
    double s = 0.0;                 /* g6lib.c:94-101: 0 + r0^2 + r1^2 + r2^2 */
    for (int k = 0; k < 3; k++) s += r[k] * r[k];
    return s;
}

ORC_EXPORT void orc_force(int ni, const int *id_i, const double *xi,
                          const double *vi, double eps2, int nj,
                          const int *id_j, const double *m, const double *xp,
                          const double *vp, double *acc, double *jerk,
                          double *pot, int *nn, double *nn_r2)
{
    for (int i = 0; i < ni; i++) {
        double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, p = 0.0, best = 0.0;
        int nearest = -1;
        for (int j = 0; j < nj; j++) {
            if (id_i[i] == id_j[j]) continue;           /* g6lib.c:135-137 */
            if (id_j[j] == -1) continue;                /* g6lib.c:138-140 */
            double r[3], w[3];
            for (int k = 0; k < 3; k++) {
                r[k] = xp[3 * j + k] - xi[3 * i + k];
                w[k] = vp[3 * j + k] - vi[3 * i + k];
            }
            double r2 = sq3(r);
            if (nearest < 0 || r2 < best) {             /* g6lib.c:151-154 */
                best = r2;
                nearest = id_j[j];
            }
            double s = r2 + eps2;
            double s32 = pow(s, 1.5), s52 = pow(s, 2.5), s12 = pow(s, 0.5);
            double mj = m[j];
            double rv = 0.0;
            for (int k = 0; k < 3; k++) rv += r[k] * w[k];
            for (int k = 0; k < 3; k++) {
                double ak = mj * r[k] / s32;
                double jkk = w[k] / s32;
                jkk -= (3 * rv * r[k]) / s52;
                jkk *= mj;
                a[k] += ak;
                jk[k] += jkk;
            }
            p += -mj / s12;
        }
        for (int k = 0; k < 3; k++) {
            acc[3 * i + k] = a[k];
            jerk[3 * i + k] = jk[k];
        }
        pot[i] = p;
        if (nn) nn[i] = nearest;
        if (nn_r2) nn_r2[i] = best;
    }
}

/* Same pair sum in 80-bit long double.  Also returns, per i, the sums of the
 * magnitudes of the terms (|a_ij|, |j_ij|, |phi_ij|), so a test can state an
 * error relative to the conditioning of a cancelling sum. */
ORC_EXPORT void orc_force_ld(int ni, const int *id_i, const double *xi,
                             const double *vi, double eps2, int nj,
                             const int *id_j, const double *m,
                             const double *xp, const double *vp, double *acc,
                             double *jerk, double *pot, double *sacc,
                             double *sjerk, double *spot)
{
    for (int i = 0; i < ni; i++) {
        long double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, p = 0;
        long double sa = 0, sj = 0, sp = 0;
        for (int j = 0; j < nj; j++) {
            if (id_i[i] == id_j[j] || id_j[j] == -1) continue;
            long double r[3], w[3];
            for (int k = 0; k < 3; k++) {
                r[k] = (long double)xp[3 * j + k] - (long double)xi[3 * i + k];
                w[k] = (long double)vp[3 * j + k] - (long double)vi[3 * i + k];
            }
            long double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            long double s = r2 + (long double)eps2;
            long double inv = 1.0L / sqrtl(s);
            long double inv3 = inv * inv * inv;
            long double rv = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
            long double mj = m[j];
            long double tj2 = 0, ta2 = 0;
            for (int k = 0; k < 3; k++) {
                long double ak = mj * r[k] * inv3;
                long double jkk = mj * (w[k] * inv3 - 3.0L * rv * r[k] * inv3 * inv * inv);
                a[k] += ak;
                jk[k] += jkk;
                ta2 += ak * ak;
                tj2 += jkk * jkk;
            }
            p -= mj * inv;
            sa += sqrtl(ta2);
            sj += sqrtl(tj2);
            sp += fabsl(mj * inv);
        }
        for (int k = 0; k < 3; k++) {
            acc[3 * i + k] = (double)a[k];
            jerk[3 * i + k] = (double)jk[k];
        }
        pot[i] = (double)p;
        if (sacc) sacc[i] = (double)sa;
        if (sjerk) sjerk[i] = (double)sj;
        if (spot) spot[i] = (double)sp;
    }
}

/* Neighbour query: ids of j with unsoftened r^2 <= h2[i], id_j != id_i and
 * id_j != -1, in ascending slot order.  counts[i] is the full count; at most
 * `cap` ids are written to lists[i*cap ...]. */
ORC_EXPORT void orc_ngb(int ni, const int *id_i, const double *xi,
                        const double *h2, int nj, const int *id_j,
                        const double *xp, int cap, int *counts, int *lists)
{
    for (int i = 0; i < ni; i++) {
        int c = 0;
        for (int j = 0; j < nj; j++) {
            if (id_i[i] == id_j[j] || id_j[j] == -1) continue;
            double r[3];
            for (int k = 0; k < 3; k++) r[k] = xp[3 * j + k] - xi[3 * i + k];
            if (sq3(r) <= h2[i]) {
                if (c < cap) lists[(size_t)i * cap + c] = id_j[j];
                c++;
            }
        }
        counts[i] = c;
    }
}

/* ------------------------------------------------------------------------ */
/* ABI state machine with g6lib's record layout (g6lib.c:11-60, 190-441).  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int id;
    double tj, dtj, mass;
    double k18[3], j6[3], a2[3];
    double v[3], x[3];
} orc_jrec;                           /* g6lib.c:11-25 */

typedef struct {
    int id;
    double v[3], x[3];
} orc_jpred;                          /* g6lib.c:43-48 */

typedef struct {
    int id;
    double x[3], v[3], eps2, h2;
    double acc[3], jerk[3], pot, nn_r2;
    int nn;
} orc_irec;                           /* g6lib.c:27-40 */

typedef struct {
    orc_jrec *j;
    orc_jpred *jp;
    orc_irec *i;
    int cap_j, cap_i;
    double ti;
    int nj, ni;
    int compat;
} orc_unit;

ORC_EXPORT orc_unit *orc_unit_open(int cap_j, int cap_i, int compat)
{
    orc_unit *u = (orc_unit *)calloc(1, sizeof(orc_unit));
    if (!u) return NULL;
    u->j = (orc_jrec *)calloc((size_t)cap_j, sizeof(orc_jrec));
    u->jp = (orc_jpred *)calloc((size_t)cap_j, sizeof(orc_jpred));
    u->i = (orc_irec *)calloc((size_t)cap_i, sizeof(orc_irec));
    if (!u->j || !u->jp || !u->i) {
        free(u->j); free(u->jp); free(u->i); free(u);
        return NULL;
    }
    u->cap_j = cap_j;
    u->cap_i = cap_i;
    u->compat = compat;
    for (int k = 0; k < cap_j; k++) u->j[k].id = u->jp[k].id = -1;  /* g6lib.c:451-460 */
    for (int k = 0; k < cap_i; k++) u->i[k].id = -1;
    return u;
}

/* Enlarges the j capacity to at least cap_j (new slots empty, id -1). */
ORC_EXPORT int orc_unit_grow(orc_unit *u, int cap_j)
{
    if (cap_j <= u->cap_j) return 0;
    int cap = u->cap_j > 0 ? u->cap_j : 1;
    while (cap < cap_j) cap *= 2;
    orc_jrec *j = (orc_jrec *)realloc(u->j, sizeof(orc_jrec) * (size_t)cap);
    if (!j) return -1;
    u->j = j;
    orc_jpred *jp = (orc_jpred *)realloc(u->jp, sizeof(orc_jpred) * (size_t)cap);
    if (!jp) return -1;
    u->jp = jp;
    for (int k = u->cap_j; k < cap; k++) {
        memset(&u->j[k], 0, sizeof(orc_jrec));
        memset(&u->jp[k], 0, sizeof(orc_jpred));
        u->j[k].id = u->jp[k].id = -1;
    }
    u->cap_j = cap;
    return 0;
}

ORC_EXPORT void orc_unit_close(orc_unit *u)
{
    if (!u) return;
    free(u->j); free(u->jp); free(u->i); free(u);
}

ORC_EXPORT void orc_unit_set_ti(orc_unit *u, double ti) { u->ti = ti; }

ORC_EXPORT int orc_unit_set_j_particle(orc_unit *u, int address, int index,
                                       double tj, double dtj, double mass,
                                       const double k18[3], const double j6[3],
                                       const double a2[3], const double v[3],
                                       const double x[3])
{
    if (address < 0 || address >= u->cap_j) return -1;
    orc_jrec *r = &u->j[address];
    r->id = index;
    r->tj = tj;
    r->dtj = dtj;
    r->mass = mass;
    for (int k = 0; k < 3; k++) {
        r->k18[k] = k18[k];
        r->j6[k] = j6[k];
        r->a2[k] = a2[k];
        r->v[k] = v[k];
        r->x[k] = x[k];
    }
    return 0;
}

ORC_EXPORT int orc_unit_firsthalf(orc_unit *u, int nj, int ni,
                                  const int *index, const double *xi,
                                  const double *vi, double eps2,
                                  const double *h2)
{
    if (ni < 0 || ni > u->cap_i || nj < 0 || nj > u->cap_j) return -1;
    u->ni = ni;
    u->nj = nj;
    for (int i = 0; i < ni; i++) {                   /* g6lib.c:346-361 */
        orc_irec *p = &u->i[i];
        p->id = index[i];
        for (int k = 0; k < 3; k++) {
            p->v[k] = vi[3 * i + k];
            p->x[k] = xi[3 * i + k];
        }
        p->eps2 = eps2;
        p->h2 = h2 ? h2[i] : 0.0;
    }
    return 0;
}

/* g6calc_lasthalf2_ (g6lib.c:411-441): predict all j, then the pair loop,
 * then copy out; nn may be NULL (g6calc_lasthalf_). */
ORC_EXPORT int orc_unit_lasthalf2(orc_unit *u, int *index, double *acc,
                                  double *jerk, double *pot, int *nn)
{
    for (int j = 0; j < u->nj; j++) {
        orc_jrec *r = &u->j[j];
        u->jp[j].id = r->id;
        predict_one(u->ti, r->tj, r->x, r->v, r->a2, r->j6, r->k18, u->compat,
                    u->jp[j].x, u->jp[j].v);
    }
    for (int i = 0; i < u->ni; i++) {
        orc_irec *ip = &u->i[i];
        double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, p = 0, best = 0;
        int nearest = -1;
        for (int j = 0; j < u->nj; j++) {
            const orc_jrec *r = &u->j[j];
            const orc_jpred *q = &u->jp[j];
            if (ip->id == r->id || r->id == -1) continue;
            double d[3], w[3];
            for (int k = 0; k < 3; k++) {
                d[k] = q->x[k] - ip->x[k];
                w[k] = q->v[k] - ip->v[k];
            }
            double r2 = sq3(d);
            if (nearest < 0 || r2 < best) {
                best = r2;
                nearest = r->id;
            }
            double s = r2 + ip->eps2;
            double s32 = pow(s, 1.5), s52 = pow(s, 2.5), s12 = pow(s, 0.5);
            double rv = 0.0;
            for (int k = 0; k < 3; k++) rv += d[k] * w[k];
            for (int k = 0; k < 3; k++) {
                double ak = r->mass * d[k] / s32;
                double jkk = w[k] / s32;
                jkk -= (3 * rv * d[k]) / s52;
                jkk *= r->mass;
                a[k] += ak;
                jk[k] += jkk;
            }
            p += -r->mass / s12;
        }
        for (int k = 0; k < 3; k++) {
            ip->acc[k] = a[k];
            ip->jerk[k] = jk[k];
        }
        ip->pot = p;
        ip->nn = nearest;
        ip->nn_r2 = best;
    }
    for (int i = 0; i < u->ni; i++) {
        const orc_irec *ip = &u->i[i];
        if (index) index[i] = ip->id;
        for (int k = 0; k < 3; k++) {
            acc[3 * i + k] = ip->acc[k];
            jerk[3 * i + k] = ip->jerk[k];
        }
        pot[i] = ip->pot;
        if (nn) nn[i] = ip->nn;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Timed CPU baseline: P threads, each an independent unit holding the full */
/* j-set, evaluating a disjoint slice of the i-set in chunks of npipes.     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nj, i0, i1, npipes, compat;
    double ti, eps2;
    const int *id;
    const double *tj, *m, *x, *v, *a2, *j6, *k18;
    double *acc, *jerk, *pot;
    int *nn;
    int status;
    double seconds;
} orc_job;

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void *orc_job_run(void *arg)
{
    orc_job *jb = (orc_job *)arg;
    orc_unit *u = orc_unit_open(jb->nj, jb->npipes, jb->compat);
    if (!u) {
        jb->status = -1;
        return NULL;
    }
    for (int j = 0; j < jb->nj; j++)
        orc_unit_set_j_particle(u, j, jb->id[j], jb->tj[j], 0.0, jb->m[j],
                                jb->k18 + 3 * j, jb->j6 + 3 * j,
                                jb->a2 + 3 * j, jb->v + 3 * j, jb->x + 3 * j);
    orc_unit_set_ti(u, jb->ti);
    double t0 = now_s();
    int *idx = (int *)malloc(sizeof(int) * (size_t)jb->npipes);
    double *xi = (double *)malloc(sizeof(double) * 3 * (size_t)jb->npipes);
    double *vi = (double *)malloc(sizeof(double) * 3 * (size_t)jb->npipes);
    for (int i = jb->i0; i < jb->i1; i += jb->npipes) {
        int n = jb->i1 - i < jb->npipes ? jb->i1 - i : jb->npipes;
        /* i-particles are the j-particles predicted to ti (the caller
         * predicts its own i-set) */
        for (int q = 0; q < n; q++) {
            int s = i + q;
            idx[q] = jb->id[s];
            double xp[3], vp[3];
            predict_one(jb->ti, jb->tj[s], jb->x + 3 * s, jb->v + 3 * s,
                        jb->a2 + 3 * s, jb->j6 + 3 * s, jb->k18 + 3 * s,
                        jb->compat, xp, vp);
            for (int k = 0; k < 3; k++) {
                xi[3 * q + k] = xp[k];
                vi[3 * q + k] = vp[k];
            }
        }
        orc_unit_firsthalf(u, jb->nj, n, idx, xi, vi, jb->eps2, NULL);
        orc_unit_lasthalf2(u, idx, jb->acc + 3 * i, jb->jerk + 3 * i,
                           jb->pot + i, jb->nn ? jb->nn + i : NULL);
    }
    free(idx); free(xi); free(vi);
    jb->seconds = now_s() - t0;
    orc_unit_close(u);
    jb->status = 0;
    return NULL;
}

/* Evaluates i-slots [0, ni) against j-slots [0, nj) with `nthreads`
 * independent units.  Returns the longest per-thread compute time in seconds
 * (j-set upload excluded, as for the device), negative on error.
 * Per-j arrays are AoS [n][3] like the ABI. */
ORC_EXPORT double orc_baseline_run(int nthreads, int npipes, int compat,
                                   int nj, int ni, double ti, double eps2,
                                   const int *id, const double *tj,
                                   const double *m, const double *x,
                                   const double *v, const double *a2,
                                   const double *j6, const double *k18,
                                   double *acc, double *jerk, double *pot,
                                   int *nn)
{
    if (nthreads < 1) nthreads = 1;
    if (npipes < 1) npipes = 1;
    orc_job *jobs = (orc_job *)calloc((size_t)nthreads, sizeof(orc_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        orc_job *jb = &jobs[t];
        jb->nj = nj;
        jb->i0 = (int)((long long)ni * t / nthreads);
        jb->i1 = (int)((long long)ni * (t + 1) / nthreads);
        jb->npipes = npipes;
        jb->compat = compat;
        jb->ti = ti;
        jb->eps2 = eps2;
        jb->id = id; jb->tj = tj; jb->m = m; jb->x = x; jb->v = v;
        jb->a2 = a2; jb->j6 = j6; jb->k18 = k18;
        jb->acc = acc; jb->jerk = jerk; jb->pot = pot; jb->nn = nn;
        pthread_create(&th[t], NULL, orc_job_run, jb);
    }
    int bad = 0;
    double dt = 0.0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        bad |= jobs[t].status;
        if (jobs[t].seconds > dt) dt = jobs[t].seconds;
    }
    free(jobs);
    free(th);
    return bad ? -1.0 : dt;
}

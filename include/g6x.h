/*
 * Extensions of libg6.so beyond the GRAPE-6 ABI.
 *
 * None of these has a counterpart in the reference; they exist for callers
 * that keep their data on the device (a device-resident block-step driver,
 * the benchmark) and to select the arithmetic:
 *
 *   precision  G6X_FP64   position/velocity differences, r^-1/2 and all sums
 *                         in IEEE double (the plain path);
 *              G6X_MIXED  predictor and dx in double-single, r^-3 kernel and
 *                         per-sub-tile sums in FP32, sub-tile partials folded
 *                         into Kahan-compensated FP32 accumulators.
 *   predictor  0 = Taylor series (default); 1 = the velocity polynomial of
 *              lib/g6lib/g6lib.c:85-87 as written (dt^3, dt^4 where the
 *              series has dt^2, dt^3), for bit-level comparison with it.
 *
 * Environment (read at g6_open): G6_PRECISION=fp64|mixed,
 * G6_PREDICTOR_COMPAT=g6lib, G6_NPIPES=<n>, G6_DEVICE=<id>, G6_EAGER=1
 * (launch the force kernel, with nearest neighbour, inside firsthalf).
 *
 * Error codes returned by every entry point of libg6.so: */
#ifndef AMUSE_AMD_G6X_H
#define AMUSE_AMD_G6X_H

#ifdef __cplusplus
extern "C" {
#endif

#define G6X_OK 0
#define G6X_ERR_NOT_OPEN (-2)     /* call before g6_open */
#define G6X_ERR_ARG (-3)          /* size/address/mode out of range */
#define G6X_ERR_SEQUENCE (-4)     /* lasthalf without a matching firsthalf */
#define G6X_ERR_HIP (-5)          /* HIP runtime failure */
#define G6X_ERR_NOMEM (-6)        /* device or pinned allocation failed */
#define G6X_ERR_NODEVICE (-7)     /* no GPU or bad device id */

#define G6X_FP64 0
#define G6X_MIXED 1

/* Precision of subsequent calcs (G6X_FP64 or G6X_MIXED). */
int g6x_set_precision(int mode);
int g6x_get_precision(void);
/* 0: Taylor predictor; 1: lib/g6lib/g6lib.c:85-87 velocity polynomial. */
int g6x_set_predictor_compat(int on);
/* i-block capacity reported by g6_npipes (1 .. 1<<22). */
int g6x_set_npipes(int npipes);
/* Text of the last error, or "" . */
const char *g6x_last_error(void);

/* Bulk form of g6_set_j_particle for n slots (arrays of n entries; [n][3]
 * for vectors).  Same semantics as n calls in order. */
int g6x_set_j_particles(int n, const int address[], const int index[],
                        const double tj[], const double dtj[],
                        const double mass[], const double k18[][3],
                        const double j6[][3], const double a2[][3],
                        const double v[][3], const double x[][3]);

/* Device-resident calc: the i-block is j-slots [i0, i0+ni) predicted to the
 * current ti (a full-N pass uses i0=0, ni=nj), so nothing crosses PCIe.
 * Results stay on the device in a library-owned buffer; g6x_fetch copies
 * the last result to host arrays (nn may be NULL).  want_nn selects the
 * nearest-neighbour variant. */
int g6x_calc_self(int nj, int i0, int ni, double eps2, int want_nn);
int g6x_fetch(int ni, double acc[][3], double jerk[][3], double pot[],
              int nn[]);
/* Waits for all work of the unit. */
int g6x_sync(void);

/* Kernel timing with HIP events recorded on the unit's own stream around
 * every force, reduce and predictor launch.  g6x_timing(1) resets and
 * enables the accumulators; g6x_get_timing returns milliseconds summed over
 * launches, the number of force launches and the pairs they evaluated. */
int g6x_timing(int enable);
int g6x_get_timing(double *force_ms, double *reduce_ms, double *predict_ms,
                   long long *force_launches, double *pairs);

/* Tile configuration of the force kernel.  njs = lanes sharing one
 * i-particle (1..256, power of two), ipl = i-particles per lane (1 or 2),
 * nchunk = j-chunks (workgroups along j; 0 = automatic), smem = 1 reads the
 * j-set through scalar loads instead of LDS tiles (njs = 1 only).  njs = 0
 * restores the automatic choice (autotuned table, else heuristic). */
int g6x_set_tiles(int njs, int ipl, int nchunk, int smem);
/* Runs the autotuner for (ni, nj) shapes on the current j-store: times every
 * valid tile configuration and keeps the fastest per ni bucket.  Returns the
 * number of configurations timed. */
int g6x_autotune(int nj, int ni_max, double eps2);
/* Writes the tile table to / reads it from a text file. */
int g6x_save_tuning(const char *path);
int g6x_load_tuning(const char *path);

#ifdef __cplusplus
}
#endif
#endif

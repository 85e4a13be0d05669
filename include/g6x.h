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

/* g6x_calc_self without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The same full pass over
 * j-slots [i0, i0+ni), results written to caller-owned DEVICE arrays acc
 * [ni][3], jerk [ni][3], pot [ni] and nn [ni] (ids of the nearest j; NULL:
 * no nearest neighbour), asynchronously on the unit's stream (g6x_sync
 * waits).  Its i-block, result rows and j-chunk partials use a static device
 * array of G6X_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends:
 * ni <= G6X_NOALLOC_MAX_NI, and fewer j-chunks where the partials would not
 * fit.  It calls no hipMalloc/hipFree: nj and i0 + ni must lie within the
 * j-store's current capacity and pending set_j_particle records within its
 * staging buffers (G6X_ERR_ARG / G6X_ERR_NOMEM otherwise).  Calls share the
 * array and run in order on the unit's stream; overlapping calls on one GPU
 * (other threads or streams) would race.  g6x_fetch does not see its
 * results. */
#define G6X_NOALLOC_SCRATCH_BYTES (64u << 20)
#define G6X_NOALLOC_MAX_NI (1 << 19)
int g6x_calc_self_noalloc(int nj, int i0, int ni, double eps2, double acc[][3],
                          double jerk[][3], double pot[], int nn[]);
/* Waits for all work of the unit. */
int g6x_sync(void);
/* Full passes in spatial order (default 1): a g6x_calc_self over every slot
 * (i0 = 0, ni = nj >= 131072) in the mixed path, with every slot holding a
 * distinct id, sorts the slots along a Morton curve of their positions and
 * takes pairs of far-apart i-blocks and j-tiles with one FP32 subtraction
 * per component relative to the i-block's centre (same error bound as the
 * double-single difference).  0: slot order always. */
int g6x_set_spatial(int on);

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

/* ---- Batched ensembles of independent clusters (g6_clusters.hip) --------
 * Cluster c holds counts[c] consecutive particles of m[], x[][3], v[][3].
 * Each cluster is integrated with the shared-step 4th-order Hermite scheme
 * of the hermite code: step = dt_param x collision time (0.03 when
 * dt_param <= 0), rounded down to a power of two dividing the cluster time;
 * all clusters advance together, one step round (predict, force, correct,
 * update) for every active cluster at a time; the force launch is a work
 * queue of (cluster, i-block) items.  Returns NULL on failure (no GPU,
 * a precision other than G6X_FP64 / G6X_MIXED, a count <= 0, more than
 * 2^28 particles, or a coordinate that is not finite or exceeds 1e15 in
 * magnitude). */
void *g6c_create(int nc, const int counts[], const double m[],
                 const double x[][3], const double v[][3], double eps2,
                 double dt_param, int precision);
/* initial forces and steps; clusters run until t_end */
int g6c_start(void *ens, double t_end);
/* step rounds until every cluster has reached t_end or max_rounds (0: no
 * limit); returns the rounds run.  use_graph 0: the step kernels launched
 * one by one, steps_per_graph rounds between checks of the active count;
 * 1: steps_per_graph rounds captured once in a hipGraph and replayed;
 * 2: the whole run in one persistent launch (the four phases separated by
 * grid barriers, the work list rebuilt on the device), equal to mode 0 bit
 * for bit. */
long long g6c_run(void *ens, int steps_per_graph, int use_graph,
                  long long max_rounds);
/* per-cluster total energy from an FP64 force pass at the current state */
int g6c_energy(void *ens, double energy[]);
/* positions, velocities, per-cluster time and step count */
int g6c_state(void *ens, double x[][3], double v[][3], double t[],
              int nsteps[]);
/* forces of the current state, unpredicted */
int g6c_forces(void *ens, int precision, double acc[][3], double jerk[][3],
               double pot[]);
/* lanes per i-particle (1..64, a power of two) for each of 8 size classes
 * (<=16, <=32, ..., <=2048, more) */
int g6c_set_tiles(void *ens, const int njs_per_class[8]);
int g6c_get_tiles(void *ens, int njs_per_class[8]);
/* j per LDS tile (16..256, a power of two) for each size class */
int g6c_set_jtiles(void *ens, const int tj_per_class[8]);
int g6c_get_jtiles(void *ens, int tj_per_class[8]);
/* times every (lanes per i, j per tile) pair per size class present and
 * keeps the fastest; returns the number of shapes timed */
int g6c_autotune(void *ens);
int g6c_set_precision(void *ens, int precision);
double g6c_elapsed_ms(void *ens);
const char *g6c_last_error(void *ens);
void g6c_free(void *ens);

/* ---- All-pairs acc/jerk on caller-owned device arrays and its adjoint
 * (g6_adjoint.hip).  x, v, acc, jerk, gx, gv, ga, gj are [n][3] row-major
 * device arrays, m and gm [n]; stream is a hipStream_t (NULL: default).
 * Forward: acc_i, jerk_i of the softened pair sum (self excluded).
 * Adjoint: given ga = dL/dacc and gj = dL/djerk, writes gx = dL/dx,
 * gv = dL/dv, gm = dL/dm.  Nothing is allocated. */
int g6a_forces_f32(int n, const float *x, const float *v, const float *m,
                   float eps2, float *acc, float *jerk, void *stream);
int g6a_forces_f64(int n, const double *x, const double *v, const double *m,
                   double eps2, double *acc, double *jerk, void *stream);
int g6a_adjoint_f32(int n, const float *x, const float *v, const float *m,
                    const float *ga, const float *gj, float eps2, float *gx,
                    float *gv, float *gm, void *stream);
int g6a_adjoint_f64(int n, const double *x, const double *v, const double *m,
                    const double *ga, const double *gj, double eps2,
                    double *gx, double *gv, double *gm, void *stream);

/* Fixed-step Hermite scheme around the pair sum (src/amuse_hermite/
 * interface.cc:266-281 predictor, :536-556 corrector) on the same [n][3]
 * device arrays, and the transposes of both linear maps for the backward
 * pass: g6a_step_correct_t_* maps dL/dx1, dL/dv1 and the outer dL/da1,
 * dL/dj1 (NULL: zero) to dL/dx, dv, da, dj and the total dL/da1, dL/dj1 that
 * enter g6a_adjoint_*; g6a_step_predict_t_* adds the predictor's part for
 * dL/dxp, dL/dvp into dL/dx, dv, da, dj.  Nothing is allocated. */
int g6a_step_predict_f32(int n, const float *x, const float *v, const float *a,
                         const float *j, double dt, float *xp, float *vp, void *stream);
int g6a_step_predict_f64(int n, const double *x, const double *v, const double *a,
                         const double *j, double dt, double *xp, double *vp, void *stream);
int g6a_step_correct_f32(int n, const float *x, const float *v, const float *a,
                         const float *j, const float *a1, const float *j1, double dt,
                         float *x1, float *v1, void *stream);
int g6a_step_correct_f64(int n, const double *x, const double *v, const double *a,
                         const double *j, const double *a1, const double *j1, double dt,
                         double *x1, double *v1, void *stream);
int g6a_step_correct_t_f32(int n, const float *gx1, const float *gv1, const float *ga1e,
                           const float *gj1e, double dt, float *gx, float *gv, float *ga,
                           float *gj, float *ga1, float *gj1, void *stream);
int g6a_step_correct_t_f64(int n, const double *gx1, const double *gv1, const double *ga1e,
                           const double *gj1e, double dt, double *gx, double *gv,
                           double *ga, double *gj, double *ga1, double *gj1, void *stream);
int g6a_step_predict_t_f32(int n, const float *gxp, const float *gvp, double dt, float *gx,
                           float *gv, float *ga, float *gj, void *stream);
int g6a_step_predict_t_f64(int n, const double *gxp, const double *gvp, double dt,
                           double *gx, double *gv, double *ga, double *gj, void *stream);

/* ---- Device-resident block-step Hermite integrator (g6_blocks.hip): the
 * integration loop of ph4's GPU path (scheduler, i-predictor, force call,
 * Hermite corrector with the Aarseth step; src/amuse_ph4/src/jdata.cc:
 * 752-793, idata.cc:330-514, scheduler.cc:66-140) run on the GPU with no
 * host round trip per block step.  The state stays on the device.
 *
 * g6d_create loads n particles (ids may be NULL: 1..n; m [n], x and v
 * [n][3]), computes the initial forces and steps (ph4 defaults: dt_fac
 * 0.0625, dt_limit 0.03125, dt_limitm 8) and returns a handle (NULL on
 * failure).  ph4_initial_load = 1 keeps acc = jerk = 0 in the j-set until
 * a particle's first step, as ph4 does.
 * g6d_run advances block steps while the system time is below t_end (ph4's
 * evolve_model loop, src/amuse_ph4/interface.cc:673-674: the system time
 * ends >= t_end, past it when t_end is not commensurate with the steps), or
 * for max_blocks (> 0) block steps, replaying `steps_per_graph` captured
 * steps at a time; returns the block steps taken or a negative code, and
 * sets *reached (if not NULL) to 1 when it stopped at t_end, 0 when the
 * block budget ran out first.
 * g6d_energy: kinetic and potential energy at the system time (FP64 pass).
 * g6d_state: x and v predicted to the system time.  g6d_state_raw: x and v
 * as last corrected, at each particle's own time (ph4's get_state,
 * src/amuse_ph4/interface.cc:485-502).  g6d_steps: per-particle
 * time and step.  g6d_counters: block steps, particle steps, system time.
 * g6d_synchronize: every particle behind the system time steps to it (ph4's
 * synchronize_all, src/amuse_ph4/src/jdata.cc:874-900), one block step. */
void *g6d_create(int n, const int ids[], const double m[], const double x[][3],
                 const double v[][3], double eps2, double eta, double dt_fac,
                 double dt_limit, double dt_limitm, int precision,
                 int ph4_initial_load, int device);
long long g6d_run(void *drv, double t_end, long long max_blocks, int steps_per_graph,
                  int *reached);
int g6d_energy(void *drv, double *kinetic, double *potential);
int g6d_state(void *drv, double x[][3], double v[][3]);
int g6d_state_raw(void *drv, double x[][3], double v[][3]);
int g6d_steps(void *drv, double time[], double step[]);
int g6d_counters(void *drv, long long *block_steps, long long *particle_steps,
                 double *system_time);
int g6d_set_precision(void *drv, int precision);
int g6d_synchronize(void *drv);
const char *g6d_last_error(void *drv);
void g6d_free(void *drv);

#ifdef __cplusplus
}
#endif
#endif

/*
 * GRAPE-6 C ABI exported by libg6.so (MI355X / gfx950).
 *
 * Each declaration replaces the reference function of the same name in the
 * CPU emulator lib/g6lib/g6lib.h:8-128 (implementation lib/g6lib/g6lib.c)
 * and in the CUDA library lib/sapporo_light/sapporoG6lib.cpp:5-82; ph4 binds
 * the same names through src/amuse_ph4/src/grape.h:6-120, phiGRAPE through
 * gfortran's trailing-underscore, by-reference convention
 * (src/amuse_phigrape/src/gravity.F:54-111, update_grape.F:30-53).
 *
 * Conventions (same as the reference):
 *   - x[][3] arrays are row-major [n][3] (Fortran x(3,n));
 *   - `index` / ids are caller-chosen ints; -1 marks an empty j-slot and
 *     i-particles whose id equals a j-slot's id do not interact with it;
 *   - `address` is a 0-based j-slot; calc calls use slots [0, nj);
 *   - return 0 on success.  Where the reference exits or returns -1 this
 *     library returns a negative code (see g6x.h) and never exits.
 *
 * The by-value C names are exported as weak symbols: ph4's grape.cc defines
 * the same wrappers strongly and must win when both are linked.
 */
#ifndef AMUSE_AMD_G6LIB_H
#define AMUSE_AMD_G6LIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Fortran / by-pointer set (g6lib.h:8-53; grape.h:8-108) ---------- */

/* Opens the unit on device `*id` (sapporo.cpp:19-41; g6lib.c:194-208). */
int g6_open_(int *id);
/* Releases the unit (g6lib.c:215-222).  Device buffers are cached for a
 * cheap re-open (phiGRAPE reopens per evolve, interface.F:573-597). */
int g6_close_(int *id);
/* i-block capacity per firsthalf/lasthalf pair (g6lib.c:228-231). */
int g6_npipes_(void);
/* Unit scalings; ignored like the reference (g6lib.c:236-248).  The pointee
 * is int in g6lib.h:11 and double in grape.h:17, hence void*. */
int g6_set_tunit_(void *unit);
int g6_set_xunit_(void *unit);
/* Prediction time for the next calc (g6lib.c:254-258). */
int g6_set_ti_(int *id, double *ti);
/* Stores j-slot *address (g6lib.c:275-315): id=*index, time tj, dtj (kept,
 * unused), mass, k18 = snap/18, j6 = jerk/6, a2 = acc/2, v, x. */
int g6_set_j_particle_(int *cluster_id, int *address, int *index,
                       double *tj, double *dtj, double *mass,
                       double k18[3], double j6[3], double a2[3],
                       double v[3], double x[3]);
/* Latches the i-block (g6lib.c:334-362).  aold/j6old/phiold are hardware
 * scaling hints, ignored like the reference. */
void g6calc_firsthalf_(int *cluster_id, int *nj, int *ni, int index[],
                       double xi[][3], double vi[][3], double aold[][3],
                       double j6old[][3], double phiold[], double *eps2,
                       double h2[]);
/* acc, jerk, pot of the latched i-block (g6lib.c:382-410); writes the
 * latched ids back into index[]. */
int g6calc_lasthalf_(int *cluster_id, int *nj, int *ni, int index[],
                     double xi[][3], double vi[][3], double *eps2,
                     double h2[], double acc[][3], double jerk[][3],
                     double pot[]);
/* As above plus nnbindex[i] = id of the nearest j (g6lib.c:411-441). */
int g6calc_lasthalf2_(int *cluster_id, int *nj, int *ni, int index[],
                      double xi[][3], double vi[][3], double *eps2,
                      double h2[], double acc[][3], double jerk[][3],
                      double pot[], int nnbindex[]);
/* Buffer hints; no-ops (g6lib.c:443-450). */
int g6_initialize_jp_buffer_(int *cluster_id, int *buf_size);
int g6_flush_jp_buffer_(int *cluster_id);
/* Marks every j-slot empty (id -1) (g6lib.c:451-463). */
int g6_reset_(int *cluster_id);
int g6_reset_fofpga_(int *cluster_id);
/* Neighbour lists of the last calc (sapporo.cpp:234-272): j ids with
 * unsoftened r^2 <= h2[i].  read returns 1 if some list overflowed. */
int g6_read_neighbour_list_(int *cluster_id);
int g6_get_neighbour_list_(int *cluster_id, int *ipipe, int *maxlength,
                           int *n_neighbours, int neighbour_list[]);

/* ---- C by-value set (g6lib.h:58-128; grape.cc:8-146) --- weak --------- */
int g6_open(int clusterid);
int g6_close(int clusterid);
int g6_npipes(void);
int g6_set_tunit(double newtunit);
int g6_set_xunit(double newxunit);
int g6_set_ti(int clusterid, double ti);
int g6_set_j_particle(int clusterid, int address, int index, double tj,
                      double dtj, double mass, double a2by18[3],
                      double a1by6[3], double aby2[3], double v[3],
                      double x[3]);
void g6calc_firsthalf(int clusterid, int nj, int ni, int index[],
                      double xi[][3], double vi[][3], double fold[][3],
                      double jold[][3], double phiold[], double eps2,
                      double h2[]);
int g6calc_lasthalf(int clusterid, int nj, int ni, int index[],
                    double xi[][3], double vi[][3], double eps2, double h2[],
                    double acc[][3], double jerk[][3], double pot[]);
int g6calc_lasthalf2(int clusterid, int nj, int ni, int index[],
                     double xi[][3], double vi[][3], double eps2,
                     double h2[], double acc[][3], double jerk[][3],
                     double pot[], int nnbindex[]);
int g6_initialize_jp_buffer(int clusterid, int buf_size);
int g6_flush_jp_buffer(int clusterid);
int g6_reset(int clusterid);
int g6_reset_fofpga(int clusterid);
int g6_read_neighbour_list(int clusterid);
int g6_get_neighbour_list(int clusterid, int ipipe, int maxlength,
                          int *n_neighbours, int neighbour_list[]);
void g6_set_neighbour_list_sort_mode(int mode);      /* grape.cc:141 */
int g6_get_neighbour_list_sort_mode(void);           /* g6lib.h:82 */
int g6_set_overflow_flag_test_mode(int aflag, int jflag, int pflag); /* grape.cc:143-146 */
int g6_get_number_of_pipelines(void);                /* g6lib.h:89 */
void g6_reinitialize(int clusterid);                 /* g6lib.h:126 */

/* ---- sapporo_light drop-in extras --------------------------------------- */
/* Autoconf probe of -lsapporo (support/shared/m4/amuse_lib.m4:135-136). */
int get_device_count(void);
/* ph4 debug hooks declared in grape.h:114-119 (dead code in gpu.cc:241,423);
 * weak no-ops. */
void force_j_particle_send(void);
int get_j_part_data(int, int, double *, double *, double *, double *,
                    double *, double *);

#ifdef __cplusplus
}
#endif
#endif

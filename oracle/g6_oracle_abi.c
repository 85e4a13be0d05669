/*
 * The oracle's ABI face — TEST INFRASTRUCTURE ONLY.
 *
 * Exports the Fortran-style GRAPE-6 entry points (lib/g6lib/g6lib.h:8-53)
 * over the oracle's orc_unit (g6_oracle.c), so a driver bound to the ABI at
 * run time (amuse_amd/csrc/hermite4.cpp) can be run on the CPU oracle and
 * compared with the same driver on libg6.so.  Differences from the
 * reference g6lib.c: dynamic capacity, the Taylor velocity predictor unless
 * G6_PREDICTOR_COMPAT=g6lib, npipes = 4096.
 */
#include <stdlib.h>
#include <string.h>

typedef struct orc_unit orc_unit;
orc_unit *orc_unit_open(int cap_j, int cap_i, int compat);
void orc_unit_close(orc_unit *u);
void orc_unit_set_ti(orc_unit *u, double ti);
int orc_unit_set_j_particle(orc_unit *u, int address, int index, double tj, double dtj,
                            double mass, const double k18[3], const double j6[3],
                            const double a2[3], const double v[3], const double x[3]);
int orc_unit_firsthalf(orc_unit *u, int nj, int ni, const int *index, const double *xi,
                       const double *vi, double eps2, const double *h2);
int orc_unit_lasthalf2(orc_unit *u, int *index, double *acc, double *jerk, double *pot,
                       int *nn);
int orc_unit_grow(orc_unit *u, int cap_j);

#define EXPORT __attribute__((visibility("default")))
#define NPIPES 4096

static orc_unit *U = NULL;

EXPORT int g6_open_(int *id)
{
    (void)id;
    if (U) orc_unit_close(U);
    const char *c = getenv("G6_PREDICTOR_COMPAT");
    U = orc_unit_open(1 << 12, NPIPES, c && !strcmp(c, "g6lib"));
    return U ? 0 : -1;
}
EXPORT int g6_close_(int *id)
{
    (void)id;
    orc_unit_close(U);
    U = NULL;
    return 0;
}
EXPORT int g6_npipes_(void) { return NPIPES; }
EXPORT int g6_set_ti_(int *id, double *ti)
{
    (void)id;
    if (!U) return -2;
    orc_unit_set_ti(U, *ti);
    return 0;
}
EXPORT int g6_set_j_particle_(int *cid, int *address, int *index, double *tj, double *dtj,
                              double *mass, double k18[3], double j6[3], double a2[3],
                              double v[3], double x[3])
{
    (void)cid;
    if (!U) return -2;
    if (orc_unit_grow(U, *address + 1)) return -1;
    return orc_unit_set_j_particle(U, *address, *index, *tj, *dtj, *mass, k18, j6, a2, v, x);
}
EXPORT void g6calc_firsthalf_(int *cid, int *nj, int *ni, int index[], double xi[][3],
                              double vi[][3], double aold[][3], double j6old[][3],
                              double phiold[], double *eps2, double h2[])
{
    (void)cid; (void)aold; (void)j6old; (void)phiold;
    if (!U) return;
    if (orc_unit_grow(U, *nj)) return;
    orc_unit_firsthalf(U, *nj, *ni, index, &xi[0][0], &vi[0][0], *eps2, h2);
}
EXPORT int g6calc_lasthalf_(int *cid, int *nj, int *ni, int index[], double xi[][3],
                            double vi[][3], double *eps2, double h2[], double acc[][3],
                            double jerk[][3], double pot[])
{
    (void)cid; (void)nj; (void)ni; (void)xi; (void)vi; (void)eps2; (void)h2;
    if (!U) return -2;
    return orc_unit_lasthalf2(U, index, &acc[0][0], &jerk[0][0], pot, NULL);
}
EXPORT int g6calc_lasthalf2_(int *cid, int *nj, int *ni, int index[], double xi[][3],
                             double vi[][3], double *eps2, double h2[], double acc[][3],
                             double jerk[][3], double pot[], int nnb[])
{
    (void)cid; (void)nj; (void)ni; (void)xi; (void)vi; (void)eps2; (void)h2;
    if (!U) return -2;
    return orc_unit_lasthalf2(U, index, &acc[0][0], &jerk[0][0], pot, nnb);
}

/* Remaining entry points with g6lib's no-op behaviour (g6lib.c:236-248,
 * 443-481); reset empties the unit. */
EXPORT int g6_set_tunit_(void *p) { (void)p; return 0; }
EXPORT int g6_set_xunit_(void *p) { (void)p; return 0; }
EXPORT int g6_initialize_jp_buffer_(int *c, int *b) { (void)c; (void)b; return 0; }
EXPORT int g6_flush_jp_buffer_(int *c) { (void)c; return 0; }
EXPORT int g6_reset_fofpga_(int *c) { (void)c; return 0; }
EXPORT int g6_reset_(int *c)
{
    return g6_open_(c);
}
EXPORT int g6_read_neighbour_list_(int *c) { (void)c; return 0; }
EXPORT int g6_get_neighbour_list_(int *c, int *ipipe, int *maxlength, int *n, int list[])
{
    (void)c; (void)ipipe; (void)maxlength; (void)list;
    *n = 0;
    return 0;
}

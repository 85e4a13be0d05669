// Runs the reference ph4 integrator (src/amuse_ph4/src, built unchanged with
// -DGPU) on whichever GRAPE-6 library it is linked against — TEST
// INFRASTRUCTURE.  The call sequence restates the AMUSE interface
// (src/amuse_ph4/interface.cc:82-115 initialize_code, :317-347
// commit_particles, :614-660 evolve_model, :758-769 energies) without the
// AMUSE worker layer:
//
//   jdata::setup_gpu -> add_particle per body -> initialize_arrays (opens the
//   library, loads every j: gpu.cc:36-92) -> new idata (initial forces) ->
//   force_initial_timestep(0.0625, 0.03125, 8) -> new scheduler ->
//   repeated jdata::advance until system_time >= t_out, energies
//   from jdata::get_kin / get_pot (a full pass through the library) and,
//   when dtmin is not NULL, the smallest particle time step.
//
// Encounter management is off (manage_encounters = 0) so only the Hermite
// block-step path runs.  ph4h_run_sync also synchronises every particle to
// the system time after each output (AMUSE's synchronize_model,
// src/amuse_ph4/interface.cc:713-724, i.e. jdata::synchronize_all), before
// the energies are taken.
#include "idata.h"
#include "jdata.h"
#include "scheduler.h"

extern "C" __attribute__((visibility("default"))) int
ph4h_run_sync(int n, const double *m, const double *x, const double *v, double eps2,
              double eta, int nout, const double *t_out, double *kin, double *pot,
              long long *blocks, long long *steps, double *dtmin, int sync)
{
    jdata *jd = new jdata;
    jd->setup_gpu();
    jd->use_gpu = jd->have_gpu;
    jd->eps2 = eps2;
    jd->eta = eta;
    jd->system_time = 0;
    jd->sync_time = 0;
    jd->manage_encounters = 0;
    for (int j = 0; j < n; j++)
        jd->add_particle(m[j], 0.0, vec(x[3 * j], x[3 * j + 1], x[3 * j + 2]),
                         vec(v[3 * j], v[3 * j + 1], v[3 * j + 2]), j + 1);
    jd->initialize_arrays();
    idata *id = new idata(jd);
    jd->force_initial_timestep(0.0625, 0.03125, 8.0);
    scheduler *s = new scheduler(jd);
    for (int k = 0; k < nout; k++) {
        while (jd->system_time < t_out[k]) jd->advance();
        if (sync) jd->synchronize_all();
        kin[k] = jd->get_kin();
        pot[k] = jd->get_pot();
        blocks[k] = jd->block_steps;
        steps[k] = jd->total_steps;
        if (dtmin) {
            double lo = jd->timestep[0];
            for (int j = 1; j < jd->nj; j++) lo = jd->timestep[j] < lo ? jd->timestep[j] : lo;
            dtmin[k] = lo;
        }
    }
    const int used_library = jd->use_gpu ? 1 : 0;
    delete s;
    delete id;
    delete jd;
    return used_library;
}

extern "C" __attribute__((visibility("default"))) int
ph4h_run(int n, const double *m, const double *x, const double *v, double eps2, double eta,
         int nout, const double *t_out, double *kin, double *pot, long long *blocks,
         long long *steps, double *dtmin)
{
    return ph4h_run_sync(n, m, x, v, eps2, eta, nout, t_out, kin, pot, blocks, steps, dtmin, 0);
}

// Block-timestep 4th-order Hermite integrator driving a GRAPE-6 library,
// restating the ph4 GPU code path:
//
//   initialisation     src/amuse_ph4/src/gpu.cc:36-92 (g6_open, load every
//                      j with a2 = acc/2, j6 = jerk/6, k18 = 0), full-N force,
//                      jdata::set_initial_timestep (jdata.cc:503-564)
//   block step         jdata::advance / idata::advance (jdata.cc:752-793,
//                      idata.cc:832-881): next block = particles with the
//                      smallest time + timestep (scheduler.cc:66); predict
//                      the i-set on the host (idata.cc:330-365); forces
//                      through g6_set_ti + firsthalf/lasthalf in chunks of
//                      npipes (gpu.cc:302-407); corrector with the Aarseth
//                      step forced to a power of two commensurate with the
//                      block time (idata.cc:399-514); reload the corrected
//                      particles (gpu.cc:163-230)
//   energies           jdata::get_pot / get_kin (jdata.cc:582-668): a full-N
//                      pass through the library for the potential
//
// The library is bound at run time (dlopen + dlsym of the Fortran-style
// entry points every GRAPE-6 implementation exports), so the same driver
// runs on libg6.so or on any other implementation of the ABI.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "h4_rules.h"

#define H4_API extern "C" __attribute__((visibility("default")))

namespace {

struct G6Table {
    void *handle = nullptr;
    int (*open)(int *) = nullptr;
    int (*close)(int *) = nullptr;
    int (*npipes)() = nullptr;
    int (*set_ti)(int *, double *) = nullptr;
    int (*set_j)(int *, int *, int *, double *, double *, double *, double *, double *,
                 double *, double *, double *) = nullptr;
    void (*firsthalf)(int *, int *, int *, int *, double (*)[3], double (*)[3], double (*)[3],
                      double (*)[3], double *, double *, double *) = nullptr;
    int (*lasthalf)(int *, int *, int *, int *, double (*)[3], double (*)[3], double *, double *,
                    double (*)[3], double (*)[3], double *) = nullptr;
    int (*lasthalf2)(int *, int *, int *, int *, double (*)[3], double (*)[3], double *,
                     double *, double (*)[3], double (*)[3], double *, int *) = nullptr;
    int (*set_precision)(int) = nullptr;       // extension, optional
    int (*set_compat)(int) = nullptr;          // extension, optional
};

template <class F>
bool bind(void *h, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(h, name));
    return f != nullptr;
}

struct Sim {
    G6Table g;
    int cluster = 0;
    int n = 0;
    double eps2 = 0.0, eta = 0.01;
    double dt_fac = 0.0625, dt_limit = 0.03125, dt_limitm = 8.0;
    int npipes = 1;
    std::vector<int> id;
    std::vector<double> mass, time, step, pot;
    std::vector<double> pos, vel, acc, jerk;  // [n][3]
    double system_time = 0.0;
    // ph4 does not reload the j-set after the initial force pass
    // (idata::setup, idata.cc:66-83, has no update_gpu), so until its first
    // step a particle is predicted with acc = jerk = 0; true reproduces
    // that, false reloads every particle with its initial acc and jerk.
    bool ph4_initial_load = true;
    long long block_steps = 0, particle_steps = 0, g6_calls = 0;
    // scratch
    std::vector<int> ilist, iid;
    std::vector<double> ipos, ivel, iacc, ijerk, ipot, h2, dummy;
    std::vector<int> inn;
    std::string err;
};

bool load_table(G6Table &t, const char *path, std::string &err)
{
    t.handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!t.handle) {
        err = dlerror();
        return false;
    }
    bool ok = bind(t.handle, "g6_open_", t.open) && bind(t.handle, "g6_close_", t.close) &&
              bind(t.handle, "g6_npipes_", t.npipes) && bind(t.handle, "g6_set_ti_", t.set_ti) &&
              bind(t.handle, "g6_set_j_particle_", t.set_j) &&
              bind(t.handle, "g6calc_firsthalf_", t.firsthalf) &&
              bind(t.handle, "g6calc_lasthalf_", t.lasthalf) &&
              bind(t.handle, "g6calc_lasthalf2_", t.lasthalf2);
    if (!ok) {
        err = "library lacks a GRAPE-6 entry point";
        return false;
    }
    bind(t.handle, "g6x_set_precision", t.set_precision);
    bind(t.handle, "g6x_set_predictor_compat", t.set_compat);
    return true;
}

using h4::predict_ph4;  // idata.cc:341-360

int load_j(Sim &s, int j)
{
    double a2[3], j6[3], k18[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        a2[k] = s.acc[3 * j + k] / 2;
        j6[k] = s.jerk[3 * j + k] / 6;
    }
    int addr = j, idx = s.id[j];
    double tj = s.time[j], dtj = s.step[j], m = s.mass[j];
    return s.g.set_j(&s.cluster, &addr, &idx, &tj, &dtj, &m, k18, j6, a2, &s.vel[3 * j],
                     &s.pos[3 * j]);
}

// Forces on the listed particles at time t (gpu.cc:365-407): i-set predicted
// on the host, library calls in chunks of npipes.
int forces(Sim &s, double t, const int *list, int ni, bool want_pot_only)
{
    s.ipos.resize(3 * size_t(std::max(ni, 1)));
    s.ivel.resize(3 * size_t(std::max(ni, 1)));
    s.iacc.resize(3 * size_t(std::max(ni, 1)));
    s.ijerk.resize(3 * size_t(std::max(ni, 1)));
    s.ipot.resize(std::max(ni, 1));
    s.iid.resize(std::max(ni, 1));
    for (int q = 0; q < ni; q++) {
        const int j = list[q];
        predict_ph4(&s.pos[3 * j], &s.vel[3 * j], &s.acc[3 * j], &s.jerk[3 * j], t - s.time[j],
                    &s.ipos[3 * q], &s.ivel[3 * q]);
        s.iid[q] = s.id[j];
    }
    double tt = t;
    int rc = s.g.set_ti(&s.cluster, &tt);
    if (rc < 0) return rc;
    int nj = s.n;
    double eps2 = s.eps2;
    for (int i0 = 0; i0 < ni; i0 += s.npipes) {
        int nni = std::min(s.npipes, ni - i0);
        auto xi = reinterpret_cast<double(*)[3]>(&s.ipos[3 * i0]);
        auto vi = reinterpret_cast<double(*)[3]>(&s.ivel[3 * i0]);
        auto dm = reinterpret_cast<double(*)[3]>(s.dummy.data());
        s.g.firsthalf(&s.cluster, &nj, &nni, &s.iid[i0], xi, vi, dm, dm, s.dummy.data(), &eps2,
                      s.h2.data());
        rc = s.g.lasthalf(&s.cluster, &nj, &nni, &s.iid[i0], xi, vi, &eps2, s.h2.data(),
                          reinterpret_cast<double(*)[3]>(&s.iacc[3 * i0]),
                          reinterpret_cast<double(*)[3]>(&s.ijerk[3 * i0]), &s.ipot[i0]);
        if (rc < 0) return rc;
        s.g6_calls++;
    }
    (void)want_pot_only;
    return 0;
}

// jdata::set_initial_timestep (jdata.cc:503-564) with the median limit
void initial_steps(Sim &s)
{
    for (int j = 0; j < s.n; j++)
        s.step[j] = h4::first_step_ph4(&s.acc[3 * j], &s.jerk[3 * j], s.eta, s.dt_fac,
                                       s.dt_limit, s.system_time);
    if (s.dt_limitm > 0 && s.n > 0) {
        std::vector<double> tmp(s.step);
        std::sort(tmp.begin(), tmp.end());
        const double dtmax = s.dt_limitm * tmp[s.n / 2];
        for (int j = 0; j < s.n; j++)
            while (s.step[j] > dtmax) s.step[j] /= 2;
    }
}

// idata::correct (idata.cc:431-512)
void correct(Sim &s, int q, int j, double tnext, const double *oacc, const double *ojerk)
{
    const double newstep = h4::correct_ph4(tnext, s.time[j], s.step[j], s.eta, oacc, ojerk,
                                           &s.iacc[3 * q], &s.ijerk[3 * q], &s.ipos[3 * q],
                                           &s.ivel[3 * q]);
    s.time[j] = tnext;
    s.step[j] = newstep;
}

int block_step(Sim &s)
{
    double tnext = 1e300;
    for (int j = 0; j < s.n; j++) tnext = std::min(tnext, s.time[j] + s.step[j]);
    s.ilist.clear();
    for (int j = 0; j < s.n; j++)
        if (s.time[j] + s.step[j] == tnext) s.ilist.push_back(j);
    const int ni = int(s.ilist.size());
    int rc = forces(s, tnext, s.ilist.data(), ni, false);
    if (rc < 0) return rc;
    for (int q = 0; q < ni; q++) {
        const int j = s.ilist[q];
        double oacc[3], ojerk[3];
        for (int k = 0; k < 3; k++) {
            oacc[k] = s.acc[3 * j + k];
            ojerk[k] = s.jerk[3 * j + k];
        }
        correct(s, q, j, tnext, oacc, ojerk);
        for (int k = 0; k < 3; k++) {
            s.pos[3 * j + k] = s.ipos[3 * q + k];
            s.vel[3 * j + k] = s.ivel[3 * q + k];
            s.acc[3 * j + k] = s.iacc[3 * q + k];
            s.jerk[3 * j + k] = s.ijerk[3 * q + k];
        }
        s.pot[j] = s.ipot[q];
        rc = load_j(s, j);
        if (rc < 0) return rc;
    }
    s.system_time = tnext;
    s.block_steps++;
    s.particle_steps += ni;
    return 0;
}

}  // namespace

H4_API void *h4_new(const char *g6path, int cluster, int npipes, int precision, int compat)
{
    Sim *s = new Sim;
    if (!load_table(s->g, g6path, s->err)) {
        fprintf(stderr, "h4_new: %s\n", s->err.c_str());
        delete s;
        return nullptr;
    }
    s->cluster = cluster;
    if (s->g.open(&s->cluster) < 0) {
        delete s;
        return nullptr;
    }
    int np = s->g.npipes();
    s->npipes = npipes > 0 ? npipes : std::max(np, 1);
    if (s->g.set_precision) s->g.set_precision(precision);
    if (s->g.set_compat) s->g.set_compat(compat);
    return s;
}

// Loads the particles, computes initial forces and time steps (ph4
// commit_particles: interface.cc:317-347, gpu.cc:36-92, jdata.cc:503-580).
H4_API int h4_commit(void *h, int n, const int *ids, const double *m, const double *x,
                     const double *v, double eps2, double eta, double dt_fac, double dt_limit,
                     double dt_limitm)
{
    Sim &s = *static_cast<Sim *>(h);
    s.n = n;
    s.eps2 = eps2;
    s.eta = eta;
    s.dt_fac = dt_fac;
    s.dt_limit = dt_limit;
    s.dt_limitm = dt_limitm;
    s.id.assign(ids, ids + n);
    s.mass.assign(m, m + n);
    s.pos.assign(x, x + 3 * size_t(n));
    s.vel.assign(v, v + 3 * size_t(n));
    s.acc.assign(3 * size_t(n), 0.0);
    s.jerk.assign(3 * size_t(n), 0.0);
    s.pot.assign(n, 0.0);
    s.time.assign(n, 0.0);
    s.step.assign(n, 0.0);
    s.system_time = 0.0;
    s.h2.assign(std::max(s.npipes, 1), eps2);
    s.dummy.assign(3 * size_t(std::max(s.npipes, 1)), 0.0);
    for (int j = 0; j < n; j++) {
        int rc = load_j(s, j);
        if (rc < 0) return rc;
    }
    std::vector<int> all(n);
    for (int j = 0; j < n; j++) all[j] = j;
    int rc = forces(s, 0.0, all.data(), n, false);
    if (rc < 0) return rc;
    for (int j = 0; j < n; j++) {
        for (int k = 0; k < 3; k++) {
            s.acc[3 * j + k] = s.iacc[3 * j + k];
            s.jerk[3 * j + k] = s.ijerk[3 * j + k];
        }
        s.pot[j] = s.ipot[j];
    }
    initial_steps(s);
    if (!s.ph4_initial_load) {
        for (int j = 0; j < n; j++) {
            rc = load_j(s, j);
            if (rc < 0) return rc;
        }
    }
    return 0;
}

// true (default): the j-set keeps acc = jerk = 0 until each particle's first
// step, as ph4 does; false: reload it with the initial acc and jerk.
H4_API int h4_set_ph4_initial_load(void *h, int on)
{
    static_cast<Sim *>(h)->ph4_initial_load = on != 0;
    return 0;
}

// Block steps while the system time is below t_end, as ph4's evolve_model
// loop (src/amuse_ph4/interface.cc:614-674): on return the system time is
// >= t_end, the last block ending on t_end when t_end is commensurate with
// the steps (any dyadic time) and past it otherwise; particles are not
// synchronised.  Returns 0 then, or 1 when max_blocks (> 0) block steps ran
// first.
H4_API int h4_evolve(void *h, double t_end, long long max_blocks)
{
    Sim &s = *static_cast<Sim *>(h);
    for (long long done = 0;; done++) {
        if (!(s.system_time < t_end)) return 0;
        if (max_blocks > 0 && done >= max_blocks) return 1;
        int rc = block_step(s);
        if (rc < 0) return rc;
    }
}

H4_API int h4_steps(void *h, long long nblocks)
{
    Sim &s = *static_cast<Sim *>(h);
    for (long long b = 0; b < nblocks; b++) {
        int rc = block_step(s);
        if (rc < 0) return rc;
    }
    return 0;
}

// Kinetic and potential energy at the current system time (jdata.cc:582-668):
// every particle predicted to system_time, potential from one full pass.
H4_API int h4_energy(void *h, double *kin, double *pot)
{
    Sim &s = *static_cast<Sim *>(h);
    std::vector<int> all(s.n);
    for (int j = 0; j < s.n; j++) all[j] = j;
    int rc = forces(s, s.system_time, all.data(), s.n, true);
    if (rc < 0) return rc;
    double K = 0, U = 0;
    for (int j = 0; j < s.n; j++) {
        const double *vp = &s.ivel[3 * j];
        K += 0.5 * s.mass[j] * (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
        U += 0.5 * s.mass[j] * s.ipot[j];
    }
    *kin = K;
    *pot = U;
    return 0;
}

// State predicted to the current system time.
H4_API int h4_get_state(void *h, double *x, double *v)
{
    Sim &s = *static_cast<Sim *>(h);
    for (int j = 0; j < s.n; j++)
        predict_ph4(&s.pos[3 * j], &s.vel[3 * j], &s.acc[3 * j], &s.jerk[3 * j],
                    s.system_time - s.time[j], &x[3 * j], &v[3 * j]);
    return 0;
}

H4_API int h4_get_steps(void *h, double *time, double *step)
{
    Sim &s = *static_cast<Sim *>(h);
    std::copy(s.time.begin(), s.time.end(), time);
    std::copy(s.step.begin(), s.step.end(), step);
    return 0;
}

// Arithmetic of the force library for subsequent calls (libg6.so's
// g6x_set_precision); returns -1 when the library has no such extension.
H4_API int h4_set_precision(void *h, int precision)
{
    Sim &s = *static_cast<Sim *>(h);
    if (!s.g.set_precision) return -1;
    return s.g.set_precision(precision);
}

H4_API double h4_time(void *h) { return static_cast<Sim *>(h)->system_time; }

H4_API int h4_counters(void *h, long long *blocks, long long *psteps, long long *calls)
{
    Sim &s = *static_cast<Sim *>(h);
    *blocks = s.block_steps;
    *psteps = s.particle_steps;
    *calls = s.g6_calls;
    return 0;
}

H4_API void h4_free(void *h)
{
    Sim *s = static_cast<Sim *>(h);
    if (!s) return;
    if (s->g.close) s->g.close(&s->cluster);
    // the library handle stays loaded: unloading a HIP runtime user mid-process
    // is not safe
    delete s;
}

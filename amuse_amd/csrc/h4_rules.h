// Arithmetic of ph4's block-step Hermite integrator that runs per particle:
// the i-particle predictor, the Hermite corrector with the Aarseth step
// forced to a power of two commensurate with the block time, and the first
// time step.  Shared by the host driver (hermite4.cpp, built with g++) and
// the device-resident driver (g6_blocks.hip, device code), so both round
// alike: every add and multiply rounds once (no contraction into FMA; the
// pragma covers clang's device pass, g++ on x86-64 has no FMA to form).
//
//   predict_ph4     idata::predict, src/amuse_ph4/src/idata.cc:341-360
//   correct_ph4     idata::correct, idata.cc:431-512
//   first_step_ph4  jdata::set_initial_timestep, jdata.cc:503-564 (before
//                   the median limit, which needs all particles)
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define H4_HD __host__ __device__
#else
#define H4_HD
#endif

#if defined(__clang__)
#define H4_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define H4_NO_CONTRACT
#endif

namespace h4 {

// Floor on a step (never reached with FP64 forces): with force noise above
// the third difference of the forces the rule keeps halving; the floor keeps
// block times advancing and commensurate (all steps are powers of two).
constexpr double kStepFloor = 0x1p-40;

H4_HD inline void predict_ph4(const double *x, const double *v, const double *a,
                              const double *j, double dt, double *xp, double *vp)
{
    H4_NO_CONTRACT
    if (dt == 0.0) {
        for (int k = 0; k < 3; k++) {
            xp[k] = x[k];
            vp[k] = v[k];
        }
        return;
    }
    for (int k = 0; k < 3; k++) {
        xp[k] = x[k] + dt * (v[k] + 0.5 * dt * (a[k] + dt * j[k] / 3));
        vp[k] = v[k] + dt * (a[k] + 0.5 * dt * j[k]);
    }
}

// Corrects the predicted position ip / velocity iv (in place) of a particle
// whose step ran from t_old to tnext, given the old acc/jerk and the new
// ones (ia, ij) at the predicted state; returns the new step.
H4_HD inline double correct_ph4(double tnext, double t_old, double step_old, double eta,
                                const double *oacc, const double *ojerk, const double *ia,
                                const double *ij, double *ip, double *iv)
{
    H4_NO_CONTRACT
    const double dt = tnext - t_old;
    const double dt2 = dt * dt;
    double a2 = 0, j2 = 0, k2 = 0, l2 = 0;
    for (int k = 0; k < 3; k++) {
        const double alpha = -3 * (oacc[k] - ia[k]) - dt * (2 * ojerk[k] + ij[k]);
        const double beta = 2 * (oacc[k] - ia[k]) + dt * (ojerk[k] + ij[k]);
        ip[k] += (alpha / 12 + beta / 20) * dt2;
        iv[k] += (alpha / 3 + beta / 4) * dt;
        a2 += ia[k] * ia[k];
        j2 += (dt * ij[k]) * (dt * ij[k]);
        k2 += (2 * alpha) * (2 * alpha);
        l2 += (6 * beta) * (6 * beta);
    }
    double newstep = eta * dt * std::sqrt((std::sqrt(a2 * k2) + j2) / (std::sqrt(j2 * l2) + k2));
    int e;
    double old2 = step_old / (2 * std::frexp(step_old, &e));
    // bounded: tnext is a sum of powers of two no smaller than the steps
    for (int it = 0; it < 200 && old2 > 0 && std::fmod(tnext, old2) != 0; it++) old2 /= 2;
    if (!(newstep > 0)) newstep = old2;  // degenerate derivative sums
    if (newstep < old2) {
        newstep = old2 / 2;
    } else {
        const double t2 = 2 * old2;
        if (newstep >= t2 && std::fmod(tnext, t2) == 0) newstep = t2;
        else newstep = old2;
    }
    if (newstep < kStepFloor) newstep = kStepFloor;
    return newstep;
}

// First step of one particle: a power of two commensurate with the system
// time and no larger than dt_limit.
H4_HD inline double first_step_ph4(const double *acc, const double *jerk, double eta,
                                   double dt_fac, double dt_limit, double system_time)
{
    H4_NO_CONTRACT
    double a2 = 0, j2 = 0;
    for (int k = 0; k < 3; k++) {
        a2 += acc[k] * acc[k];
        j2 += jerk[k] * jerk[k];
    }
    double first;
    if (eta == 0.0) first = dt_limit;
    else if (a2 == 0.0 || j2 == 0.0) first = dt_fac * eta;
    else first = dt_fac * eta * std::sqrt(a2 / j2);
    if (std::isnan(first)) first = dt_fac * eta;
    int e;
    first /= 2 * std::frexp(first, &e);
    for (int it = 0; it < 200 && first > 0 && std::fmod(system_time, first) != 0; it++)
        first /= 2;
    for (int it = 0; it < 2100 && first > dt_limit; it++) first /= 2;
    return first;
}

}  // namespace h4

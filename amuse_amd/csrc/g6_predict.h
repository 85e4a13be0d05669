// The GRAPE-6 j-predictor as device functions, shared by the library's
// kernels (g6_kernels.hip) and the device-resident driver (g6_blocks.hip).
#pragma once
#include "g6_kernels.h"

namespace g6k {

constexpr float kFarF = 1.0e18f;  // empty-slot position, mixed path

// ------------------------------------------------------------------------
// Predictor (lib/g6lib/g6lib.c:64-92).  Contraction is off so each add and
// multiply rounds once, as in the reference; the nearest-neighbour search
// depends on predicted positions being bit-identical to it.
// ------------------------------------------------------------------------
struct Pred {
    double x[3], v[3];
};

__device__ __forceinline__ Pred predict_slot(double ti, double tj,
                                             const double x[3], const double v[3],
                                             const double a2[3], const double j6[3],
                                             const double k18[3], int compat, bool use_k18)
{
#pragma clang fp contract(off)
    Pred r;
    const double dt = ti - tj;
    const double dt2 = dt * dt;
    const double dt3 = dt2 * dt;
    const double dt4 = dt2 * dt2;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double dx = v[k] * dt;
        dx = dx + a2[k] * dt2;
        dx = dx + j6[k] * dt3;
        if (use_k18) dx = dx + k18[k] * 3.0 / 4.0 * dt4;
        r.x[k] = dx + x[k];
        double dv = a2[k] * 2.0 * dt;
        if (compat) {
            dv = dv + j6[k] * 3.0 * dt3;
            if (use_k18) dv = dv + k18[k] * 6.0 * dt4;
        } else {
            dv = dv + j6[k] * 3.0 * dt2;
            if (use_k18) dv = dv + k18[k] * 3.0 * dt3;
        }
        r.v[k] = dv + v[k];
    }
    return r;
}

__device__ __forceinline__ void write_pred(const JStore &s, int j, int prec, int id, const Pred &p)
{
    if (prec == kF64) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s.px[k][j] = id == -1 ? kFar : p.x[k];
            s.pv[k][j] = id == -1 ? 0.0 : p.v[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double xp = p.x[k];
            const float hi = float(xp);
            const float lo = float(xp - double(hi));
            s.hx[k][j] = id == -1 ? kFarF : hi;
            s.lx[k][j] = id == -1 ? 0.0f : lo;
            s.fv[k][j] = id == -1 ? 0.0f : float(p.v[k]);
        }
    }
}

}  // namespace g6k

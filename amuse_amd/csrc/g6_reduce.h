// Compensated sums and the ordered reduction of per-chunk force partials,
// shared by the force kernels (g6_kernels.hip) and the device-resident block
// step (g6_blocks.hip, whose correction kernel folds the partials itself).
#pragma once

#include "g6_kernels.h"

#include <hip/hip_runtime.h>

namespace g6k {

// TwoSum (Knuth): hi + lo holds the sum without the rounding error of the add.
__device__ __forceinline__ void two_sum_add_d(double &hi, double &lo, double x)
{
    const double s = hi + x;
    const double bp = s - hi;
    const double err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

__device__ __forceinline__ void two_sum_add(float &hi, float &lo, float x)
{
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// Lexicographic (r2, slot) minimum; an empty key has r2 = +inf, slot = -1
// (as unsigned, the largest).
__device__ __forceinline__ void nn_merge(double &best, int &addr, double ob, int oa)
{
    const bool take = (ob < best) || (ob == best && unsigned(oa) < unsigned(addr));
    best = take ? ob : best;
    addr = take ? oa : addr;
}

__device__ __forceinline__ void nn_merge_f(float &best, int &addr, float ob, int oa)
{
    const bool take = (ob < best) || (ob == best && unsigned(oa) < unsigned(addr));
    best = take ? ob : best;
    addr = take ? oa : addr;
}

// ------------------------------------------------------------------------
// Ordered reduction of the per-chunk partials.  Partials are
// [chunk][field][PS], results [7][S] (+ nn).  Two forms: one lane folds all
// chunks of its i in chunk order; or one wave per i, lane L folding chunks
// L, L + 64, ... followed by a fixed butterfly over the lanes (for small
// i-blocks with many chunks, where the serial loop is a latency chain).
// Either way the result does not depend on timing.
// ------------------------------------------------------------------------
__device__ __forceinline__ void fold_chunk_f64(const double *__restrict__ part, int ch, int PS,
                                               int i, double *h, double *l, double &best,
                                               int &addr, int want_nn)
{
    const double *r = part + size_t(ch) * kPartFieldsD * PS;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        two_sum_add_d(h[k], l[k], r[(2 * k) * PS + i]);
        l[k] += r[(2 * k + 1) * PS + i];
    }
    if (want_nn) nn_merge(best, addr, r[14 * PS + i], int(r[15 * PS + i]));
}

__device__ __forceinline__ void fold_chunk_mx(const double *__restrict__ part, int ch, int PS,
                                              int i, float *h, float *l, float &best, int &addr,
                                              int want_nn)
{
    const float2 *r = reinterpret_cast<const float2 *>(part + size_t(ch) * kPartFields * PS);
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const float2 v = r[k * PS + i];
        two_sum_add(h[k], l[k], v.x);
        l[k] += v.y;
    }
    if (want_nn) nn_merge_f(best, addr, r[7 * PS + i].x, __float_as_int(r[8 * PS + i].x));
}

// lane form (first == 0, step == 1) or one lane of the wave form
// (first = lane, step = 64, followed by butterfly_*)
__device__ __forceinline__ void reduce_i_f64(const double *__restrict__ part, int nchunk, int PS,
                                             int i, int first, int step, double *h, double *l,
                                             double &best, int &addr, int want_nn)
{
#pragma unroll
    for (int k = 0; k < 7; k++) { h[k] = 0.0; l[k] = 0.0; }
    best = __builtin_inf();
    addr = -1;
    for (int ch = first; ch < nchunk; ch += step)
        fold_chunk_f64(part, ch, PS, i, h, l, best, addr, want_nn);
}

__device__ __forceinline__ void reduce_i_mx(const double *__restrict__ part, int nchunk, int PS,
                                            int i, int first, int step, float *h, float *l,
                                            float &best, int &addr, int want_nn)
{
#pragma unroll
    for (int k = 0; k < 7; k++) { h[k] = 0.0f; l[k] = 0.0f; }
    best = __builtin_inff();
    addr = -1;
    for (int ch = first; ch < nchunk; ch += step)
        fold_chunk_mx(part, ch, PS, i, h, l, best, addr, want_nn);
}

__device__ __forceinline__ void butterfly_f64(double *h, double *l, double &best, int &addr)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const double oh = __shfl_xor(h[k], off), ol = __shfl_xor(l[k], off);
            two_sum_add_d(h[k], l[k], oh);
            l[k] += ol;
        }
        const double ob = __shfl_xor(best, off);
        const int oa = __shfl_xor(addr, off);
        nn_merge(best, addr, ob, oa);
    }
}

__device__ __forceinline__ void butterfly_mx(float *h, float *l, float &best, int &addr)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const float oh = __shfl_xor(h[k], off), ol = __shfl_xor(l[k], off);
            two_sum_add(h[k], l[k], oh);
            l[k] += ol;
        }
        const float ob = __shfl_xor(best, off);
        const int oa = __shfl_xor(addr, off);
        nn_merge_f(best, addr, ob, oa);
    }
}

}  // namespace g6k

// Spatial order of the j-slots for full passes of the mixed path.
//
// A full pass (every slot as an i-particle) in slot order gives each
// workgroup i-particles scattered over the whole system, so every pair needs
// the double-single difference.  Sorted along a Morton curve of their
// positions, the i-block of a workgroup and most j-tiles are compact and far
// apart, and k_force_mx_sp takes such pairs with one FP32 subtraction per
// component relative to the i-block's centre.  This file builds that order:
// the bounding box of the positions, a 63-bit Morton key per slot (21 bits
// per axis), a radix sort of (key, slot) pairs, the predicted mixed arrays
// gathered in sorted order, and the box of every kTJ-slot tile.
#include "g6_kernels.h"

#include <hipcub/hipcub.hpp>

namespace g6k {

namespace {

// Order-preserving image of a float in the unsigned integers.
__device__ __forceinline__ unsigned ord(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float unord(unsigned u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_box_init(unsigned *box)
{
    const int k = threadIdx.x;
    if (k < 6) box[k] = k < 3 ? 0xffffffffu : 0u;
}

// Box of the (hi parts of the) predicted positions of slots [0, nj).
__global__ void __launch_bounds__(256) k_box(JStore s, int nj, unsigned *box)
{
    unsigned lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const unsigned o = ord(s.hx[k][j]);
            lo[k] = min(lo[k], o);
            hi[k] = max(hi[k], o);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off >= 1; off >>= 1) {
            lo[k] = min(lo[k], unsigned(__shfl_xor(int(lo[k]), off)));
            hi[k] = max(hi[k], unsigned(__shfl_xor(int(hi[k]), off)));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            atomicMin(&box[k], lo[k]);
            atomicMax(&box[3 + k], hi[k]);
        }
    }
}

// 21 bits of x spread to every third bit of a 63-bit word.
__device__ __forceinline__ unsigned long long spread21(unsigned x)
{
    unsigned long long v = x & 0x1fffffu;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void __launch_bounds__(256) k_morton(JStore s, int nj, const unsigned *box,
                                                unsigned long long *key, int *val)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    unsigned long long code = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float lo = unord(box[k]), hi = unord(box[3 + k]);
        const float ext = hi > lo ? hi - lo : 1.0f;
        float u = (s.hx[k][j] - lo) / ext * 2097151.0f;
        u = fminf(fmaxf(u, 0.0f), 2097151.0f);
        code |= spread21(unsigned(u)) << k;
    }
    key[j] = code;
    val[j] = j;
}

// Mixed predicted arrays, id and pid of slot perm[k] to position k.
__global__ void __launch_bounds__(256) k_gather(JStore s, JStore o, const int *__restrict__ perm,
                                                int nj)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nj) return;
    const int j = perm[k];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.hx[c][k] = s.hx[c][j];
        o.lx[c][k] = s.lx[c][j];
        o.fv[c][k] = s.fv[c][j];
    }
    o.fm[k] = s.fm[j];
    o.id[k] = s.id[j];
    o.pid[k] = s.pid[j];
}

// Box (min xyz, max xyz of the hi parts) of each kTJ-slot tile: one
// workgroup of kTJ lanes per tile.
__global__ void __launch_bounds__(kTJ) k_tile_box(JStore o, int nj, float *tbox)
{
    __shared__ float red[6][kTJ / 64];
    const int j = blockIdx.x * kTJ + threadIdx.x;
    float v[6];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float x = j < nj ? o.hx[k][j] : o.hx[k][blockIdx.x * kTJ];
        v[k] = x;
        v[3 + k] = x;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], __shfl_xor(v[k], off));
            v[3 + k] = fmaxf(v[3 + k], __shfl_xor(v[3 + k], off));
        }
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; k++) red[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float r = red[k][0];
        for (int w = 1; w < kTJ / 64; w++) r = k < 3 ? fminf(r, red[k][w]) : fmaxf(r, red[k][w]);
        tbox[6 * blockIdx.x + k] = r;
    }
}

__global__ void __launch_bounds__(256) k_unsort_rows(const double *__restrict__ in,
                                                     const int *__restrict__ in_nn, int S_in,
                                                     const int *__restrict__ perm, int ni,
                                                     double *out, int *out_nn, int S_out, int nn)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int j = perm[i];
#pragma unroll
    for (int k = 0; k < 7; k++) out[k * size_t(S_out) + j] = in[k * size_t(S_in) + i];
    if (nn) out_nn[j] = in_nn[i];
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

}  // namespace

hipError_t spatial_order(const JStore &s, int nj, const JStore &o, int *perm, float *tbox,
                         SpatialBuf &buf, hipStream_t st)
{
    if (nj <= 0) return hipSuccess;
    const size_t n = size_t(nj);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t need = al(n * 8) * 2 + al(n * 4) + 256;
    hipError_t e;
    if (need > buf.bytes) {
        if (buf.base) (void)hipFree(buf.base);
        buf.base = nullptr;
        buf.bytes = 0;
        e = hipMalloc(&buf.base, need);
        if (e != hipSuccess) return e;
        buf.bytes = need;
    }
    char *p = static_cast<char *>(buf.base);
    auto *kin = reinterpret_cast<unsigned long long *>(p);
    auto *kout = reinterpret_cast<unsigned long long *>(p + al(n * 8));
    auto *vin = reinterpret_cast<int *>(p + 2 * al(n * 8));
    auto *box = reinterpret_cast<unsigned *>(p + 2 * al(n * 8) + al(n * 4));
    size_t tmp_need = 0;
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_need, kin, kout, vin, perm, nj, 0, 63,
                                           st);
    if (e != hipSuccess) return e;
    if (tmp_need > buf.tmp_bytes) {
        if (buf.tmp) (void)hipFree(buf.tmp);
        buf.tmp = nullptr;
        buf.tmp_bytes = 0;
        e = hipMalloc(&buf.tmp, tmp_need);
        if (e != hipSuccess) return e;
        buf.tmp_bytes = tmp_need;
    }
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, st, box);
    hipLaunchKernelGGL(k_box, dim3(std::min(nblk(nj, 256), 1024)), dim3(256), 0, st, s, nj, box);
    hipLaunchKernelGGL(k_morton, dim3(nblk(nj, 256)), dim3(256), 0, st, s, nj, box, kin, vin);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t tb = buf.tmp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(buf.tmp, tb, kin, kout, vin, perm, nj, 0, 63, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather, dim3(nblk(nj, 256)), dim3(256), 0, st, s, o, perm, nj);
    hipLaunchKernelGGL(k_tile_box, dim3(nblk(nj, kTJ)), dim3(kTJ), 0, st, o, nj, tbox);
    return hipGetLastError();
}

hipError_t launch_unsort_rows(const double *in, const int *in_nn, int S_in, const int *perm,
                              int ni, double *out, int *out_nn, int S_out, bool nn,
                              hipStream_t st)
{
    if (ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unsort_rows, dim3(nblk(ni, 256)), dim3(256), 0, st, in, in_nn, S_in,
                       perm, ni, out, out_nn, S_out, int(nn));
    return hipGetLastError();
}

}  // namespace g6k

// All-pairs acc/jerk on caller-owned device arrays, and its adjoint.
//
// Forward (the pair sum of lib/g6lib/g6lib.c:165-183, without potential):
//   a_i = sum_j m_j r / s^1.5,  j_i = sum_j m_j (v / s^1.5 - 3 (r.v) r / s^2.5),
//   r = x_j - x_i, v = v_j - v_i, s = r^2 + eps2, j != i.
//
// Adjoint: given gA_i = dL/da_i and gJ_i = dL/dj_i, with y = s^-1/2,
//   Gr(r, v; gA, gJ) = y^3 gA - 3 y^5 (r.v) gJ - 3 y^5 (r.gJ) v
//                      + r [-3 y^5 (r.gA + v.gJ) + 15 y^7 (r.v)(r.gJ)]
//   Gv(r, v; gJ)     = y^3 gJ - 3 y^5 (r.gJ) r
// are the pair's vector-Jacobian products for d/dr and d/dv (both even in
// (r, v) -> (-r, -v)), and for particle k, summing over p != k with
// r = x_p - x_k, v = v_p - v_k:
//   dL/dx_k = m_k sum_p Gr(r, v; gA_p, gJ_p) - sum_p m_p Gr(r, v; gA_k, gJ_k)
//   dL/dv_k = m_k sum_p Gv(r, v; gJ_p)       - sum_p m_p Gv(r, v; gJ_k)
//   dL/dm_k = - sum_p [gA_p . r y^3 + gJ_p . (v y^3 - 3 (r.v) r y^5)]
// (the m_k terms are k acting as the source of p's force, the m_p terms k
// receiving p's force; the sign of r flips between the two views).  Gr and
// Gv are linear in (gA, gJ), so both views of a pair are one Gr and one Gv
// of (m_k gA_p - m_p gA_k, m_k gJ_p - m_p gJ_k).
//
// Arrays are [n][3] row-major (torch's contiguous (n, 3)); T is float or
// double.  Each workgroup covers 256/NJS i-particles, NJS lanes share one
// i and split the j-range (NJS from a workgroup target per kernel); j-tiles go
// through LDS as rows of four values.  Each lane's plain partial sums are
// folded into compensated (TwoSum) sums every ~64 pairs.  No allocation:
// the caller passes every buffer and its stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#define G6A_API extern "C" __attribute__((visibility("default")))

namespace g6a {

constexpr int kWG = 256;
constexpr int kTJ = 256;
// Tiles between folds of a lane's plain partial sums into its compensated
// ones: ~64 pairs per lane, as in the force kernels (NJS lanes share an i,
// so a lane takes kTJ / NJS pairs of a tile; folding after every tile cost
// the FP32 adjoint ~10 instructions per pair at NJS = 32).
template <int NJS>
constexpr int kFoldTiles = (64 * NJS) / kTJ > 1 ? (64 * NJS) / kTJ : 1;

template <class T>
__device__ __forceinline__ T rsq(T s);
template <>
__device__ __forceinline__ float rsq<float>(float s)
{
    return __builtin_amdgcn_rsqf(s);
}
template <>
__device__ __forceinline__ double rsq<double>(double s)
{
    const double y = __builtin_amdgcn_rsq(s);
    const double h = s * y;
    const double e = __builtin_fma(-h, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// One LDS row of four values (ds_read_b128 for float; two for double).
template <class T>
struct Row4 {
    using type = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
};
template <class T>
__device__ __forceinline__ typename Row4<T>::type row4(T a, T b, T c, T d)
{
    typename Row4<T>::type r;
    r.x = a; r.y = b; r.z = c; r.w = d;
    return r;
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T>
__device__ __forceinline__ void two_sum(T &hi, T &lo, T x)
{
    const T s = hi + x;
    const T bp = s - hi;
    lo += (hi - (s - bp)) + (x - bp);
    hi = s;
}

// Sums of the NJS lanes of each i-row.  A workgroup holds kWG / NJS rows,
// row il in lanes t = il + (kWG / NJS) js, js = 0 .. NJS - 1, so a wave
// holds every row (and reads only 64 NJS / kWG distinct j per step):
// butterflies over the lanes of a row within the wave, then the waves'
// totals in wave order through `fold` (LDS no longer read as a j-tile,
// (kWG / 64) (kWG / NJS) Q 2 values).  Threads t < kWG / NJS end with the
// totals of row t; the order of the additions is fixed.
template <class T, int NJS, int Q>
__device__ __forceinline__ void fold_rows(T (&h)[Q], T (&l)[Q], T *fold)
{
    constexpr int RW = kWG / NJS;
    static_assert(RW <= 64 && 64 % RW == 0, "every wave holds every row");
#pragma unroll
    for (int o = 1; o * RW < 64; o <<= 1)
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const T oh = __shfl_xor(h[q], o * RW), ol = __shfl_xor(l[q], o * RW);
            two_sum(h[q], l[q], oh);
            l[q] += ol;
        }
    const int t = threadIdx.x, w = t >> 6, u = t & 63;
    __syncthreads();
    if (u < RW) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            fold[((w * RW + u) * Q + q) * 2] = h[q];
            fold[((w * RW + u) * Q + q) * 2 + 1] = l[q];
        }
    }
    __syncthreads();
    if (t < RW) {
        for (int ww = 1; ww < kWG / 64; ww++) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                two_sum(h[q], l[q], fold[((ww * RW + t) * Q + q) * 2]);
                l[q] += fold[((ww * RW + t) * Q + q) * 2 + 1];
            }
        }
    }
}

// IPL i-particles per lane (rows il + p RW, p < IPL): each j read from LDS
// serves IPL independent pairs, which the scheduler interleaves.
template <class T, int NJS, int IPL>
__global__ void __launch_bounds__(kWG) k_forces(int n, const T *__restrict__ x,
                                                const T *__restrict__ v,
                                                const T *__restrict__ m, T eps2,
                                                T *__restrict__ acc, T *__restrict__ jerk)
{
    using R = typename Row4<T>::type;
    constexpr int RW = kWG / NJS;  // lane rows of the workgroup
    constexpr int Q = 6 * IPL;     // sums per lane: acc3 jerk3 of each of its i
    // the j-tile (x y z m), (vx vy vz -), then the fold of the rows' lanes
    __shared__ union {
        struct {
            R a[kTJ], b[kTJ];
        } tile;
        T fold[(kWG / 64) * RW * Q * 2];
    } sh;
    R *sa = sh.tile.a, *sb = sh.tile.b;
    const int t = threadIdx.x;
    const int js = t / RW, il = t % RW;
    int ic[IPL];
    T xi[IPL], yi[IPL], zi[IPL], vxi[IPL], vyi[IPL], vzi[IPL];
#pragma unroll
    for (int p = 0; p < IPL; p++) {
        const int i = blockIdx.x * RW * IPL + il + p * RW;
        ic[p] = i < n ? i : n - 1;
        xi[p] = x[3 * ic[p]]; yi[p] = x[3 * ic[p] + 1]; zi[p] = x[3 * ic[p] + 2];
        vxi[p] = v[3 * ic[p]]; vyi[p] = v[3 * ic[p] + 1]; vzi[p] = v[3 * ic[p] + 2];
    }
    T h[Q], l[Q], s6[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) { h[k] = 0; l[k] = 0; s6[k] = 0; }
    for (int tb = 0, nt = 0; tb < n; tb += kTJ) {
        const int jn = min(kTJ, n - tb);
        __syncthreads();
        if (t < jn) {
            const int q = tb + t;
            sa[t] = row4<T>(x[3 * q], x[3 * q + 1], x[3 * q + 2], m[q]);
            sb[t] = row4<T>(v[3 * q], v[3 * q + 1], v[3 * q + 2], T(0));
        }
        __syncthreads();
        for (int jj = js; jj < jn; jj += NJS) {
            const R A = sa[jj], B = sb[jj];
#pragma unroll
            for (int p = 0; p < IPL; p++) {
                const T dx = A.x - xi[p], dy = A.y - yi[p], dz = A.z - zi[p];
                const T s = fma_t(dz, dz, fma_t(dy, dy, fma_t(dx, dx, eps2)));
                T y = rsq<T>(s);
                y = (tb + jj == ic[p]) ? T(0) : y;
                const T y2 = y * y, mi3 = A.w * y * y2;
                const T dvx = B.x - vxi[p], dvy = B.y - vyi[p], dvz = B.z - vzi[p];
                const T rv = dx * dvx + dy * dvy + dz * dvz;
                const T c = T(-3) * rv * y2;
                T *sp = s6 + 6 * p;
                sp[0] += mi3 * dx; sp[1] += mi3 * dy; sp[2] += mi3 * dz;
                sp[3] += mi3 * (dvx + c * dx); sp[4] += mi3 * (dvy + c * dy); sp[5] += mi3 * (dvz + c * dz);
            }
        }
        if (++nt == kFoldTiles<NJS> || tb + kTJ >= n) {
            nt = 0;
#pragma unroll
            for (int k = 0; k < Q; k++) { two_sum(h[k], l[k], s6[k]); s6[k] = 0; }
        }
    }
    fold_rows<T, NJS, Q>(h, l, sh.fold);
    if (t < RW) {
#pragma unroll
        for (int p = 0; p < IPL; p++) {
            const int i = blockIdx.x * RW * IPL + t + p * RW;
            if (i >= n) continue;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                acc[3 * i + k] = h[6 * p + k] + l[6 * p + k];
                jerk[3 * i + k] = h[6 * p + 3 + k] + l[6 * p + 3 + k];
            }
        }
    }
}

template <class T, int NJS>
__global__ void __launch_bounds__(kWG) k_adjoint(int n, const T *__restrict__ x,
                                                 const T *__restrict__ v,
                                                 const T *__restrict__ m,
                                                 const T *__restrict__ ga,
                                                 const T *__restrict__ gj, T eps2,
                                                 T *__restrict__ gx, T *__restrict__ gv,
                                                 T *__restrict__ gm)
{
    using R = typename Row4<T>::type;
    constexpr int RW = kWG / NJS;  // k-rows of the workgroup
    // (x y z m), (vx vy vz -), (gA), (gJ) of each j, then the fold of the
    // rows' lanes
    __shared__ union {
        struct {
            R a[kTJ], b[kTJ], c[kTJ], d[kTJ];
        } tile;
        T fold[(kWG / 64) * RW * 9 * 2];
    } sh;
    R *sa = sh.tile.a, *sb = sh.tile.b, *sc = sh.tile.c, *sd = sh.tile.d;
    const int t = threadIdx.x;
    const int js = t / RW, il = t % RW;
    const int k0 = blockIdx.x * RW + il;
    const int kc = k0 < n ? k0 : n - 1;
    T X[3], V[3], GA[3], GJ[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        X[q] = x[3 * kc + q];
        V[q] = v[3 * kc + q];
        GA[q] = ga[3 * kc + q];
        GJ[q] = gj[3 * kc + q];
    }
    const T mk = m[kc];
    // Gr and Gv are linear in their adjoint arguments, so the two views of
    // a pair take one evaluation on the combined adjoint
    //   gA = m_k gA_p - m_p gA_k,   gJ = m_k gJ_p - m_p gJ_k;
    // of the terms that multiply k's own gA_k, gJ_k only two scalar sums
    // remain, T1 = sum_p m_p y^3 and T2 = sum_p m_p (-3 y^5 r.v), applied
    // at the end.  Sums (hi/lo): gx part[3], gv part[3], T1, T2, dm.
    T h[9], l[9];
#pragma unroll
    for (int q = 0; q < 9; q++) { h[q] = 0; l[q] = 0; }
    T s[9];
#pragma unroll
    for (int q = 0; q < 9; q++) s[q] = 0;
    for (int tb = 0, nt = 0; tb < n; tb += kTJ) {
        const int jn = min(kTJ, n - tb);
        __syncthreads();
        if (t < jn) {
            const int p = tb + t;
            sa[t] = row4<T>(x[3 * p], x[3 * p + 1], x[3 * p + 2], m[p]);
            sb[t] = row4<T>(v[3 * p], v[3 * p + 1], v[3 * p + 2], T(0));
            sc[t] = row4<T>(ga[3 * p], ga[3 * p + 1], ga[3 * p + 2], T(0));
            sd[t] = row4<T>(gj[3 * p], gj[3 * p + 1], gj[3 * p + 2], T(0));
        }
        __syncthreads();
        for (int jj = js; jj < jn; jj += NJS) {
            const R A = sa[jj], B = sb[jj], Cg = sc[jj], Dg = sd[jj];
            const T r0 = A.x - X[0], r1 = A.y - X[1], r2 = A.z - X[2];
            const T w0 = B.x - V[0], w1 = B.y - V[1], w2 = B.z - V[2];
            const T ss = fma_t(r2, r2, fma_t(r1, r1, fma_t(r0, r0, eps2)));
            T y = rsq<T>(ss);
            y = (tb + jj == kc) ? T(0) : y;
            const T y2 = y * y, y3 = y * y2, y5 = y3 * y2;
            const T y5m3 = T(-3) * y5, y7x15 = T(15) * (y5 * y2);
            const T rv = r0 * w0 + r1 * w1 + r2 * w2;
            const T mp = A.w;
            // r.gA, v.gJ, r.gJ with p's adjoint inputs and with k's own
            const T rga_p = r0 * Cg.x + r1 * Cg.y + r2 * Cg.z;
            const T vgj_p = w0 * Dg.x + w1 * Dg.y + w2 * Dg.z;
            const T rgj_p = r0 * Dg.x + r1 * Dg.y + r2 * Dg.z;
            const T rga_k = r0 * GA[0] + r1 * GA[1] + r2 * GA[2];
            const T vgj_k = w0 * GJ[0] + w1 * GJ[1] + w2 * GJ[2];
            const T rgj_k = r0 * GJ[0] + r1 * GJ[1] + r2 * GJ[2];
            const T sp = rga_p + vgj_p;
            // the same dots of the combined adjoint
            const T sgc = fma_t(mk, sp, -mp * (rga_k + vgj_k));
            const T rgj = fma_t(mk, rgj_p, -mp * rgj_k);
            const T cj = y5m3 * rv;
            const T cr = fma_t(y5m3, sgc, (y7x15 * rv) * rgj);
            const T cv = y5m3 * rgj;
            const T mky3 = mk * y3, mkcj = mk * cj;
            // Every sum takes its terms one fused multiply-add at a time
            // (s = fma(u, v, s)): the sum is the last operand, so no term
            // needs a separate add.
            s[0] = fma_t(cr, r0, fma_t(cv, w0, fma_t(mkcj, Dg.x, fma_t(mky3, Cg.x, s[0]))));
            s[1] = fma_t(cr, r1, fma_t(cv, w1, fma_t(mkcj, Dg.y, fma_t(mky3, Cg.y, s[1]))));
            s[2] = fma_t(cr, r2, fma_t(cv, w2, fma_t(mkcj, Dg.z, fma_t(mky3, Cg.z, s[2]))));
            s[3] = fma_t(cv, r0, fma_t(mky3, Dg.x, s[3]));
            s[4] = fma_t(cv, r1, fma_t(mky3, Dg.y, s[4]));
            s[5] = fma_t(cv, r2, fma_t(mky3, Dg.z, s[5]));
            s[6] = fma_t(mp, y3, s[6]);
            s[7] = fma_t(mp, cj, s[7]);
            // dm_k: -(gA_p . r y3 + gJ_p . (w y3 - 3 rv r y5))
            s[8] = fma_t(-cj, rgj_p, fma_t(-sp, y3, s[8]));
        }
        if (++nt == kFoldTiles<NJS> || tb + kTJ >= n) {
            nt = 0;
#pragma unroll
            for (int q = 0; q < 9; q++) { two_sum(h[q], l[q], s[q]); s[q] = 0; }
        }
    }
    fold_rows<T, NJS, 9>(h, l, sh.fold);
    if (t < RW && k0 < n) {
        const T t1 = h[6] + l[6], t2 = h[7] + l[7];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            gx[3 * k0 + q] = (h[q] + l[q]) - fma_t(GA[q], t1, GJ[q] * t2);
            gv[3 * k0 + q] = (h[3 + q] + l[3 + q]) - GJ[q] * t1;
        }
        gm[k0] = h[8] + l[8];
    }
}

// Lanes per i: the smallest power of two from 8 that gives `wgs`
// workgroups of kWG / njs i-rows for n particles (`wgs` times the i per lane
// for the forward), at most 64 (one wave per i).  Timed at n = 16384 over
// njs = 8..64: the forward is fastest at 1024 workgroups (FP32 with two i
// per lane: 215 µs, against 221 µs with one i per lane at its best, 2048
// workgroups), the adjoint at 512 (FP64 805 against 908 µs at 2048): more
// lanes per i mean a longer cross-lane fold at the end and fewer rows per
// workgroup, which costs the heavier kernels more than the extra waves
// gain.
inline int lanes_per_i(int n, long long wgs)
{
    int njs = 8;
    while (njs < 64 && (long long)n * njs < wgs * kWG) njs *= 2;
    return njs;
}

// FP32: two i per lane; FP64: one (twice the registers would cost the FP64
// kernel half its waves, and it issues near the FP64 rate already)
template <class T>
constexpr int kForceIPL = sizeof(T) == 4 ? 2 : 1;

template <class T, int NJS>
void launch_forces_t(int n, const T *x, const T *v, const T *m, T eps2, T *acc, T *jerk,
                     hipStream_t st)
{
    constexpr int rows = (kWG / NJS) * kForceIPL<T>;
    const int grid = (n + rows - 1) / rows;
    hipLaunchKernelGGL((k_forces<T, NJS, kForceIPL<T>>), dim3(grid), dim3(kWG), 0, st, n, x, v,
                       m, eps2, acc, jerk);
}

template <class T>
int launch_forces(int n, const T *x, const T *v, const T *m, T eps2, T *acc, T *jerk,
                  hipStream_t st)
{
    if (n <= 0) return 0;
    switch (lanes_per_i(n, 1024ll * kForceIPL<T>)) {
    case 8: launch_forces_t<T, 8>(n, x, v, m, eps2, acc, jerk, st); break;
    case 16: launch_forces_t<T, 16>(n, x, v, m, eps2, acc, jerk, st); break;
    case 32: launch_forces_t<T, 32>(n, x, v, m, eps2, acc, jerk, st); break;
    default: launch_forces_t<T, 64>(n, x, v, m, eps2, acc, jerk, st); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <class T, int NJS>
void launch_adjoint_t(int n, const T *x, const T *v, const T *m, const T *ga, const T *gj,
                      T eps2, T *gx, T *gv, T *gm, hipStream_t st)
{
    const int grid = (n + kWG / NJS - 1) / (kWG / NJS);
    hipLaunchKernelGGL((k_adjoint<T, NJS>), dim3(grid), dim3(kWG), 0, st, n, x, v, m, ga, gj, eps2,
                       gx, gv, gm);
}

template <class T>
int launch_adjoint(int n, const T *x, const T *v, const T *m, const T *ga, const T *gj, T eps2,
                   T *gx, T *gv, T *gm, hipStream_t st)
{
    if (n <= 0) return 0;
    switch (lanes_per_i(n, 512)) {
    case 8: launch_adjoint_t<T, 8>(n, x, v, m, ga, gj, eps2, gx, gv, gm, st); break;
    case 16: launch_adjoint_t<T, 16>(n, x, v, m, ga, gj, eps2, gx, gv, gm, st); break;
    case 32: launch_adjoint_t<T, 32>(n, x, v, m, ga, gj, eps2, gx, gv, gm, st); break;
    default: launch_adjoint_t<T, 64>(n, x, v, m, ga, gj, eps2, gx, gv, gm, st); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

// ---------------------------------------------------------------------------
// The fixed-step Hermite scheme around the pair sum (src/amuse_hermite/
// interface.cc: predictor :266-281, time-symmetric corrector :536-556) as
// elementwise kernels over the 3n components, and the transposes of both
// maps for the backward pass (all linear in their inputs):
//   xp = x + v dt + a dt^2/2 + j dt^3/6,   vp = v + a dt + j dt^2/2
//   v1 = v + (a + a1) dt/2 + (j - j1) dt^2/12
//   x1 = x + (v + v1) dt/2 + (a - a1) dt^2/12
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_step_predict(int m3, const T *__restrict__ x,
                                                      const T *__restrict__ v,
                                                      const T *__restrict__ a,
                                                      const T *__restrict__ j, T dt, T c2, T c3,
                                                      T *__restrict__ xp, T *__restrict__ vp)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m3; k += gridDim.x * blockDim.x) {
        xp[k] = x[k] + v[k] * dt + a[k] * c2 + j[k] * c3;
        vp[k] = v[k] + a[k] * dt + j[k] * c2;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_step_correct(int m3, const T *__restrict__ x,
                                                      const T *__restrict__ v,
                                                      const T *__restrict__ a,
                                                      const T *__restrict__ j,
                                                      const T *__restrict__ a1,
                                                      const T *__restrict__ j1, T h, T q,
                                                      T *__restrict__ x1, T *__restrict__ v1)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m3; k += gridDim.x * blockDim.x) {
        const T vn = v[k] + (a[k] + a1[k]) * h + (j[k] - j1[k]) * q;
        x1[k] = x[k] + (v[k] + vn) * h + (a[k] - a1[k]) * q;
        v1[k] = vn;
    }
}

// Transpose of the corrector (h = dt/2, q = dt^2/12): from dL/dx1, dL/dv1
// and the outer dL/da1, dL/dj1 (NULL: zero) to dL/dx, dv, da, dj and the
// totals dL/da1, dL/dj1 that enter the pair-sum adjoint.
template <class T>
__global__ void __launch_bounds__(256) k_step_correct_t(int m3, const T *__restrict__ gx1,
                                                        const T *__restrict__ gv1,
                                                        const T *__restrict__ ga1e,
                                                        const T *__restrict__ gj1e, T h, T q,
                                                        T *__restrict__ gx, T *__restrict__ gv,
                                                        T *__restrict__ ga, T *__restrict__ gj,
                                                        T *__restrict__ ga1, T *__restrict__ gj1)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m3; k += gridDim.x * blockDim.x) {
        const T bx = gx1[k];
        const T bv = gv1[k] + bx * h;  // dL/dv1 with x1's dependence on v1
        gx[k] = bx;
        gv[k] = bx * h + bv;
        ga[k] = bx * q + bv * h;
        gj[k] = bv * q;
        ga1[k] = (ga1e ? ga1e[k] : T(0)) - bx * q + bv * h;
        gj1[k] = (gj1e ? gj1e[k] : T(0)) - bv * q;
    }
}

// Transpose of the predictor, added into gx, gv, ga, gj.
template <class T>
__global__ void __launch_bounds__(256) k_step_predict_t(int m3, const T *__restrict__ gxp,
                                                        const T *__restrict__ gvp, T dt, T c2,
                                                        T c3, T *__restrict__ gx,
                                                        T *__restrict__ gv, T *__restrict__ ga,
                                                        T *__restrict__ gj)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m3; k += gridDim.x * blockDim.x) {
        const T bx = gxp[k], bv = gvp[k];
        gx[k] += bx;
        gv[k] += bx * dt + bv;
        ga[k] += bx * c2 + bv * dt;
        gj[k] += bx * c3 + bv * c2;
    }
}

inline int elem_grid(int m3) { return std::min((m3 + 255) / 256, 2048); }

template <class T>
int step_predict(int n, const T *x, const T *v, const T *a, const T *j, double dt, T *xp, T *vp,
                 hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_step_predict<T>, dim3(elem_grid(3 * n)), dim3(256), 0, st, 3 * n, x, v,
                       a, j, T(dt), T(dt * dt / 2), T(dt * dt * dt / 6), xp, vp);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <class T>
int step_correct(int n, const T *x, const T *v, const T *a, const T *j, const T *a1,
                 const T *j1, double dt, T *x1, T *v1, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_step_correct<T>, dim3(elem_grid(3 * n)), dim3(256), 0, st, 3 * n, x, v,
                       a, j, a1, j1, T(dt / 2), T(dt * dt / 12), x1, v1);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <class T>
int step_correct_t(int n, const T *gx1, const T *gv1, const T *ga1e, const T *gj1e, double dt,
                   T *gx, T *gv, T *ga, T *gj, T *ga1, T *gj1, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_step_correct_t<T>, dim3(elem_grid(3 * n)), dim3(256), 0, st, 3 * n, gx1,
                       gv1, ga1e, gj1e, T(dt / 2), T(dt * dt / 12), gx, gv, ga, gj, ga1, gj1);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <class T>
int step_predict_t(int n, const T *gxp, const T *gvp, double dt, T *gx, T *gv, T *ga, T *gj,
                   hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_step_predict_t<T>, dim3(elem_grid(3 * n)), dim3(256), 0, st, 3 * n, gxp,
                       gvp, T(dt), T(dt * dt / 2), T(dt * dt * dt / 6), gx, gv, ga, gj);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

}  // namespace g6a

#define G6A_STEP_API(S, T)                                                                      \
    G6A_API int g6a_step_predict_##S(int n, const T *x, const T *v, const T *a, const T *j,      \
                                     double dt, T *xp, T *vp, void *stream)                     \
    {                                                                                          \
        return g6a::step_predict<T>(n, x, v, a, j, dt, xp, vp, static_cast<hipStream_t>(stream)); \
    }                                                                                          \
    G6A_API int g6a_step_correct_##S(int n, const T *x, const T *v, const T *a, const T *j,      \
                                     const T *a1, const T *j1, double dt, T *x1, T *v1,         \
                                     void *stream)                                              \
    {                                                                                          \
        return g6a::step_correct<T>(n, x, v, a, j, a1, j1, dt, x1, v1,                          \
                                    static_cast<hipStream_t>(stream));                          \
    }                                                                                          \
    G6A_API int g6a_step_correct_t_##S(int n, const T *gx1, const T *gv1, const T *ga1e,         \
                                       const T *gj1e, double dt, T *gx, T *gv, T *ga, T *gj,    \
                                       T *ga1, T *gj1, void *stream)                            \
    {                                                                                          \
        return g6a::step_correct_t<T>(n, gx1, gv1, ga1e, gj1e, dt, gx, gv, ga, gj, ga1, gj1,    \
                                      static_cast<hipStream_t>(stream));                        \
    }                                                                                          \
    G6A_API int g6a_step_predict_t_##S(int n, const T *gxp, const T *gvp, double dt, T *gx,     \
                                       T *gv, T *ga, T *gj, void *stream)                       \
    {                                                                                          \
        return g6a::step_predict_t<T>(n, gxp, gvp, dt, gx, gv, ga, gj,                          \
                                      static_cast<hipStream_t>(stream));                        \
    }
G6A_STEP_API(f32, float)
G6A_STEP_API(f64, double)

G6A_API int g6a_forces_f32(int n, const float *x, const float *v, const float *m, float eps2,
                           float *acc, float *jerk, void *stream)
{
    return g6a::launch_forces<float>(n, x, v, m, eps2, acc, jerk, static_cast<hipStream_t>(stream));
}
G6A_API int g6a_forces_f64(int n, const double *x, const double *v, const double *m,
                           double eps2, double *acc, double *jerk, void *stream)
{
    return g6a::launch_forces<double>(n, x, v, m, eps2, acc, jerk, static_cast<hipStream_t>(stream));
}
G6A_API int g6a_adjoint_f32(int n, const float *x, const float *v, const float *m,
                            const float *ga, const float *gj, float eps2, float *gx, float *gv,
                            float *gm, void *stream)
{
    return g6a::launch_adjoint<float>(n, x, v, m, ga, gj, eps2, gx, gv, gm,
                                      static_cast<hipStream_t>(stream));
}
G6A_API int g6a_adjoint_f64(int n, const double *x, const double *v, const double *m,
                            const double *ga, const double *gj, double eps2, double *gx,
                            double *gv, double *gm, void *stream)
{
    return g6a::launch_adjoint<double>(n, x, v, m, ga, gj, eps2, gx, gv, gm,
                                       static_cast<hipStream_t>(stream));
}

// Batched ensembles: thousands of independent small N-body clusters advanced
// together on one GPU.
//
// Scheme per cluster: the shared-timestep Hermite code
// src/amuse_hermite/interface.cc — predictor :266-281, force with the
// collision-time estimate :496-531, time-symmetric corrector :536-556, step
// = dt_param x collision time :618-637 (dt_param = 0.03, :86) — with the step
// rounded down to a power of two that divides the cluster time, so every
// cluster lands exactly on a dyadic t_end:
//
//   predict   x_p = x + v dt + a dt^2/2 + j dt^3/6, v_p = v + a dt + j dt^2/2
//             (FP64), stored again as double-single + FP32 for the force
//   force     one persistent launch for every active cluster: workgroups pull
//             (cluster, i-block) items off a queue sorted by cost; NJS lanes
//             share an i-particle (NJS from the tuned table of the cluster's
//             size class); j-tiles of the cluster staged in LDS; mixed
//             (double-single dx, FP32 r^-3, compensated sums) or FP64; the
//             inverse quartic collision time max(v^4/s^2, m_ij^2 r^2/s^4) is
//             reduced per cluster with an atomic maximum
//   correct   the time-symmetric corrector (FP64)
//   update    t += dt, next step from the collision time, retire clusters
//             that reached t_end
//
// The step criterion uses positions and velocities only: criteria built
// from differences of successive forces (Aarseth's) turn FP32 force noise
// into ever smaller steps, and would make the mixed path unusable here.
//
// The four kernels of a step contain no memset and no host synchronisation,
// so a sequence of steps is captured once into a hipGraph and replayed.
#include "g6_kernels.h"
#include "g6_gridbar.h"
#include "../../include/g6x.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#define G6C_API extern "C" __attribute__((visibility("default")))

namespace g6c {

constexpr int kWG = 256;
constexpr int kTJ = 256;
constexpr int kClasses = 8;  // size classes: <=16, <=32, ..., <=2048 (and above)
constexpr int kFlushC = 64;  // pairs per lane between folds, FP32 sums
constexpr int kFlushD = 32;  // the same for FP64 sums (clusters are not the hot FP64 path)
constexpr double kMaxCoord = 1e15;  // largest |coordinate| g6c_create accepts

struct Dev {
    int nc = 0, np = 0;
    int *off = nullptr, *cnt = nullptr, *pc = nullptr;  // cluster offsets/counts, particle->cluster
    double *t = nullptr, *dt = nullptr;
    unsigned int *cmax = nullptr;  // per cluster: max over pairs of 1/collision-time^4 (float bits)
    int *active = nullptr, *nsteps = nullptr;
    double *x[3] = {}, *v[3] = {}, *a[3] = {}, *j[3] = {}, *m = nullptr;
    double *px[3] = {}, *pv[3] = {};
    float *hx[3] = {}, *lx[3] = {}, *fv[3] = {}, *fm = nullptr;
    double *na[3] = {}, *nj[3] = {}, *npot = nullptr;
    double *energy = nullptr;  // [nc]
    int *item_c = nullptr, *item_i0 = nullptr, *item_njs = nullptr;
    int nitems = 0;
    // particles of the clusters the work items cover (the predictor and
    // corrector run over these only, so retired clusters cost nothing)
    int *plist = nullptr;
    int nlist = 0;
    int *queue = nullptr, *nactive = nullptr;
    double t_end = 1.0, dt_param = 0.03, eps2 = 1e-6, dt_limit = 0.03125, dt_floor = 0x1p-40;
    // j per LDS tile (<= kTJ) and lanes per i of each size class
    int tj_class[kClasses] = {kTJ, kTJ, kTJ, kTJ, kTJ, kTJ, kTJ, kTJ};
    int njs_class[kClasses] = {16, 16, 4, 4, 1, 1, 1, 1};
    // work list of the persistent run (kc_run), built on the device: items
    // (cluster, first row, lanes per row) of the active clusters, clusters
    // taken in the order corder (largest first); ctl = {items, clusters
    // covered, boost, built}
    int *wl_c = nullptr, *wl_i0 = nullptr, *wl_njs = nullptr, *corder = nullptr, *ctl = nullptr;
    int wl_cap = 0;
    unsigned int *bar = nullptr;  // grid barrier of kc_run (g6bar::kWords, zero at each launch)
    int fill_target = 512;  // items a rebuilt list aims at (see build_items)
};

// size class of a cluster of n particles: <=16, <=32, ..., <=2048, more
__host__ __device__ inline int size_class_d(int n)
{
    int k = 0, lim = 16;
    while (n > lim && k < kClasses - 1) { lim *= 2; k++; }
    return k;
}

__device__ __forceinline__ unsigned long long dbits(double x)
{
    return static_cast<unsigned long long>(__double_as_longlong(x));
}

// ---------------------------------------------------------------------------
// predictor (FP64), also resets the work queue of the force kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void predict_particle(const Dev &d, int p, double dt)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double x = d.x[k][p], v = d.v[k][p], a = d.a[k][p], j = d.j[k][p];
        // src/amuse_hermite/interface.cc:276-278
        const double xp = x + v * dt + a * dt * dt / 2 + j * dt * dt * dt / 6;
        const double vp = v + a * dt + j * dt * dt / 2;
        d.px[k][p] = xp;
        d.pv[k][p] = vp;
        const float hi = float(xp);
        d.hx[k][p] = hi;
        d.lx[k][p] = float(xp - double(hi));
        d.fv[k][p] = float(vp);
    }
}

__global__ void __launch_bounds__(256) kc_predict(Dev d, int zero_step)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) *d.queue = 0;
    if (q >= d.nlist) return;
    const int p = d.plist[q];
    const int c = d.pc[p];
    if (!d.active[c]) return;
    predict_particle(d, p, zero_step ? 0.0 : d.dt[c]);
}

// ---------------------------------------------------------------------------
// persistent force kernel
// ---------------------------------------------------------------------------
struct TileF {
    float4 a[kTJ];  // hx hy hz m
    float4 b[kTJ];  // lx ly lz vx
    float4 c[kTJ];  // vy vz - -: the stride of a and b, so one LDS address serves all three
};
struct TileD {
    double4 a[kTJ];  // x y z m
    double4 b[kTJ];  // vx vy vz -
};

__device__ __forceinline__ void two_sum_f(float &hi, float &lo, float x)
{
    const float s = hi + x;
    const float bp = s - hi;
    lo += (hi - (s - bp)) + (x - bp);
    hi = s;
}
__device__ __forceinline__ void two_sum_d(double &hi, double &lo, double x)
{
    const double s = hi + x;
    const double bp = s - hi;
    lo += (hi - (s - bp)) + (x - bp);
    hi = s;
}

union TileU {
    TileF f;
    TileD dd;
};

// max of two floats known to be >= +0 or NaN, on their bits: one integer
// instruction instead of fmaxf's canonicalising pair, and a NaN term wins
// (then the step rule falls back to dt_limit, as the oracle's max does)
__device__ __forceinline__ float umax_f(float a, float b)
{
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// Work item (cluster c, rows i0 .. i0 + kWG/njs - 1 of it, njs lanes per
// row) of one force pass; every thread of the workgroup takes part (tile
// loads and barriers), so the caller's control flow is workgroup-uniform.
template <bool MIXED>
__device__ __forceinline__ void force_item(const Dev &d, TileU &tile, int c, int i0, int njs)
{
    const int t = threadIdx.x;
    {
        const int off = d.off[c], n = d.cnt[c];
        const int js = t % njs;
        const int il = t / njs;
        const int i = i0 + il;  // local index in the cluster
        const bool ok = i < n;
        const int ip = off + (ok ? i : 0);
        // accumulators: hi/lo pairs of acc3 jerk3 pot
        float fh[7], fl[7], fs[7];
        double dh[7], dl[7], ds[7];
#pragma unroll
        for (int k = 0; k < 7; k++) { fh[k] = fl[k] = fs[k] = 0.0f; dh[k] = dl[k] = ds[k] = 0.0; }
        const float eps2f = float(d.eps2);
        float ihx = d.hx[0][ip], ihy = d.hx[1][ip], ihz = d.hx[2][ip];
        float ilx = d.lx[0][ip], ily = d.lx[1][ip], ilz = d.lx[2][ip];
        float ivx = d.fv[0][ip], ivy = d.fv[1][ip], ivz = d.fv[2][ip];
        const double ix = d.px[0][ip], iy = d.px[1][ip], iz = d.px[2][ip];
        const double ivxd = d.pv[0][ip], ivyd = d.pv[1][ip], ivzd = d.pv[2][ip];
        const float imf = d.fm[ip];
        // max over j of 1 / (collision time)^4 = max(v^4/s^2, m_ij^2 r^2/s^4),
        // kept as max(v^2/s) (squared at the end: squaring is monotonic,
        // rounding included) and max(m_ij^2 r^2/s^4)
        float cq1 = 0.0f, cq2 = 0.0f;
        const int tj = d.tj_class[size_class_d(n)];
        const int lg_njs = __ffs(njs) - 1;  // njs is a power of two
        for (int tb = 0; tb < n; tb += tj) {
            const int jn = min(tj, n - tb);
            // every lane runs kn iterations over j = js + k njs; entries jn ..
            // kn njs - 1 of the tile (fewer than njs, at most kTJ in all) are
            // padding: zero mass at 1e17 (1e100 in FP64) on every axis, where
            // each term of the pair comes out an exact zero for any i within
            // ~1e16 of the origin (r^-3 times a zero mass; the collision terms
            // underflow), so the sums and maxima are those of the jn real j
            const int kn = (jn + njs - 1) >> lg_njs;
            __syncthreads();
            if (t < (kn << lg_njs)) {
                const bool real = t < jn;
                const int q = off + tb + (real ? t : 0);
                if (MIXED) {
                    if (real) {
                        tile.f.a[t] = make_float4(d.hx[0][q], d.hx[1][q], d.hx[2][q], d.fm[q]);
                        tile.f.b[t] = make_float4(d.lx[0][q], d.lx[1][q], d.lx[2][q], d.fv[0][q]);
                        tile.f.c[t] = make_float4(d.fv[1][q], d.fv[2][q], 0.0f, 0.0f);
                    } else {
                        tile.f.a[t] = make_float4(1e17f, 1e17f, 1e17f, 0.0f);
                        tile.f.b[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        tile.f.c[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                } else {
                    if (real) {
                        tile.dd.a[t] = make_double4(d.px[0][q], d.px[1][q], d.px[2][q], d.m[q]);
                        tile.dd.b[t] = make_double4(d.pv[0][q], d.pv[1][q], d.pv[2][q], 0.0);
                    } else {
                        tile.dd.a[t] = make_double4(1e100, 1e100, 1e100, 0.0);
                        tile.dd.b[t] = make_double4(0.0, 0.0, 0.0, 0.0);
                    }
                }
            }
            __syncthreads();
            // the lane's own i sits at tile entry i - tb, i.e. iteration ks
            const int rel = i - tb - js;
            const int ks = (rel >= 0 && rel < jn && !(rel & (njs - 1))) ? rel >> lg_njs : -1;
            int cntp = 0;
            for (int k = 0; k < kn; k++) {
                const int jj = js + (k << lg_njs);
                const bool self = k == ks;
                if (MIXED) {
                    const float4 A = tile.f.a[jj], B = tile.f.b[jj];
                    const float4 Cc = tile.f.c[jj];
                    const float dx = (A.x - ihx) + (B.x - ilx);
                    const float dy = (A.y - ihy) + (B.y - ily);
                    const float dz = (A.z - ihz) + (B.z - ilz);
                    const float s = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2f)));
                    float y = __builtin_amdgcn_rsqf(s);
                    y = self ? 0.0f : y;
                    const float y2 = y * y, mi = A.w * y, mi3 = mi * y2;
                    fs[6] -= mi;
                    {
                        // interface.cc:496-516: (s/v^2)^2 and s^4/(m_ij^2 r^2), inverted
                        const float dvx = B.w - ivx, dvy = Cc.x - ivy, dvz = Cc.y - ivz;
                        const float v2 = dvx * dvx + dvy * dvy + dvz * dvz;
                        // both terms are >= +0 (products of squares; s >=
                        // eps2f as every fma adds a square): their maxima
                        // are taken on the bits, as the cluster's atomicMax
                        cq1 = umax_f(cq1, v2 * y2);
                        const float u = (imf + A.w) * y2 * y2;
                        cq2 = umax_f(cq2, u * u * (s - eps2f));
                    }
                    fs[0] = __builtin_fmaf(mi3, dx, fs[0]);
                    fs[1] = __builtin_fmaf(mi3, dy, fs[1]);
                    fs[2] = __builtin_fmaf(mi3, dz, fs[2]);
                    const float dvx = B.w - ivx, dvy = Cc.x - ivy, dvz = Cc.y - ivz;
                    const float rv = __builtin_fmaf(dz, dvz, __builtin_fmaf(dy, dvy, dx * dvx));
                    const float cc = (-3.0f * rv) * y2;
                    fs[3] = __builtin_fmaf(mi3, __builtin_fmaf(cc, dx, dvx), fs[3]);
                    fs[4] = __builtin_fmaf(mi3, __builtin_fmaf(cc, dy, dvy), fs[4]);
                    fs[5] = __builtin_fmaf(mi3, __builtin_fmaf(cc, dz, dvz), fs[5]);
                    if (++cntp == kFlushC) {
                        cntp = 0;
#pragma unroll
                        for (int k = 0; k < 7; k++) { two_sum_f(fh[k], fl[k], fs[k]); fs[k] = 0.0f; }
                    }
                } else {
                    const double4 A = tile.dd.a[jj], B = tile.dd.b[jj];
                    const double dx = A.x - ix, dy = A.y - iy, dz = A.z - iz;
                    const double s = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, d.eps2)));
                    double y = __builtin_amdgcn_rsq(s);
                    const double h = s * y;
                    const double e = __builtin_fma(-h, y, 1.0);
                    y = __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
                    y = self ? 0.0 : y;
                    const double y2 = y * y, mi = A.w * y, mi3 = mi * y2;
                    ds[6] -= mi;
                    {
                        const double dvx = B.x - ivxd, dvy = B.y - ivyd, dvz = B.z - ivzd;
                        const double v2 = dvx * dvx + dvy * dvy + dvz * dvz;
                        const double q1 = (v2 * y2) * (v2 * y2);
                        const double u = (double(imf) + A.w) * y2 * y2;
                        const double q2 = u * u * (s - d.eps2);
                        // both >= +0 or NaN: the maximum on the bits, as
                        // the mixed path and the cluster's atomicMax
                        cq2 = umax_f(cq2, float(fmax(q1, q2)));
                    }
                    ds[0] = __builtin_fma(mi3, dx, ds[0]);
                    ds[1] = __builtin_fma(mi3, dy, ds[1]);
                    ds[2] = __builtin_fma(mi3, dz, ds[2]);
                    const double dvx = B.x - ivxd, dvy = B.y - ivyd, dvz = B.z - ivzd;
                    const double rv = __builtin_fma(dz, dvz, __builtin_fma(dy, dvy, dx * dvx));
                    const double cc = (-3.0 * rv) * y2;
                    ds[3] = __builtin_fma(mi3, __builtin_fma(cc, dx, dvx), ds[3]);
                    ds[4] = __builtin_fma(mi3, __builtin_fma(cc, dy, dvy), ds[4]);
                    ds[5] = __builtin_fma(mi3, __builtin_fma(cc, dz, dvz), ds[5]);
                    if (++cntp == kFlushD) {
                        cntp = 0;
#pragma unroll
                        for (int k = 0; k < 7; k++) { two_sum_d(dh[k], dl[k], ds[k]); ds[k] = 0.0; }
                    }
                }
            }
            if (MIXED) {
#pragma unroll
                for (int k = 0; k < 7; k++) { two_sum_f(fh[k], fl[k], fs[k]); fs[k] = 0.0f; }
            } else {
#pragma unroll
                for (int k = 0; k < 7; k++) { two_sum_d(dh[k], dl[k], ds[k]); ds[k] = 0.0; }
            }
        }
        // combine the njs lanes of each i
        for (int o = 1; o < njs; o <<= 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                if (MIXED) {
                    const float oh = __shfl_xor(fh[k], o), ol = __shfl_xor(fl[k], o);
                    two_sum_f(fh[k], fl[k], oh);
                    fl[k] += ol;
                } else {
                    const double oh = __shfl_xor(dh[k], o), ol = __shfl_xor(dl[k], o);
                    two_sum_d(dh[k], dl[k], oh);
                    dl[k] += ol;
                }
            }
        }
        float cq = umax_f(cq1 * cq1, cq2);
        for (int o = 1; o < njs; o <<= 1) cq = umax_f(cq, __shfl_xor(cq, o));
        if (ok && js == 0) atomicMax(&d.cmax[c], __float_as_uint(cq));
        if (ok && js == 0) {
            double r[7];
#pragma unroll
            for (int k = 0; k < 7; k++) r[k] = MIXED ? double(fh[k]) + double(fl[k]) : dh[k] + dl[k];
            d.na[0][ip] = r[0]; d.na[1][ip] = r[1]; d.na[2][ip] = r[2];
            d.nj[0][ip] = r[3]; d.nj[1][ip] = r[4]; d.nj[2][ip] = r[5];
            d.npot[ip] = r[6];
        }
    }
}

// One force pass over the host-built item list: workgroups pull items off
// the queue (reset by kc_predict) until it is empty.
template <bool MIXED>
__global__ void __launch_bounds__(kWG) kc_force(Dev d)
{
    __shared__ TileU tile;
    __shared__ int s_item;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_item = atomicAdd(d.queue, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= d.nitems) break;
        const int c = d.item_c[it];
        if (!d.active[c]) continue;
        force_item<MIXED>(d, tile, c, d.item_i0[it], d.item_njs[it]);
    }
}

// ---------------------------------------------------------------------------
// time-symmetric corrector (FP64), src/amuse_hermite/interface.cc:536-556
// ---------------------------------------------------------------------------
__device__ __forceinline__ void correct_particle(const Dev &d, int p, double dt)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double oa = d.a[k][p], oj = d.j[k][p], ov = d.v[k][p], ox = d.x[k][p];
        const double na = d.na[k][p], nj = d.nj[k][p];
        const double v1 = ov + (oa + na) * dt / 2 + (oj - nj) * dt * dt / 12;
        d.x[k][p] = ox + (ov + v1) * dt / 2 + (oa - na) * dt * dt / 12;
        d.v[k][p] = v1;
        d.a[k][p] = na;
        d.j[k][p] = nj;
    }
}

__global__ void __launch_bounds__(256) kc_correct(Dev d)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.nlist) return;
    const int p = d.plist[q];
    const int c = d.pc[p];
    if (!d.active[c]) return;
    correct_particle(d, p, d.dt[c]);
}

// first forces of each cluster become its current acc and jerk
__global__ void __launch_bounds__(256) kc_take_forces(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d.a[k][p] = d.na[k][p];
        d.j[k][p] = d.nj[k][p];
    }
}

// Step from the collision time: dt_param x (1/cmax)^(1/4) (interface.cc:
// 618-637), rounded down to a power of two dividing t and not above
// dt_limit; never below dt_floor, so the cluster time always advances.  The
// fourth root is two correctly rounded square roots and a division (not
// pow(), whose library code costs the persistent kernel ~30 registers), the
// same operations as the oracle's step rule.
__device__ __forceinline__ double next_step(const Dev &d, double t, unsigned int cbits)
{
    const float cm = __uint_as_float(cbits);
    double cand = cm > 0.0f ? d.dt_param / sqrt(sqrt(double(cm))) : d.dt_limit;
    if (!(cand > 0.0)) cand = d.dt_floor;
    double s = d.dt_limit;
    for (int it = 0; it < 80 && s > cand && s > d.dt_floor; it++) s *= 0.5;
    for (int it = 0; it < 80 && fmod(t, s) != 0.0 && s > d.dt_floor; it++) s *= 0.5;
    return s < d.dt_floor ? d.dt_floor : s;
}

__device__ __forceinline__ void update_cluster(const Dev &d, int c, int initial);

__global__ void __launch_bounds__(256) kc_update(Dev d, int initial)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nc) return;
    update_cluster(d, c, initial);
}

__device__ __forceinline__ void update_cluster(const Dev &d, int c, int initial)
{
    const unsigned int cb = d.cmax[c];
    d.cmax[c] = 0u;
    if (initial) {
        d.t[c] = 0.0;
        d.active[c] = 1;
        d.dt[c] = next_step(d, 0.0, cb);
        return;
    }
    if (!d.active[c]) return;
    const double tn = d.t[c] + d.dt[c];
    d.t[c] = tn;
    d.nsteps[c] += 1;
    if (tn >= d.t_end) {
        d.active[c] = 0;
        atomicSub(d.nactive, 1);
        return;
    }
    d.dt[c] = next_step(d, tn, cb);
}

// Per-cluster energy from the FP64 potential of the last force pass at the
// current state: E = sum m (v^2/2 + pot/2); one workgroup per cluster, fixed
// reduction order.
__global__ void __launch_bounds__(256) kc_energy(Dev d)
{
    __shared__ double red[256];
    const int c = blockIdx.x;
    const int off = d.off[c], n = d.cnt[c];
    double e = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int p = off + i;
        const double v2 = d.v[0][p] * d.v[0][p] + d.v[1][p] * d.v[1][p] + d.v[2][p] * d.v[2][p];
        e += d.m[p] * (0.5 * v2 + 0.5 * d.npot[p]);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.energy[c] = red[0];
}

__global__ void kc_set_active_count(Dev d)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int n = 0;
        for (int c = 0; c < d.nc; c++) n += d.active[c];
        *d.nactive = n;
    }
}

// ---------------------------------------------------------------------------
// The whole run in one persistent launch (kc_run)
// ---------------------------------------------------------------------------
// A cooperative launch (every workgroup resident) repeats the four phases of
// a step round -- predict, force, correct, update -- with a grid barrier
// after each, until no cluster is active or the round budget is spent; both
// tests read values that no workgroup writes until after the next barrier,
// so every workgroup leaves the loop at the same round.  The work list is
// built on the device by workgroup 0, as build_items does on the host:
// the items of the active clusters (clusters largest first), rebuilt when
// half the covered clusters have retired, with the lanes per i raised
// ("boost") until the list holds fill_target items.

// Loads of values other workgroups wrote before the last grid barrier.
__device__ __forceinline__ int load_shared_int(const int *p)
{
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier (g6_gridbar.h): thread 0 of each workgroup arrives and waits
// with device-scope atomics; the workgroup's own writes are ordered before it
// and its later reads after it by __syncthreads.  The host zeroes d.bar before
// each launch and reports an aborted launch as an error.
struct DeviceBarOps {
    static __device__ void release_fence() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
    static __device__ void acquire_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
    static __device__ unsigned fetch_add(unsigned *p, unsigned v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ void store(unsigned *p, unsigned v)
    {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ void store_release(unsigned *p, unsigned v)
    {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ unsigned poll(unsigned *p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ void pause() { __builtin_amdgcn_s_sleep(2); }
};

// Returns false (in every thread of the workgroup) when the launch was
// aborted.  Not inlined: the persistent kernel has six barriers, and the
// spin loops inlined cost it ~10 registers.
__device__ __noinline__ bool grid_barrier(unsigned int *bar, unsigned int gen)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0)
        s_ok = g6bar::arrive_and_wait<DeviceBarOps>(bar, int(blockIdx.x), int(gridDim.x), gen);
    __syncthreads();
    return s_ok != 0;
}

__device__ __forceinline__ int lanes_of(const Dev &d, int n, int boost)
{
    return min(64, d.njs_class[size_class_d(n)] * boost);
}

// Sum over the workgroup (kWG threads) of v, returned to every thread.
__device__ __forceinline__ int wg_sum(int v, int *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int s = 0;
    for (int w = 0; w < kWG / 64; w++) s += red[w];
    __syncthreads();
    return s;
}

// Workgroup 0: the work list of the clusters active now.  fill = 0: every
// cluster at its tuned lanes per i (boost 1), as at the start of a run.
__device__ void build_work_list(const Dev &d, int fill, int *red)
{
    const int t = threadIdx.x;
    const int per = (d.nc + kWG - 1) / kWG;  // clusters per thread, in corder
    const int c0 = min(d.nc, t * per), c1 = min(d.nc, c0 + per);
    int act = 0;
    for (int k = c0; k < c1; k++) act += d.active[d.corder[k]] ? 1 : 0;
    const int covered = wg_sum(act, red);
    int boost = 1;
    if (fill) {
        const int target = max(1, d.fill_target);
        for (; boost < 64; boost *= 2) {
            int count = 0, room = 0;
            for (int k = c0; k < c1; k++) {
                const int c = d.corder[k];
                if (!d.active[c]) continue;
                const int n = d.cnt[c], njs = lanes_of(d, n, boost);
                const int rows = kWG / njs;
                count += (n + rows - 1) / rows;
                room |= njs < 64 ? 1 : 0;
            }
            count = wg_sum(count, red);
            room = wg_sum(room, red);
            if (count >= target || room == 0) break;
        }
    }
    // items of this thread's clusters, then their offsets (exclusive scan)
    int mine = 0;
    for (int k = c0; k < c1; k++) {
        const int c = d.corder[k];
        if (!d.active[c]) continue;
        const int n = d.cnt[c], rows = kWG / lanes_of(d, n, boost);
        mine += (n + rows - 1) / rows;
    }
    int incl = mine;
    const int lane = t & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    __syncthreads();
    if (lane == 63) red[t >> 6] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < (t >> 6); w++) base += red[w];
    int at = base + incl - mine;
    int total = 0;
    for (int w = 0; w < kWG / 64; w++) total += red[w];
    for (int k = c0; k < c1; k++) {
        const int c = d.corder[k];
        if (!d.active[c]) continue;
        const int n = d.cnt[c], njs = lanes_of(d, n, boost), rows = kWG / njs;
        for (int i0 = 0; i0 < n && at < d.wl_cap; i0 += rows, at++) {
            d.wl_c[at] = c;
            d.wl_i0[at] = i0;
            d.wl_njs[at] = njs;
        }
    }
    if (t == 0) {
        d.ctl[0] = min(total, d.wl_cap);
        d.ctl[1] = covered;
        d.ctl[2] = boost;
        d.ctl[3] = 1;
    }
    __syncthreads();
}

// Rows of work item `it`: the particles of (cluster, i0 .. i0 + kWG/njs - 1).
__device__ __forceinline__ int item_particle(const Dev &d, int it, int t, int &c)
{
    c = d.wl_c[it];
    const int i = d.wl_i0[it] + t;
    if (t >= kWG / d.wl_njs[it] || i >= d.cnt[c]) return -1;
    return d.off[c] + i;
}

template <bool MIXED>
__global__ void __launch_bounds__(kWG) kc_run(Dev d, long long max_rounds, long long *rounds_out)
{
    __shared__ TileU tile;
    __shared__ int s_item;
    __shared__ int red[kWG / 64];
    const int t = threadIdx.x;
    unsigned gen = 0;
    long long r = 0;
    // The list state as every workgroup keeps it in registers: read by all
    // before the barrier, written (by workgroup 0, in a rebuild) only after
    // it, and read again after each rebuild's own barrier.
    int built = load_shared_int(&d.ctl[3]);
    int covered = load_shared_int(&d.ctl[1]);
    bool ok = grid_barrier(d.bar, gen++);
    while (ok) {
        // work list: none yet, or half the covered clusters retired (nactive
        // changes only in the update phase, three barriers below)
        const int nact = load_shared_int(d.nactive);
        if (!built || nact * 2 <= covered) {
            if (blockIdx.x == 0) build_work_list(d, built, red);
            if (!grid_barrier(d.bar, gen++)) break;
            built = 1;
            covered = load_shared_int(&d.ctl[1]);
        }
        const int nwork = load_shared_int(&d.ctl[0]);
        // predict (and reset the force queue)
        if (blockIdx.x == 0 && t == 0) *d.queue = 0;
        for (int it = blockIdx.x; it < nwork; it += gridDim.x) {
            int c;
            const int p = item_particle(d, it, t, c);
            if (p >= 0 && d.active[c]) predict_particle(d, p, d.dt[c]);
        }
        if (!grid_barrier(d.bar, gen++)) break;
        // force: items off the queue
        for (;;) {
            if (t == 0) s_item = atomicAdd(d.queue, 1);
            __syncthreads();
            const int it = s_item;
            __syncthreads();
            if (it >= nwork) break;
            const int c = d.wl_c[it];
            if (!d.active[c]) continue;
            force_item<MIXED>(d, tile, c, d.wl_i0[it], d.wl_njs[it]);
        }
        if (!grid_barrier(d.bar, gen++)) break;
        // correct
        for (int it = blockIdx.x; it < nwork; it += gridDim.x) {
            int c;
            const int p = item_particle(d, it, t, c);
            if (p >= 0 && d.active[c]) correct_particle(d, p, d.dt[c]);
        }
        if (!grid_barrier(d.bar, gen++)) break;
        // update: time, next step, retire
        for (int c = blockIdx.x * kWG + t; c < d.nc; c += gridDim.x * kWG)
            update_cluster(d, c, 0);
        if (!grid_barrier(d.bar, gen++)) break;
        r++;
        if (load_shared_int(d.nactive) <= 0 || r >= max_rounds) break;
    }
    if (blockIdx.x == 0 && t == 0) *rounds_out = r;
}

__global__ void kc_zero_words(unsigned int *w, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        w[k] = 0u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Ensemble {
    Dev d;
    hipStream_t stream = nullptr;
    void *base = nullptr;
    int precision = 1;
    std::vector<int> cnt, off;
    int ib_rows = 0;
    int persistent_wgs = 0;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int graph_prec = -1;
    int items_clusters = 0;  // clusters covered by the current work items
    int *wl_base = nullptr;  // the persistent run's work list and cluster order
    double force_ms = 0.0;
    long long launches = 0;
    std::string err;
};

int size_class(int n) { return size_class_d(n); }

#define CK(expr)                                                       \
    do {                                                               \
        hipError_t e_ = (expr);                                        \
        if (e_ != hipSuccess) {                                        \
            E->err = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return G6X_ERR_HIP;                                        \
        }                                                              \
    } while (0)

// Two timing events, destroyed with the holder.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    bool create() { return hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess; }
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Work items: every (cluster, i-block) pair, i-block = 256/njs rows, sorted
// by descending cost (rows x cluster size) so the queue ends with the cheap
// ones.  `only` (host flags, one per cluster) restricts the list to some
// clusters: every workgroup pulls items one atomic at a time, so items of
// retired clusters would cost a serial chain of atomics per step.
// fill: when the covered clusters give fewer than two items per compute unit
// (the last clusters of a run, often with the smallest steps), raise the
// lanes per i (up to 64) until they do, so a step's latency is not that of
// a few long items.
int build_items(Ensemble *E, const int *only = nullptr, bool fill = false)
{
    // the previous item arrays may still be read by queued work
    CK(hipStreamSynchronize(E->stream));
    auto njs_of = [&](int n, int boost) {
        return std::min(64, E->d.njs_class[size_class(n)] * boost);
    };
    int boost = 1;
    if (fill) {
        const int target = std::max(1, E->persistent_wgs / 2);
        for (; boost < 64; boost *= 2) {
            long long count = 0;
            bool room = false;
            for (int c = 0; c < E->d.nc; c++) {
                if (only && !only[c]) continue;
                const int rows = kWG / njs_of(E->cnt[c], boost);
                count += (E->cnt[c] + rows - 1) / rows;
                room |= njs_of(E->cnt[c], boost) < 64;
            }
            if (count >= target || !room) break;
        }
    }
    std::vector<int> ic, ii0, injs, plist;
    std::vector<double> cost;
    int covered = 0;
    for (int c = 0; c < E->d.nc; c++) {
        if (only && !only[c]) continue;
        covered++;
        const int n = E->cnt[c];
        for (int i = 0; i < n; i++) plist.push_back(E->off[c] + i);
        const int njs = njs_of(n, boost);
        const int rows = kWG / njs;
        for (int i0 = 0; i0 < n; i0 += rows) {
            ic.push_back(c);
            ii0.push_back(i0);
            injs.push_back(njs);
            cost.push_back(double(std::min(rows, n - i0)) * n / njs + 64.0 * n / kTJ);
        }
    }
    std::vector<int> order(ic.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int> a(order.size()), b(order.size()), cc(order.size());
    for (size_t k = 0; k < order.size(); k++) {
        a[k] = ic[order[k]];
        b[k] = ii0[order[k]];
        cc[k] = injs[order[k]];
    }
    const int n = int(order.size());
    if (E->d.item_c) {
        (void)hipFree(E->d.item_c);
        E->d.item_c = nullptr;
    }
    int *items = nullptr;
    CK(hipMalloc(reinterpret_cast<void **>(&items), sizeof(int) * 3 * size_t(std::max(n, 1))));
    E->d.item_c = items;
    E->d.item_i0 = E->d.item_c + n;
    E->d.item_njs = E->d.item_c + 2 * n;
    CK(hipMemcpy(E->d.item_c, a.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    CK(hipMemcpy(E->d.item_i0, b.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    CK(hipMemcpy(E->d.item_njs, cc.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (!plist.empty())
        CK(hipMemcpy(E->d.plist, plist.data(), sizeof(int) * plist.size(), hipMemcpyHostToDevice));
    E->d.nitems = n;
    E->d.nlist = int(plist.size());
    E->items_clusters = covered;
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
    return 0;
}

int launch_force(Ensemble *E, int prec)
{
    // no more workgroups than items: the surplus would only take the atomic
    // that tells it the queue is empty
    const int wgs = std::max(1, std::min(E->persistent_wgs, E->d.nitems));
    if (prec == G6X_MIXED)
        hipLaunchKernelGGL(kc_force<true>, dim3(wgs), dim3(kWG), 0, E->stream, E->d);
    else
        hipLaunchKernelGGL(kc_force<false>, dim3(wgs), dim3(kWG), 0, E->stream, E->d);
    return hipGetLastError() == hipSuccess ? 0 : G6X_ERR_HIP;
}

// blocks of the predictor / corrector over the listed particles (one at least,
// which resets the work queue)
int list_blocks(const Ensemble *E) { return std::max(1, (E->d.nlist + 255) / 256); }

int enqueue_step(Ensemble *E, int prec)
{
    const int pb = list_blocks(E), cb = (E->d.nc + 255) / 256;
    hipLaunchKernelGGL(kc_predict, dim3(pb), dim3(256), 0, E->stream, E->d, 0);
    if (launch_force(E, prec)) return G6X_ERR_HIP;
    hipLaunchKernelGGL(kc_correct, dim3(pb), dim3(256), 0, E->stream, E->d);
    hipLaunchKernelGGL(kc_update, dim3(cb), dim3(256), 0, E->stream, E->d, 0);
    return hipGetLastError() == hipSuccess ? 0 : G6X_ERR_HIP;
}

}  // namespace g6c

using namespace g6c;

// Frees whatever of E has been created (also a partly created ensemble).
static void release(Ensemble *E)
{
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    if (E->d.item_c) (void)hipFree(E->d.item_c);
    if (E->wl_base) (void)hipFree(E->wl_base);
    if (E->base) (void)hipFree(E->base);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

// Creates an ensemble of nc clusters; cluster c holds counts[c] particles,
// consecutive in m[], x[][3], v[][3].  precision: G6X_MIXED or G6X_FP64.
G6C_API void *g6c_create(int nc, const int counts[], const double m[], const double xv[][3],
                         const double vv[][3], double eps2, double dt_param, int precision)
{
    const double *x = &xv[0][0];
    const double *v = &vv[0][0];
    if (nc <= 0 || (precision != G6X_FP64 && precision != G6X_MIXED)) return nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return nullptr;
    // coordinates within kMaxCoord of the origin: the force kernel pads its
    // j-tiles with zero-mass entries at 1e17, whose terms are exact zeros
    // only for i far inside that
    long long ntot = 0;
    for (int c = 0; c < nc; c++) {
        if (counts[c] <= 0) return nullptr;
        ntot += counts[c];
    }
    if (ntot > (1ll << 28)) return nullptr;
    for (long long k = 0; k < 3 * ntot; k++)
        if (!(std::fabs(x[k]) <= kMaxCoord)) return nullptr;
    Ensemble *E = new Ensemble;
    E->precision = precision;
    E->cnt.assign(counts, counts + nc);
    E->off.resize(nc);
    long long np = 0;
    for (int c = 0; c < nc; c++) {
        E->off[c] = int(np);
        np += counts[c];
    }
    Dev &d = E->d;
    d.nc = nc;
    d.np = int(np);
    d.eps2 = eps2;
    d.dt_param = dt_param > 0.0 ? dt_param : 0.03;
    for (int k = 0; k < kClasses; k++) E->d.njs_class[k] = k <= 1 ? 16 : (k <= 3 ? 4 : 1);
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        E->stream = nullptr;
        release(E);
        return nullptr;
    }
    // one allocation: per-particle doubles (x v a j px pv na nj = 24, m, npot), floats (10),
    // ints (pc plist), per-cluster doubles (t dt energy) + uint (cmax) + ints (off cnt active nsteps)
    const size_t P = size_t(np), Cn = size_t(nc);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = al(P * 8) * 26 + al(P * 4) * 12 + al(Cn * 8) * 4 + al(Cn * 4) * 4 + 256;
    if (hipMalloc(&E->base, bytes) != hipSuccess) {
        E->base = nullptr;
        release(E);
        return nullptr;
    }
    char *q = static_cast<char *>(E->base);
    auto D = [&](double *&p) { p = reinterpret_cast<double *>(q); q += al(P * 8); };
    auto F = [&](float *&p) { p = reinterpret_cast<float *>(q); q += al(P * 4); };
    for (int k = 0; k < 3; k++) { D(d.x[k]); D(d.v[k]); D(d.a[k]); D(d.j[k]); D(d.px[k]); D(d.pv[k]); D(d.na[k]); D(d.nj[k]); }
    D(d.m); D(d.npot);
    for (int k = 0; k < 3; k++) { F(d.hx[k]); F(d.lx[k]); F(d.fv[k]); }
    F(d.fm);
    d.pc = reinterpret_cast<int *>(q); q += al(P * 4);
    d.plist = reinterpret_cast<int *>(q); q += al(P * 4);
    d.t = reinterpret_cast<double *>(q); q += al(Cn * 8);
    d.dt = reinterpret_cast<double *>(q); q += al(Cn * 8);
    d.energy = reinterpret_cast<double *>(q); q += al(Cn * 8);
    d.cmax = reinterpret_cast<unsigned int *>(q); q += al(Cn * 8);
    d.off = reinterpret_cast<int *>(q); q += al(Cn * 4);
    d.cnt = reinterpret_cast<int *>(q); q += al(Cn * 4);
    d.active = reinterpret_cast<int *>(q); q += al(Cn * 4);
    d.nsteps = reinterpret_cast<int *>(q); q += al(Cn * 4);
    d.queue = reinterpret_cast<int *>(q);
    d.nactive = d.queue + 1;

    std::vector<double> h(P);
    std::vector<float> hf(P);
    std::vector<int> pc(P);
    for (int c = 0; c < nc; c++)
        for (int i = 0; i < counts[c]; i++) pc[E->off[c] + i] = c;
    bool ok = true;
    for (int k = 0; k < 3 && ok; k++) {
        for (size_t p = 0; p < P; p++) h[p] = x[3 * p + k];
        ok &= hipMemcpy(d.x[k], h.data(), P * 8, hipMemcpyHostToDevice) == hipSuccess;
        for (size_t p = 0; p < P; p++) h[p] = v[3 * p + k];
        ok &= hipMemcpy(d.v[k], h.data(), P * 8, hipMemcpyHostToDevice) == hipSuccess;
    }
    for (size_t p = 0; p < P; p++) hf[p] = float(m[p]);
    ok &= hipMemcpy(d.m, m, P * 8, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.fm, hf.data(), P * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.pc, pc.data(), P * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.off, E->off.data(), Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.cnt, counts, Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
    std::vector<int> ones(nc, 1), zeros(nc, 0);
    std::vector<unsigned int> zero_u(nc, 0u);
    ok &= hipMemcpy(d.active, ones.data(), Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.nsteps, zeros.data(), Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(d.cmax, zero_u.data(), Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
    for (int k = 0; k < 3; k++) {
        ok &= hipMemset(d.a[k], 0, P * 8) == hipSuccess;
        ok &= hipMemset(d.j[k], 0, P * 8) == hipSuccess;
    }
    ok &= hipDeviceSynchronize() == hipSuccess;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    E->persistent_wgs = cus * 4;
    d.fill_target = std::max(1, E->persistent_wgs / 2);
    // work list of the persistent run: at least 4 rows per item, so at most
    // n/4 + 1 items per cluster
    d.wl_cap = int(std::min<long long>(P / 4 + Cn + 8, 1ll << 30));
    d.ctl = d.queue + 2;
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return counts[a] > counts[b]; });
    // work list, cluster order, then the grid barrier's words on a 256-byte
    // boundary
    const size_t bar_at = (3 * size_t(d.wl_cap) + Cn + 63) / 64 * 64;
    if (ok) {
        ok = hipMalloc(reinterpret_cast<void **>(&E->wl_base),
                       sizeof(int) * (bar_at + g6bar::kWords)) == hipSuccess;
        if (!ok) E->wl_base = nullptr;
    }
    if (ok) {
        d.wl_c = E->wl_base;
        d.wl_i0 = d.wl_c + d.wl_cap;
        d.wl_njs = d.wl_i0 + d.wl_cap;
        d.corder = d.wl_njs + d.wl_cap;
        d.bar = reinterpret_cast<unsigned int *>(E->wl_base + bar_at);
        ok &= hipMemcpy(d.corder, order.data(), Cn * 4, hipMemcpyHostToDevice) == hipSuccess;
        const int zero8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        ok &= hipMemcpy(d.ctl, zero8, sizeof(zero8), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok || build_items(E)) {
        release(E);
        return nullptr;
    }
    return E;
}

// The device-built work list no longer matches the tiles or the active set.
static int invalidate_work_list(Ensemble *E)
{
    const int zero = 0;
    CK(hipMemcpyAsync(E->d.ctl + 3, &zero, sizeof(int), hipMemcpyHostToDevice, E->stream));
    CK(hipStreamSynchronize(E->stream));
    return 0;
}

G6C_API const char *g6c_last_error(void *h) { return static_cast<Ensemble *>(h)->err.c_str(); }

// Forces at the current state (no prediction) in `precision`, then the
// per-cluster energies from their FP64 potential when precision is FP64.
static int force_now(Ensemble *E, int precision)
{
    // a run may have dropped the items of retired clusters
    if (E->items_clusters < E->d.nc && build_items(E)) return G6X_ERR_HIP;
    hipLaunchKernelGGL(kc_predict, dim3(list_blocks(E)), dim3(256), 0, E->stream, E->d, 1);
    if (launch_force(E, precision)) return G6X_ERR_HIP;
    return 0;
}

// Initial forces and first steps; every cluster becomes active at t = 0.
G6C_API int g6c_start(void *h, double t_end)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    E->d.t_end = t_end;
    hipLaunchKernelGGL(kc_set_active_count, dim3(1), dim3(1), 0, E->stream, E->d);
    // all clusters active for the initial pass
    std::vector<int> ones(E->d.nc, 1);
    CK(hipMemcpyAsync(E->d.active, ones.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice, E->stream));
    if (force_now(E, E->precision)) return G6X_ERR_HIP;
    const int pb = (E->d.np + 255) / 256, cb = (E->d.nc + 255) / 256;
    hipLaunchKernelGGL(kc_take_forces, dim3(pb), dim3(256), 0, E->stream, E->d);
    hipLaunchKernelGGL(kc_update, dim3(cb), dim3(256), 0, E->stream, E->d, 1);
    hipLaunchKernelGGL(kc_set_active_count, dim3(1), dim3(1), 0, E->stream, E->d);
    CK(hipStreamSynchronize(E->stream));
    return invalidate_work_list(E);
}

// use_graph == 2: up to max_rounds step rounds (0: until every cluster has
// reached t_end) in persistent cooperative launches of kc_run, one unless a
// launch's round budget (2^20 rounds) runs out first.
static long long run_persistent(Ensemble *E, long long max_rounds)
{
    int dev = 0, cus = 0, per_cu = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *kern = E->precision == G6X_MIXED ? reinterpret_cast<const void *>(&kc_run<true>)
                                                 : reinterpret_cast<const void *>(&kc_run<false>);
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kWG, 0));
    if (per_cu <= 0) {
        E->err = "kc_run: no workgroup fits a compute unit";
        return G6X_ERR_HIP;
    }
    const int grid = std::min(per_cu * cus, g6bar::kMaxGroups * g6bar::kGroup);
    long long *d_rounds = reinterpret_cast<long long *>(E->d.ctl + 4);
    const unsigned int *d_abort = E->d.bar + g6bar::abort_word();
    long long total = 0;
    for (;;) {
        long long budget = 1ll << 20;
        if (max_rounds > 0) budget = std::min(budget, max_rounds - total);
        if (budget <= 0) break;
        hipLaunchKernelGGL(kc_zero_words, dim3((g6bar::kWords + 255) / 256), dim3(256), 0,
                           E->stream, E->d.bar, g6bar::kWords);
        CK(hipGetLastError());
        void *args[] = {&E->d, &budget, &d_rounds};
        CK(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(kWG), args, 0, E->stream));
        long long done = 0;
        int nact = 0;
        unsigned int aborted = 0;
        CK(hipMemcpyAsync(&done, d_rounds, sizeof(done), hipMemcpyDeviceToHost, E->stream));
        CK(hipMemcpyAsync(&nact, E->d.nactive, sizeof(nact), hipMemcpyDeviceToHost, E->stream));
        CK(hipMemcpyAsync(&aborted, d_abort, sizeof(aborted), hipMemcpyDeviceToHost, E->stream));
        CK(hipStreamSynchronize(E->stream));
        total += done;
        E->launches++;
        if (aborted) {
            E->err = "kc_run: a grid barrier timed out; the run stopped after " +
                     std::to_string(total) + " rounds";
            return G6X_ERR_HIP;
        }
        if (nact <= 0) break;
    }
    return total;
}

static int capture_graph(Ensemble *E, int steps_per_graph)
{
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    E->graph = nullptr;
    hipGraph_t g;
    CK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps_per_graph; s++)
        if (enqueue_step(E, E->precision)) {
            (void)hipStreamEndCapture(E->stream, &g);
            return G6X_ERR_HIP;
        }
    CK(hipStreamEndCapture(E->stream, &g));
    CK(hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    E->graph_steps = steps_per_graph;
    E->graph_prec = E->precision;
    return 0;
}

// Advances every cluster to t_end.  use_graph = 1 captures `steps_per_graph`
// steps into one hipGraph and replays it; 0 launches the same kernels
// eagerly.  Returns the number of step rounds executed.
G6C_API long long g6c_run(void *h, int steps_per_graph, int use_graph, long long max_rounds)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    if (steps_per_graph < 1) steps_per_graph = 1;
    EventPair ev;
    if (!ev.create()) {
        E->err = "g6c_run: hipEventCreate failed";
        return G6X_ERR_HIP;
    }
    if (use_graph == 2) {
        (void)hipEventRecord(ev.a, E->stream);
        const long long r = run_persistent(E, max_rounds);
        (void)hipEventRecord(ev.b, E->stream);
        (void)hipEventSynchronize(ev.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev.a, ev.b);
        E->force_ms += ms;
        return r;
    }
    // the active count, read back after each launch (pinned, freed on return)
    struct Pinned {
        int *p = nullptr;
        ~Pinned()
        {
            if (p) (void)hipHostFree(p);
        }
    } pin;
    CK(hipHostMalloc(reinterpret_cast<void **>(&pin.p), sizeof(int)));
    int *h_active = pin.p;
    std::vector<int> flags(E->d.nc);
    long long rounds = 0;
    (void)hipEventRecord(ev.a, E->stream);
    long long rc = 0;
    for (;;) {
        if (max_rounds > 0 && rounds >= max_rounds) break;
        hipError_t e = hipSuccess;
        if (use_graph) {
            if (!E->graph || E->graph_steps != steps_per_graph || E->graph_prec != E->precision)
                if (capture_graph(E, steps_per_graph)) { rc = G6X_ERR_HIP; break; }
            e = hipGraphLaunch(E->graph, E->stream);
        } else {
            for (int s = 0; s < steps_per_graph && e == hipSuccess; s++)
                if (enqueue_step(E, E->precision)) e = hipErrorUnknown;
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_active, E->d.nactive, sizeof(int), hipMemcpyDeviceToHost, E->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
        if (e != hipSuccess) {
            E->err = std::string("step rounds: ") + hipGetErrorString(e);
            rc = G6X_ERR_HIP;
            break;
        }
        rounds += steps_per_graph;
        if (*h_active <= 0) break;
        if (*h_active * 2 <= E->items_clusters) {
            // half the covered clusters have retired: drop their items
            e = hipMemcpy(flags.data(), E->d.active, sizeof(int) * E->d.nc, hipMemcpyDeviceToHost);
            if (e != hipSuccess || build_items(E, flags.data(), true)) {
                if (e != hipSuccess) E->err = std::string("active flags: ") + hipGetErrorString(e);
                rc = G6X_ERR_HIP;
                break;
            }
        }
    }
    (void)hipEventRecord(ev.b, E->stream);
    (void)hipEventSynchronize(ev.b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev.a, ev.b);
    E->force_ms += ms;
    return rc < 0 ? rc : rounds;
}

// Per-cluster total energy at the current state, from an FP64 force pass.
G6C_API int g6c_energy(void *h, double *energy)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    // force_now requires the clusters to be marked active; keep their flags
    std::vector<int> saved(E->d.nc);
    CK(hipMemcpyAsync(saved.data(), E->d.active, sizeof(int) * E->d.nc, hipMemcpyDeviceToHost, E->stream));
    CK(hipStreamSynchronize(E->stream));
    std::vector<int> ones(E->d.nc, 1);
    CK(hipMemcpyAsync(E->d.active, ones.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice, E->stream));
    if (force_now(E, G6X_FP64)) return G6X_ERR_HIP;
    hipLaunchKernelGGL(kc_energy, dim3(E->d.nc), dim3(256), 0, E->stream, E->d);
    CK(hipMemcpyAsync(energy, E->d.energy, sizeof(double) * E->d.nc, hipMemcpyDeviceToHost, E->stream));
    CK(hipMemcpyAsync(E->d.active, saved.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice, E->stream));
    CK(hipStreamSynchronize(E->stream));
    return 0;
}

G6C_API int g6c_state(void *h, double xo[][3], double vo[][3], double t[], int nsteps[])
{
    double *x = &xo[0][0];
    double *v = &vo[0][0];
    Ensemble *E = static_cast<Ensemble *>(h);
    const size_t P = size_t(E->d.np);
    std::vector<double> buf(P);
    for (int k = 0; k < 3; k++) {
        CK(hipMemcpy(buf.data(), E->d.x[k], P * 8, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) x[3 * p + k] = buf[p];
        CK(hipMemcpy(buf.data(), E->d.v[k], P * 8, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) v[3 * p + k] = buf[p];
    }
    if (t) CK(hipMemcpy(t, E->d.t, sizeof(double) * E->d.nc, hipMemcpyDeviceToHost));
    if (nsteps) CK(hipMemcpy(nsteps, E->d.nsteps, sizeof(int) * E->d.nc, hipMemcpyDeviceToHost));
    return 0;
}

// Forces of the current state (no prediction) for every particle, in the
// given precision: acc, jerk [np][3], pot [np].
G6C_API int g6c_forces(void *h, int precision, double acco[][3], double jerko[][3], double pot[])
{
    double *acc = &acco[0][0];
    double *jerk = &jerko[0][0];
    Ensemble *E = static_cast<Ensemble *>(h);
    std::vector<int> saved(E->d.nc), ones(E->d.nc, 1);
    CK(hipMemcpy(saved.data(), E->d.active, sizeof(int) * E->d.nc, hipMemcpyDeviceToHost));
    CK(hipMemcpy(E->d.active, ones.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice));
    if (force_now(E, precision)) return G6X_ERR_HIP;
    CK(hipStreamSynchronize(E->stream));
    const size_t P = size_t(E->d.np);
    std::vector<double> buf(P);
    for (int k = 0; k < 3; k++) {
        CK(hipMemcpy(buf.data(), E->d.na[k], P * 8, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) acc[3 * p + k] = buf[p];
        CK(hipMemcpy(buf.data(), E->d.nj[k], P * 8, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < P; p++) jerk[3 * p + k] = buf[p];
    }
    CK(hipMemcpy(pot, E->d.npot, P * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(E->d.active, saved.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice));
    return 0;
}

// Sets the lanes-per-i of each size class (<=16, <=32, ..., <=2048, more).
G6C_API int g6c_set_tiles(void *h, const int *njs_per_class)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    for (int k = 0; k < kClasses; k++) {
        const int v = njs_per_class[k];
        if (v < 1 || v > 64 || (v & (v - 1))) return G6X_ERR_ARG;
    }
    std::copy(njs_per_class, njs_per_class + kClasses, E->d.njs_class);
    if (invalidate_work_list(E)) return G6X_ERR_HIP;
    return build_items(E);
}

// j per LDS tile of each size class (16 .. 256, a power of two).
G6C_API int g6c_set_jtiles(void *h, const int *tj_per_class)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    for (int k = 0; k < kClasses; k++) {
        const int v = tj_per_class[k];
        if (v < 16 || v > kTJ || (v & (v - 1))) return G6X_ERR_ARG;
    }
    std::copy(tj_per_class, tj_per_class + kClasses, E->d.tj_class);
    return 0;
}

G6C_API int g6c_get_jtiles(void *h, int *tj_per_class)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    std::copy(E->d.tj_class, E->d.tj_class + kClasses, tj_per_class);
    return kClasses;
}

G6C_API int g6c_get_tiles(void *h, int *njs_per_class)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    std::copy(E->d.njs_class, E->d.njs_class + kClasses, njs_per_class);
    return kClasses;
}

// The autotuner's search (times every shape for each size class present,
// keeps the fastest): the number of shapes timed, or an error code.
static int tune_classes(Ensemble *E, hipEvent_t a, hipEvent_t b)
{
    int timed = 0;
    for (int k = 0; k < kClasses; k++) {
        std::vector<int> act(E->d.nc, 0);
        bool any = false;
        for (int c = 0; c < E->d.nc; c++)
            if (size_class(E->cnt[c]) == k) { act[c] = 1; any = true; }
        if (!any) continue;
        CK(hipMemcpy(E->d.active, act.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice));
        int nmax = 0;
        for (int c = 0; c < E->d.nc; c++)
            if (act[c]) nmax = std::max(nmax, E->cnt[c]);
        float best = 1e30f;
        int best_njs = E->d.njs_class[k], best_tj = E->d.tj_class[k];
        // tiles shorter than the largest cluster of the class (a longer one
        // holds the whole cluster, like kTJ)
        for (int tj = kTJ; tj >= 64; tj /= 2) {
            if (tj < kTJ && tj >= nmax) continue;
            E->d.tj_class[k] = tj;
            for (int njs = 1; njs <= 64; njs *= 2) {
                E->d.njs_class[k] = njs;
                if (build_items(E, act.data())) return G6X_ERR_HIP;
                hipLaunchKernelGGL(kc_predict, dim3(list_blocks(E)), dim3(256), 0, E->stream, E->d, 1);
                if (launch_force(E, E->precision)) return G6X_ERR_HIP;
                // the best of two timings of `reps` passes each, reps raised
                // until a timing covers >= 1 ms (classes of small clusters
                // take a few µs per pass, where launch jitter would decide)
                int reps = 3;
                float ms_best = 1e30f;
                for (int trial = 0; trial < 2; trial++) {
                    CK(hipEventRecord(a, E->stream));
                    for (int r = 0; r < reps; r++) {
                        hipLaunchKernelGGL(kc_predict, dim3(list_blocks(E)), dim3(256), 0,
                                           E->stream, E->d, 1);
                        if (launch_force(E, E->precision)) return G6X_ERR_HIP;
                    }
                    CK(hipEventRecord(b, E->stream));
                    CK(hipEventSynchronize(b));
                    float ms = 0;
                    CK(hipEventElapsedTime(&ms, a, b));
                    ms_best = std::min(ms_best, ms / float(reps));
                    if (trial == 0 && ms < 1.0f)
                        reps = std::min(64, int(reps * (1.0f / std::max(ms, 1e-3f))) + 1);
                }
                timed++;
                if (ms_best < best) { best = ms_best; best_njs = njs; best_tj = tj; }
            }
        }
        E->d.njs_class[k] = best_njs;
        E->d.tj_class[k] = best_tj;
    }
    return timed;
}

// Autotuner: for each size class present, times the force pass restricted
// to that class's clusters for every (lanes per i, j per tile) pair -- njs
// in {1, 2, ..., 64} (i-block rows 256/njs), tile 64, 128 or 256 j where
// shorter than the class's largest cluster -- and keeps the fastest.  The
// caller's active flags come back as they were, and after an error so do
// the shape tables.
G6C_API int g6c_autotune(void *h)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    std::vector<int> saved(E->d.nc);
    CK(hipMemcpy(saved.data(), E->d.active, sizeof(int) * E->d.nc, hipMemcpyDeviceToHost));
    int njs0[kClasses], tj0[kClasses];
    std::copy(E->d.njs_class, E->d.njs_class + kClasses, njs0);
    std::copy(E->d.tj_class, E->d.tj_class + kClasses, tj0);
    int rc = G6X_ERR_HIP;
    {
        EventPair ev;
        if (ev.create())
            rc = tune_classes(E, ev.a, ev.b);
        else
            E->err = "g6c_autotune: hipEventCreate failed";
    }
    if (rc < 0) {
        std::copy(njs0, njs0 + kClasses, E->d.njs_class);
        std::copy(tj0, tj0 + kClasses, E->d.tj_class);
    }
    // items of every cluster again, and the caller's active flags
    const int rb = build_items(E);
    CK(hipMemcpy(E->d.active, saved.data(), sizeof(int) * E->d.nc, hipMemcpyHostToDevice));
    if (invalidate_work_list(E)) return G6X_ERR_HIP;
    if (rc < 0) return rc;
    return rb ? rb : rc;
}

G6C_API int g6c_set_precision(void *h, int precision)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    if (precision != G6X_FP64 && precision != G6X_MIXED) return G6X_ERR_ARG;
    E->precision = precision;
    return 0;
}

G6C_API double g6c_elapsed_ms(void *h) { return static_cast<Ensemble *>(h)->force_ms; }

G6C_API void g6c_free(void *h)
{
    Ensemble *E = static_cast<Ensemble *>(h);
    if (E) release(E);
}

// GRAPE-6 force path kernels for gfx950 (CDNA4, wave64).
//
//   reset / scatter / predict   j-store maintenance, HBM-streaming
//   iblock_self                 gathers predicted j-slots as an i-block
//   force_f64<NJS,IPL,NN>       plain path: FP64 differences, r^-1/2 and sums
//   force_mx<NJS,IPL,NN>        mixed path: double-single dx, FP32 r^-3 and
//                               sub-tile sums, Kahan/TwoSum outer accumulators
//   reduce                      ordered sum of per-chunk partials + nn argmin
//
// Arithmetic follows lib/g6lib/g6lib.c: predictor :64-92, pair loop
// :123-186 (skip id_i == id_j and id_j == -1, nearest neighbour on the
// unsoftened r^2 with strict '<' so the lowest slot wins a tie, potential
// -m/sqrt(r^2+eps2), acc m r / s^1.5, jerk m (v/s^1.5 - 3 (r.v) r / s^2.5)).
//
// Force kernel shape: a workgroup of 256 lanes covers IB = IPL*256/NJS
// i-particles; NJS lanes share one i and take every NJS-th j of a 256-j tile
// staged in LDS as structure-of-arrays (conflict-free ds_read_b128, a
// broadcast when NJS = 1).  The NJS partial sums of one i are combined with
// cross-lane shuffles (DPP / ds_swizzle), across waves through LDS when
// NJS > 64.  Workgroups along grid.y split j into chunks; their partials are
// summed in chunk order by the reduce kernel, so results do not depend on
// scheduling.  The nearest-neighbour key is the pair (r^2, slot) compared
// lexicographically, which reproduces the sequential first-wins rule for any
// j-split.
#include "g6_kernels.h"
#include "g6_predict.h"
#include "g6_reduce.h"

#include <algorithm>
#include <cfloat>

namespace g6k {


// inner-loop unroll depths (build-time tunables)
#ifndef G6_MX_UNROLL
#define G6_MX_UNROLL 2
#endif
#ifndef G6_F64_UNROLL
#define G6_F64_UNROLL 2
#endif
// spatial-order mixed kernel: its far/near branch per tile already gives the
// scheduler two loop bodies; unrolled once more it ran 3 % slower at N = 2^20
#ifndef G6_SP_UNROLL
#define G6_SP_UNROLL 1
#endif

// ------------------------------------------------------------------------
// Memory carving
// ------------------------------------------------------------------------
static inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t jstore_bytes(int cap)
{
    size_t n = size_t(cap);
    size_t b = 0;
    b += align256(n * sizeof(double)) * (15 + 3 + 7);  // base 15+tj,dtj,m ; pred 7
    b += align256(n * sizeof(float)) * 10;             // mixed pred
    b += align256(n * sizeof(int)) * 2;                // id, pid
    return b;
}

void jstore_carve(JStore &s, void *base, int cap)
{
    char *p = static_cast<char *>(base);
    size_t n = size_t(cap);
    auto takeD = [&](double *&d) { d = reinterpret_cast<double *>(p); p += align256(n * sizeof(double)); };
    auto takeF = [&](float *&d) { d = reinterpret_cast<float *>(p); p += align256(n * sizeof(float)); };
    auto takeI = [&](int *&d) { d = reinterpret_cast<int *>(p); p += align256(n * sizeof(int)); };
    for (int k = 0; k < 3; k++) { takeD(s.x[k]); takeD(s.v[k]); takeD(s.a2[k]); takeD(s.j6[k]); takeD(s.k18[k]); }
    takeD(s.tj); takeD(s.dtj); takeD(s.m);
    for (int k = 0; k < 3; k++) { takeD(s.px[k]); takeD(s.pv[k]); }
    takeD(s.pm);
    for (int k = 0; k < 3; k++) { takeF(s.hx[k]); takeF(s.lx[k]); takeF(s.fv[k]); }
    takeF(s.fm);
    takeI(s.id); takeI(s.pid);
    s.cap = cap;
    s.base = base;
}

size_t iblock_bytes(int cap)
{
    return align256(size_t(cap) * sizeof(double)) * 16;
}

// Sixteen rows of cap doubles: x y z vx vy vz h2 id | acc3 jerk3 pot nn, so
// that one pitched copy moves a whole i-block in and one moves the results
// out (the id and nn rows hold ints in the first half of their row).
void iblock_carve(IBlock &b, void *base, int cap)
{
    char *p = static_cast<char *>(base);
    const size_t row = align256(size_t(cap) * sizeof(double));
    for (int k = 0; k < 3; k++) b.x[k] = reinterpret_cast<double *>(p + k * row);
    for (int k = 0; k < 3; k++) b.v[k] = reinterpret_cast<double *>(p + (3 + k) * row);
    b.h2 = reinterpret_cast<double *>(p + 6 * row);
    b.id = reinterpret_cast<int *>(p + 7 * row);
    b.out = reinterpret_cast<double *>(p + 8 * row);
    b.nn = reinterpret_cast<int *>(p + 15 * row);
    b.cap = cap;
    b.base = base;
}

__global__ void __launch_bounds__(256) k_reset(JStore s, int lo, int hi)
{
    const int j = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= hi) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s.x[k][j] = 0.0; s.v[k][j] = 0.0; s.a2[k][j] = 0.0; s.j6[k][j] = 0.0; s.k18[k][j] = 0.0;
        s.px[k][j] = kFar; s.pv[k][j] = 0.0;
        s.hx[k][j] = kFarF; s.lx[k][j] = 0.0f; s.fv[k][j] = 0.0f;
    }
    s.tj[j] = 0.0; s.dtj[j] = 0.0; s.m[j] = 0.0; s.pm[j] = 0.0; s.fm[j] = 0.0f;
    s.id[j] = -1; s.pid[j] = -1;
}

__global__ void __launch_bounds__(256) k_copy_store(JStore d, JStore s, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d.x[k][j] = s.x[k][j]; d.v[k][j] = s.v[k][j]; d.a2[k][j] = s.a2[k][j];
        d.j6[k][j] = s.j6[k][j]; d.k18[k][j] = s.k18[k][j];
        d.px[k][j] = s.px[k][j]; d.pv[k][j] = s.pv[k][j];
        d.hx[k][j] = s.hx[k][j]; d.lx[k][j] = s.lx[k][j]; d.fv[k][j] = s.fv[k][j];
    }
    d.tj[j] = s.tj[j]; d.dtj[j] = s.dtj[j]; d.m[j] = s.m[j]; d.pm[j] = s.pm[j]; d.fm[j] = s.fm[j];
    d.id[j] = s.id[j]; d.pid[j] = s.pid[j];
}

// Applies staged set_j_particle records; when the predicted arrays are
// current for time ti they are updated for the touched slots as well, so
// a later calc at the same ti needs no full predictor pass.
__global__ void __launch_bounds__(256) k_scatter(JStore s, const JRecord *__restrict__ rec, int n,
                                                 int prec, int pred_valid, double ti, int compat)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const JRecord q = rec[r];
    const int j = q.addr;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s.x[k][j] = q.x[k]; s.v[k][j] = q.v[k]; s.a2[k][j] = q.a2[k];
        s.j6[k][j] = q.j6[k]; s.k18[k][j] = q.k18[k];
    }
    s.tj[j] = q.tj; s.dtj[j] = q.dtj; s.m[j] = q.m; s.id[j] = q.id;
    const double mp = q.id == -1 ? 0.0 : q.m;
    s.pm[j] = mp;
    s.fm[j] = float(mp);
    s.pid[j] = q.id;
    if (pred_valid) {
        const Pred p = predict_slot(ti, q.tj, q.x, q.v, q.a2, q.j6, q.k18, compat, true);
        write_pred(s, j, prec, q.id, p);
    }
}

// Full predictor pass over slots [0, nj): streams 13 doubles + 1 int in
// and 6 doubles (plain) or 9 floats (mixed) out per slot; HBM-bound.
__global__ void __launch_bounds__(256) k_predict(JStore s, int nj, double ti, int prec, int compat,
                                                 int use_k18)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    double x[3], v[3], a2[3], j6[3], k18[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        x[k] = s.x[k][j]; v[k] = s.v[k][j]; a2[k] = s.a2[k][j]; j6[k] = s.j6[k][j];
    }
    if (use_k18) {
#pragma unroll
        for (int k = 0; k < 3; k++) k18[k] = s.k18[k][j];
    }
    const int id = s.id[j];
    const Pred p = predict_slot(ti, s.tj[j], x, v, a2, j6, k18, compat, use_k18 != 0);
    write_pred(s, j, prec, id, p);
}

__global__ void __launch_bounds__(256) k_iblock_self(JStore s, IBlock b, int i0, int ni, int prec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int j = i0 + i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (prec == kF64) {
            b.x[k][i] = s.px[k][j];
            b.v[k][i] = s.pv[k][j];
        } else {
            b.x[k][i] = double(s.hx[k][j]) + double(s.lx[k][j]);
            b.v[k][i] = double(s.fv[k][j]);
        }
    }
    b.h2[i] = 0.0;
    b.id[i] = s.id[j];
}

// Result rows out[7][S] (acc xyz, jerk xyz, pot) and nn rows -> [ni][3]
// acc and jerk, pot[ni] and nn[ni] (nn may be null).
__global__ void __launch_bounds__(256) k_rows_to_aos(const double *__restrict__ out,
                                                     const int *__restrict__ nnr, int S, int ni,
                                                     double *acc, double *jerk, double *pot,
                                                     int *nn)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        acc[3 * size_t(i) + k] = out[k * size_t(S) + i];
        jerk[3 * size_t(i) + k] = out[(3 + k) * size_t(S) + i];
    }
    pot[i] = out[6 * size_t(S) + i];
    if (nn) nn[i] = nnr[i];
}

// ------------------------------------------------------------------------
// Pair interaction, plain FP64.
// ------------------------------------------------------------------------
// r^-1/2 to full double precision: v_rsq_f64 seed, then one third-order
// correction y(1 + e/2 + 3e^2/8) with e = 1 - s y^2.
__device__ __forceinline__ double rsqrt_f64(double s)
{
    const double y = __builtin_amdgcn_rsq(s);
    const double h = s * y;
    const double e = __builtin_fma(-h, y, 1.0);
    const double t = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y * e, t, y);
}

// r^2 rounded exactly as the reference: ((dx*dx + dy*dy) + dz*dz).
__device__ __forceinline__ double exact_r2(double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    double r2 = dx * dx;
    r2 = r2 + dy * dy;
    r2 = r2 + dz * dz;
    return r2;
}

struct AccD {
    double ax, ay, az, jx, jy, jz, pot;
    double best;   // nearest-neighbour r^2 so far (blocks before the current one)
    int addr;      // its slot (NNM 2: found at the end, from blo/bhi)
    double tbest;  // NNM 2: running minimum of r^2 over the current block of j
    int blo, bhi;  // NNM 2: slot range of the lane's block that holds best
};

struct ID {
    double x, y, z, vx, vy, vz;
    int id;
};

// double-double accumulator (hi + lo) fed with TwoSum: the plain per-lane
// sums run over at most a few hundred pairs before being folded in, so the
// error of a full sum does not grow with nj.
struct DDAcc {
    double h[7], l[7];
};


__device__ __forceinline__ void flush_f64(AccD &a, DDAcc &d)
{
    two_sum_add_d(d.h[0], d.l[0], a.ax);
    two_sum_add_d(d.h[1], d.l[1], a.ay);
    two_sum_add_d(d.h[2], d.l[2], a.az);
    two_sum_add_d(d.h[3], d.l[3], a.jx);
    two_sum_add_d(d.h[4], d.l[4], a.jy);
    two_sum_add_d(d.h[5], d.l[5], a.jz);
    two_sum_add_d(d.h[6], d.l[6], a.pot);
    a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.pot = 0.0;
}

// NNM: 0 no nearest neighbour; 1 per-pair compare, first slot wins; 2 a
// running minimum of r^2 over a block of j, the slot found at the end (as
// pair_mx).  CHK: pairs with equal ids are skipped (y = 0, never the nearest
// neighbour); without it the caller knows no j of the item shares an id
// with an i of it.
template <int NNM, bool CHK = true>
__device__ __forceinline__ void pair_f64(const ID &I, AccD &a, double xj, double yj, double zj,
                                         double mj, double vxj, double vyj, double vzj, int idj,
                                         int addr, double eps2)
{
    const double dx = xj - I.x, dy = yj - I.y, dz = zj - I.z;
    // r^2 and s = r^2 + eps2 rounded as lib/g6lib/g6lib.c:149,156 does, in
    // every variant, so forces do not depend on whether nn was asked for
    const double r2 = exact_r2(dx, dy, dz);
    const double s = r2 + eps2;
    const bool valid = !CHK || idj != I.id;
    double y = rsqrt_f64(s);
    if (CHK) y = valid ? y : 0.0;
    if (NNM == 1) {
        const bool better = valid && (r2 < a.best);
        a.best = better ? r2 : a.best;
        a.addr = better ? addr : a.addr;
    }
    if (NNM == 2) a.tbest = __builtin_fmin(a.tbest, valid ? r2 : __builtin_inf());
    const double y2 = y * y;
    const double mi = mj * y;
    const double mi3 = mi * y2;
    a.pot -= mi;
    a.ax = __builtin_fma(mi3, dx, a.ax);
    a.ay = __builtin_fma(mi3, dy, a.ay);
    a.az = __builtin_fma(mi3, dz, a.az);
    const double dvx = vxj - I.vx, dvy = vyj - I.vy, dvz = vzj - I.vz;
    const double rv = __builtin_fma(dz, dvz, __builtin_fma(dy, dvy, dx * dvx));
    const double c = (-3.0 * rv) * y2;
    a.jx = __builtin_fma(mi3, __builtin_fma(c, dx, dvx), a.jx);
    a.jy = __builtin_fma(mi3, __builtin_fma(c, dy, dvy), a.jy);
    a.jz = __builtin_fma(mi3, __builtin_fma(c, dz, dvz), a.jz);
}


struct FParams {
    const double *ix, *iy, *iz, *ivx, *ivy, *ivz;
    const int *iid;
    int ni;
    // plain-path j
    const double *jx, *jy, *jz, *jm, *jvx, *jvy, *jvz;
    // mixed-path j
    const float *hx, *hy, *hz, *lx, *ly, *lz, *fm, *fvx, *fvy, *fvz;
    const int *jid;
    int nj, jchunk;
    double eps2;
    double *out;  // [7][ostride]
    int *nn;
    int ostride;
    double *part;  // [chunk][fields][pstride]
    int pstride;
    // j-slots predicted as tiles are loaded (items with PRED): the base
    // records and the time they are predicted to
    const double *bx[3], *bv[3], *ba2[3], *bj6[3], *btj;
    double tpred;
    int compat;
    // >= 0: i row r is j-slot self_i0 + r, and no two of the nj slots share
    // an id (none is empty), so only an item whose j-range holds one of its
    // i-slots can meet a pair with equal ids; -1: ids compared in every pair.
    int self_i0 = -1;
    // spatially ordered full pass: per j-tile (kTJ slots) the box of the
    // positions, min xyz then max xyz (k_force_mx_sp)
    const float *tbox = nullptr;
};

// Lanes' j per item from which the nearest neighbour is kept as a running
// minimum per block of a lane's j with one slot search at the end (below: a
// compare per pair).
constexpr int kLongItem = 16384;
// a lane's j per such block in the FP64 kernels (the final search gives one
// lane of a wave to each j of a block)
constexpr int kBlockJ = 64;

// Whether an item (i rows [r0, r0 + rows), j-chunk by) must compare ids.
__device__ __forceinline__ bool needs_id_check(const FParams &p, int r0, int rows, int by)
{
    if (p.self_i0 < 0) return true;
    const int j0 = by * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);
    const int s0 = p.self_i0 + r0;
    const int s1 = p.self_i0 + min(p.ni, r0 + rows);
    return s0 < j1 && j0 < s1;
}

// j-slot j predicted to p.tpred from its base record, as the predictor
// kernels do (k18 = 0); the caller converts it as write_pred does.
__device__ __forceinline__ Pred predict_base(const FParams &p, int j)
{
    double x[3], v[3], a2[3], j6[3];
    const double k18[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        x[k] = p.bx[k][j]; v[k] = p.bv[k][j]; a2[k] = p.ba2[k][j]; j6[k] = p.bj6[k][j];
    }
    return predict_slot(p.tpred, p.btj[j], x, v, a2, j6, k18, p.compat, false);
}

// Shared-memory layout of one j tile.
struct TileD {
    double2 xy[kTJ], zm[kTJ], vxy[kTJ], vzid[kTJ];
};
struct TileF {
    float4 a[kTJ];  // hx hy hz m
    float4 b[kTJ];  // lx ly lz vx
    float4 c[kTJ];  // vy vz id -
};

// One (i-block bx, j-chunk by) work item; the kernel below runs one item
// per workgroup, kd_* below loops over items whose count lives on the device.
// PRED: the j-slots are predicted to p.tpred as the tiles are loaded (the
// values write_pred would store), not read from the predicted store.
// LONG (with NN): the nearest neighbour is a running minimum per block of
// kBlockJ of a lane's j, its slot found once at the end (as force_mx_item);
// otherwise a compare and selects per pair.  CHK: as in pair_f64.
template <int NJS, int IPL, bool NN, bool LONG = false, bool PRED = false, bool CHK = true>
__device__ __forceinline__ void force_f64_item(const FParams &p, TileD &tile, int bx, int by,
                                       int nchunk)
{
    constexpr int NNM = NN ? (LONG ? 2 : 1) : 0;
    constexpr int IBL = kWG / NJS;
    constexpr int IB = IBL * IPL;
    // tiles between two folds of the plain sums: >= 128 pairs per lane
    constexpr int FT = NJS >= 4 ? NJS / 2 : 1;

    const int t = threadIdx.x;
    const int js = t % NJS;
    const int il = t / NJS;
    const int ibase = bx * IB;

    ID I[IPL];
    AccD A[IPL];
    DDAcc D[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = min(ibase + il + q * IBL, p.ni - 1);
        I[q].x = p.ix[i]; I[q].y = p.iy[i]; I[q].z = p.iz[i];
        I[q].vx = p.ivx[i]; I[q].vy = p.ivy[i]; I[q].vz = p.ivz[i];
        I[q].id = p.iid[i];
        A[q] = AccD{0, 0, 0, 0, 0, 0, 0, __builtin_inf(), -1, __builtin_inf(), 0, 0};
#pragma unroll
        for (int k = 0; k < 7; k++) { D[q].h[k] = 0.0; D[q].l[k] = 0.0; }
    }
    const double eps2 = p.eps2;
    const int j0 = by * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);

    int tcount = 0;
    for (int tb = j0; tb < j1; tb += kTJ) {
        const int jn = min(kTJ, j1 - tb);
        __syncthreads();
        if (t < jn) {
            const int j = tb + t;
            if constexpr (PRED) {
                const Pred pr = predict_base(p, j);
                const bool empty = p.jid[j] == -1;
                tile.xy[t] = make_double2(empty ? kFar : pr.x[0], empty ? kFar : pr.x[1]);
                tile.zm[t] = make_double2(empty ? kFar : pr.x[2], p.jm[j]);
                tile.vxy[t] = make_double2(empty ? 0.0 : pr.v[0], empty ? 0.0 : pr.v[1]);
                tile.vzid[t] = make_double2(empty ? 0.0 : pr.v[2],
                                            __longlong_as_double((long long)p.jid[j]));
            } else {
                tile.xy[t] = make_double2(p.jx[j], p.jy[j]);
                tile.zm[t] = make_double2(p.jz[j], p.jm[j]);
                tile.vxy[t] = make_double2(p.jvx[j], p.jvy[j]);
                tile.vzid[t] = make_double2(p.jvz[j], __longlong_as_double((long long)p.jid[j]));
            }
        }
        __syncthreads();
        for (int jb = js; jb < jn; jb += NJS * kBlockJ) {
            const int je = min(jn, jb + NJS * kBlockJ);
#pragma unroll G6_F64_UNROLL
            for (int jj = jb; jj < je; jj += NJS) {
                const double2 a = tile.xy[jj], b = tile.zm[jj], c = tile.vxy[jj],
                              d = tile.vzid[jj];
                const int idj = int(__double_as_longlong(d.y));
#pragma unroll
                for (int q = 0; q < IPL; q++)
                    pair_f64<NNM, CHK>(I[q], A[q], a.x, a.y, b.x, b.y, c.x, c.y, d.x, idj,
                                       tb + jj, eps2);
            }
            if (NNM == 2) {
                // a block minimum below the best so far (strictly: an equal
                // r^2 keeps the earlier block) moves the best to this block
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    const bool imp = A[q].tbest < A[q].best;
                    A[q].best = imp ? A[q].tbest : A[q].best;
                    A[q].blo = imp ? tb + jb : A[q].blo;
                    A[q].bhi = imp ? tb + je : A[q].bhi;
                    A[q].tbest = __builtin_inf();
                }
            }
        }
        if (++tcount == FT) {
            tcount = 0;
#pragma unroll
            for (int q = 0; q < IPL; q++) flush_f64(A[q], D[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < IPL; q++) flush_f64(A[q], D[q]);
    if (NNM == 2) {
        // The slot: the first j of the best block, in the lane's stride, with
        // exactly that r^2 (recomputed from the arrays the tiles were loaded
        // from, so bit-identical); the wave searches one lane's block at a
        // time, lane k taking the block's k-th j, and the first match is the
        // lowest set bit of the ballot.
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            for (int L = 0; L < 64; L++) {
                const double best = __shfl(A[q].best, L);
                if (!(best < __builtin_inf())) continue;  // wave-uniform
                const int blo = __shfl(A[q].blo, L), bhi = __shfl(A[q].bhi, L);
                const double xi = __shfl(I[q].x, L), yi = __shfl(I[q].y, L),
                             zi = __shfl(I[q].z, L);
                const int idi = __shfl(I[q].id, L);
                const int j = blo + lane * NJS;
                bool hit = false;
                if (j < bhi) {
                    const double r2 = exact_r2(p.jx[j] - xi, p.jy[j] - yi, p.jz[j] - zi);
                    hit = r2 == best && (!CHK || p.jid[j] != idi);
                }
                const unsigned long long mask = __ballot(hit);
                if (lane == L && mask) A[q].addr = blo + (__ffsll((long long)mask) - 1) * NJS;
            }
        }
    }

    // Combine the NJS lanes of each i (double-double adds).
#pragma unroll
    for (int q = 0; q < IPL; q++) {
#pragma unroll
        for (int off = 1; off < (NJS < 64 ? NJS : 64); off <<= 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                const double oh = __shfl_xor(D[q].h[k], off);
                const double ol = __shfl_xor(D[q].l[k], off);
                two_sum_add_d(D[q].h[k], D[q].l[k], oh);
                D[q].l[k] += ol;
            }
            if (NN) {
                const double ob = __shfl_xor(A[q].best, off);
                const int oa = __shfl_xor(A[q].addr, off);
                nn_merge(A[q].best, A[q].addr, ob, oa);
            }
        }
    }
    if (NJS > 64) {
        // Lanes 0 of the NJS/64 waves of each i meet in LDS.
        constexpr int W = NJS > 64 ? NJS / 64 : 1;
        __syncthreads();
        double *red = reinterpret_cast<double *>(&tile);  // [wave][IPL][16]
        const int wave = t / 64;
        if ((t & 63) == 0) {
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                double *r = red + (wave * IPL + q) * 16;
#pragma unroll
                for (int k = 0; k < 7; k++) { r[k] = D[q].h[k]; r[7 + k] = D[q].l[k]; }
                r[14] = A[q].best;
                r[15] = double(A[q].addr);
            }
        }
        __syncthreads();
        if (js == 0) {
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                for (int w = 1; w < W; w++) {
                    const double *r = red + ((wave + w) * IPL + q) * 16;
#pragma unroll
                    for (int k = 0; k < 7; k++) {
                        two_sum_add_d(D[q].h[k], D[q].l[k], r[k]);
                        D[q].l[k] += r[7 + k];
                    }
                    if (NN) nn_merge(A[q].best, A[q].addr, r[14], int(r[15]));
                }
            }
        }
    }
    if (js != 0) return;
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = ibase + il + q * IBL;
        if (i >= p.ni) continue;
        const int S = p.ostride;
        if (nchunk == 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) p.out[k * S + i] = D[q].h[k] + D[q].l[k];
            if (NN) p.nn[i] = A[q].addr >= 0 ? p.jid[A[q].addr] : -1;
        } else {
            // partial record: hi and lo of each sum in consecutive field pairs
            const int PS = p.pstride;
            double *r = p.part + size_t(by) * kPartFieldsD * PS;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                r[(2 * k) * PS + i] = D[q].h[k];
                r[(2 * k + 1) * PS + i] = D[q].l[k];
            }
            r[14 * PS + i] = NN ? A[q].best : __builtin_inf();
            r[15 * PS + i] = NN ? double(A[q].addr) : -1.0;
        }
    }
}

// LONG (with NN) is chosen by the host (long_items): one kernel holding all
// four (LONG, id compare) forms needs ~220 VGPRs against ~166 for two.
template <int NJS, int IPL, bool NN, bool LONG>
__global__ void __launch_bounds__(kWG) k_force_f64(FParams p)
{
    __shared__ TileD tile;
    if (needs_id_check(p, blockIdx.x * (kWG / NJS) * IPL, (kWG / NJS) * IPL, blockIdx.y))
        force_f64_item<NJS, IPL, NN, LONG>(p, tile, blockIdx.x, blockIdx.y, gridDim.y);
    else
        force_f64_item<NJS, IPL, NN, LONG, false, false>(p, tile, blockIdx.x, blockIdx.y,
                                                         gridDim.y);
}

// Short j-ranges with the j-slots predicted as they are loaded (one i per
// wave, one chunk): the step of a small system needs no predictor launch.
template <bool NN>
__global__ void __launch_bounds__(kWG) k_force_f64_p(FParams p)
{
    __shared__ TileD tile;
    force_f64_item<64, 1, NN, false, true>(p, tile, blockIdx.x, 0, 1);
}

// ------------------------------------------------------------------------
// Pair interaction, mixed precision.
// ------------------------------------------------------------------------
struct IF {
    float hx, hy, hz, lx, ly, lz, vx, vy, vz;
    int id;
};
struct AccF {
    // sub-tile FP32 sums
    float ax, ay, az, jx, jy, jz, pot;
    float best;   // nearest-neighbour r^2 so far (blocks before the current one)
    int addr;     // its slot (found at the end, from blo/bhi)
    float tbest;  // running minimum of r^2 over the current block of j
    int blo, bhi; // slot range of the lane's block that holds best
};
// double-single accumulator (hi + lo), fed with TwoSum
struct DSAcc {
    float h[7], l[7];
};


// Difference vector and r^2 of one pair, mixed path: dx from double-single
// positions, r^2 = (dx dx + dy dy) + dz dz with FMAs.  CHK: the self pair
// (same id) gets r^2 = +inf, so s = +inf, r^-1/2 = 0 and every term is
// exactly zero, and it can never be the nearest neighbour; without CHK the
// caller knows no j of the item shares an id with an i of it.
template <bool CHK = true>
__device__ __forceinline__ float pair_r2_mx(const IF &I, float4 ja, float4 jb, float4 jc,
                                            float &dx, float &dy, float &dz)
{
    dx = (ja.x - I.hx) + (jb.x - I.lx);
    dy = (ja.y - I.hy) + (jb.y - I.ly);
    dz = (ja.z - I.hz) + (jb.z - I.lz);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    if (!CHK) return r2;
    return __float_as_int(jc.z) != I.id ? r2 : __builtin_inff();
}

// NNM: 0 no nearest neighbour; 1 per-pair compare, first slot wins; 2 only a
// running minimum of r^2 over a block of j: the caller keeps the block that
// first reached the overall minimum and finds the slot in it once, at the
// end, which takes the compare and two selects out of the pair.
template <int NNM, bool CHK = true>
__device__ __forceinline__ void pair_mx(const IF &I, AccF &a, float4 ja, float4 jb, float4 jc,
                                        int addr, float eps2)
{
    float dx, dy, dz;
    const float r2 = pair_r2_mx<CHK>(I, ja, jb, jc, dx, dy, dz);
    // same s in every variant (forces independent of the nn request)
    const float s = r2 + eps2;
    const float y = __builtin_amdgcn_rsqf(s);
    if (NNM == 1) {
        const bool better = r2 < a.best;
        a.best = better ? r2 : a.best;
        a.addr = better ? addr : a.addr;
    }
    if (NNM == 2) a.tbest = fminf(a.tbest, r2);
    const float y2 = y * y;
    const float mi = ja.w * y;
    const float mi3 = mi * y2;
    a.pot -= mi;
    a.ax = __builtin_fmaf(mi3, dx, a.ax);
    a.ay = __builtin_fmaf(mi3, dy, a.ay);
    a.az = __builtin_fmaf(mi3, dz, a.az);
    const float dvx = jb.w - I.vx, dvy = jc.x - I.vy, dvz = jc.y - I.vz;
    const float rv = __builtin_fmaf(dz, dvz, __builtin_fmaf(dy, dvy, dx * dvx));
    const float c = (-3.0f * rv) * y2;
    a.jx = __builtin_fmaf(mi3, __builtin_fmaf(c, dx, dvx), a.jx);
    a.jy = __builtin_fmaf(mi3, __builtin_fmaf(c, dy, dvy), a.jy);
    a.jz = __builtin_fmaf(mi3, __builtin_fmaf(c, dz, dvz), a.jz);
}

__device__ __forceinline__ void flush_mx(AccF &a, DSAcc &d)
{
    two_sum_add(d.h[0], d.l[0], a.ax);
    two_sum_add(d.h[1], d.l[1], a.ay);
    two_sum_add(d.h[2], d.l[2], a.az);
    two_sum_add(d.h[3], d.l[3], a.jx);
    two_sum_add(d.h[4], d.l[4], a.jy);
    two_sum_add(d.h[5], d.l[5], a.jz);
    two_sum_add(d.h[6], d.l[6], a.pot);
    a.ax = a.ay = a.az = a.jx = a.jy = a.jz = a.pot = 0.0f;
}

// pairs per lane between two flushes of the FP32 sub-tile sums
#ifndef G6_MX_FLUSH
#define G6_MX_FLUSH 64
#endif
constexpr int kFlush = G6_MX_FLUSH;
// the final nearest-neighbour search gives one lane to each j of a block
static_assert(kFlush <= 64, "a block of a lane's j must fit one wave");


// End of a mixed work item: the NJS lanes of each i combined (double-single
// adds, across waves through LDS when NJS > 64), then the sums and nearest
// neighbour written as results (one chunk) or as the chunk's partial record.
template <int NJS, int IPL, bool NN>
__device__ __forceinline__ void mx_item_epilogue(const FParams &p, TileF &tile, IF *I, AccF *A,
                                                 DSAcc *D, int ibase, int il, int js, int by,
                                                 int nchunk)
{
    constexpr int IBL = kWG / NJS;
    const int t = threadIdx.x;
    // Combine the NJS lanes of each i (double-single adds).
#pragma unroll
    for (int q = 0; q < IPL; q++) {
#pragma unroll
        for (int off = 1; off < (NJS < 64 ? NJS : 64); off <<= 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                const float oh = __shfl_xor(D[q].h[k], off);
                const float ol = __shfl_xor(D[q].l[k], off);
                two_sum_add(D[q].h[k], D[q].l[k], oh);
                D[q].l[k] += ol;
            }
            if (NN) {
                const float ob = __shfl_xor(A[q].best, off);
                const int oa = __shfl_xor(A[q].addr, off);
                nn_merge_f(A[q].best, A[q].addr, ob, oa);
            }
        }
    }
    if (NJS > 64) {
        constexpr int W = NJS > 64 ? NJS / 64 : 1;
        __syncthreads();
        float *red = reinterpret_cast<float *>(&tile);  // [wave][IPL][16]
        const int wave = t / 64;
        if ((t & 63) == 0) {
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                float *r = red + (wave * IPL + q) * 16;
#pragma unroll
                for (int k = 0; k < 7; k++) { r[k] = D[q].h[k]; r[7 + k] = D[q].l[k]; }
                r[14] = A[q].best;
                r[15] = __int_as_float(A[q].addr);
            }
        }
        __syncthreads();
        if (js == 0) {
#pragma unroll
            for (int q = 0; q < IPL; q++) {
                for (int w = 1; w < W; w++) {
                    const float *r = red + ((wave + w) * IPL + q) * 16;
#pragma unroll
                    for (int k = 0; k < 7; k++) {
                        two_sum_add(D[q].h[k], D[q].l[k], r[k]);
                        D[q].l[k] += r[7 + k];
                    }
                    if (NN) nn_merge_f(A[q].best, A[q].addr, r[14], __float_as_int(r[15]));
                }
            }
        }
    }
    if (js != 0) return;
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = ibase + il + q * IBL;
        if (i >= p.ni) continue;
        const int S = p.ostride;
        if (nchunk == 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) p.out[k * S + i] = double(D[q].h[k]) + double(D[q].l[k]);
            if (NN) p.nn[i] = A[q].addr >= 0 ? p.jid[A[q].addr] : -1;
        } else {
            // partial record: (hi, lo) float pairs packed into the double slots
            const int PS = p.pstride;
            float2 *r = reinterpret_cast<float2 *>(p.part + size_t(by) * kPartFields * PS);
#pragma unroll
            for (int k = 0; k < 7; k++) r[k * PS + i] = make_float2(D[q].h[k], D[q].l[k]);
            r[7 * PS + i] = make_float2(NN ? A[q].best : __builtin_inff(), 0.0f);
            r[8 * PS + i] = make_float2(__int_as_float(NN ? A[q].addr : -1), 0.0f);
        }
    }
}

// One (i-block bx, j-chunk by) work item; the kernel below runs one item
// per workgroup, kd_* below loops over items whose count lives on the device.
// LONG (with NN): a lane's j in the item number >= kLongItem, so the
// nearest neighbour is a running minimum per block of kFlush j with the slot
// found once at the end; otherwise a compare and two selects per pair, as a
// final search (up to kFlush dependent loads) would dominate a short item.
// Both give the first slot with the lane's minimal r^2.
// PRED: as in force_f64_item.  CHK: as in pair_r2_mx.
template <int NJS, int IPL, bool NN, bool LONG, bool PRED = false, bool CHK = true>
__device__ __forceinline__ void force_mx_item(const FParams &p, TileF &tile, int bx, int by,
                                       int nchunk)
{
    constexpr int NNM = NN ? (LONG ? 2 : 1) : 0;
    constexpr int IBL = kWG / NJS;
    constexpr int IB = IBL * IPL;

    const int t = threadIdx.x;
    const int js = t % NJS;
    const int il = t / NJS;
    const int ibase = bx * IB;

    IF I[IPL];
    AccF A[IPL];
    DSAcc D[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = min(ibase + il + q * IBL, p.ni - 1);
        const double x = p.ix[i], y = p.iy[i], z = p.iz[i];
        I[q].hx = float(x); I[q].lx = float(x - double(I[q].hx));
        I[q].hy = float(y); I[q].ly = float(y - double(I[q].hy));
        I[q].hz = float(z); I[q].lz = float(z - double(I[q].hz));
        I[q].vx = float(p.ivx[i]); I[q].vy = float(p.ivy[i]); I[q].vz = float(p.ivz[i]);
        I[q].id = p.iid[i];
        A[q] = AccF{0, 0, 0, 0, 0, 0, 0, __builtin_inff(), -1, __builtin_inff(), 0, 0};
#pragma unroll
        for (int k = 0; k < 7; k++) { D[q].h[k] = 0.0f; D[q].l[k] = 0.0f; }
    }
    const float eps2 = float(p.eps2);
    const int j0 = by * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);

    for (int tb = j0; tb < j1; tb += kTJ) {
        const int jn = min(kTJ, j1 - tb);
        __syncthreads();
        if (t < jn) {
            const int j = tb + t;
            if constexpr (PRED) {
                const Pred pr = predict_base(p, j);
                const bool empty = p.jid[j] == -1;
                float hi[3], lo[3], fv[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const float h = float(pr.x[k]);
                    hi[k] = empty ? kFarF : h;
                    lo[k] = empty ? 0.0f : float(pr.x[k] - double(h));
                    fv[k] = empty ? 0.0f : float(pr.v[k]);
                }
                tile.a[t] = make_float4(hi[0], hi[1], hi[2], p.fm[j]);
                tile.b[t] = make_float4(lo[0], lo[1], lo[2], fv[0]);
                tile.c[t] = make_float4(fv[1], fv[2], __int_as_float(p.jid[j]), 0.0f);
            } else {
                tile.a[t] = make_float4(p.hx[j], p.hy[j], p.hz[j], p.fm[j]);
                tile.b[t] = make_float4(p.lx[j], p.ly[j], p.lz[j], p.fvx[j]);
                tile.c[t] = make_float4(p.fvy[j], p.fvz[j], __int_as_float(p.jid[j]), 0.0f);
            }
        }
        __syncthreads();
        for (int jb = js; jb < jn; jb += NJS * kFlush) {
            const int je = min(jn, jb + NJS * kFlush);
#pragma unroll G6_MX_UNROLL
            for (int jj = jb; jj < je; jj += NJS) {
                const float4 a = tile.a[jj], b = tile.b[jj], c = tile.c[jj];
#pragma unroll
                for (int q = 0; q < IPL; q++)
                    pair_mx<NNM, CHK>(I[q], A[q], a, b, c, tb + jj, eps2);
            }
#pragma unroll
            for (int q = 0; q < IPL; q++) flush_mx(A[q], D[q]);
            if (NNM == 2) {
                // A block minimum below the best so far (strictly: an equal
                // r^2 keeps the earlier block) moves the best to this block.
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    const bool imp = A[q].tbest < A[q].best;
                    A[q].best = imp ? A[q].tbest : A[q].best;
                    A[q].blo = imp ? tb + jb : A[q].blo;
                    A[q].bhi = imp ? tb + je : A[q].bhi;
                    A[q].tbest = __builtin_inff();
                }
            }
        }
    }
    if (NNM == 2) {
        // The slot: the first j of the best block, in the lane's stride,
        // with exactly that r^2 (recomputed from the same arrays the tiles
        // were loaded from, so bit-identical).  The wave searches one lane's
        // block at a time, lane k taking the block's k-th j (a block holds
        // at most kFlush of a lane's j), so each block is read once,
        // coalesced, and the first match is the lowest set bit of the ballot.
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            for (int L = 0; L < 64; L++) {
                const float best = __shfl(A[q].best, L);
                if (!(best < __builtin_inff())) continue;  // wave-uniform
                const int blo = __shfl(A[q].blo, L), bhi = __shfl(A[q].bhi, L);
                IF ii{};
                ii.hx = __shfl(I[q].hx, L); ii.hy = __shfl(I[q].hy, L);
                ii.hz = __shfl(I[q].hz, L); ii.lx = __shfl(I[q].lx, L);
                ii.ly = __shfl(I[q].ly, L); ii.lz = __shfl(I[q].lz, L);
                ii.id = __shfl(I[q].id, L);
                const int j = blo + lane * NJS;
                bool hit = false;
                if (j < bhi) {
                    const float4 a = make_float4(p.hx[j], p.hy[j], p.hz[j], 0.0f);
                    const float4 b = make_float4(p.lx[j], p.ly[j], p.lz[j], 0.0f);
                    const float4 c = make_float4(0.0f, 0.0f, __int_as_float(p.jid[j]), 0.0f);
                    float dx, dy, dz;
                    hit = pair_r2_mx<CHK>(ii, a, b, c, dx, dy, dz) == best;
                }
                const unsigned long long mask = __ballot(hit);
                if (lane == L && mask) A[q].addr = blo + (__ffsll((long long)mask) - 1) * NJS;
            }
        }
    }

    mx_item_epilogue<NJS, IPL, NN>(p, tile, I, A, D, ibase, il, js, by, nchunk);
}

// ------------------------------------------------------------------------
// Mixed path over spatially ordered slots (full passes, g6x_calc_self): the
// i-block and the j-tiles are runs of slots in Morton order, with the
// bounding box of each tile's positions in p.tbox.  A tile whose box lies at
// least R from the i-block's box (R: the i-box's half-diagonal) is "far":
// it is staged with positions relative to the i-box centre c, rounded once
// to FP32 (x_j - c in double, then float), and the pair takes
// dx = (x_j - c) - (x_i - c) in one FP32 subtraction instead of the three of
// the double-single difference.  Its error is at most 2^-24 (|x_j - c| +
// |x_i - c|) <= 2^-24 (|dx| + 2 R) <= 3 * 2^-24 |dx| per component (|dx| >= R
// there), the order of the double-single path's final rounding.  Other
// tiles take the double-single difference with the id compare.  The slots
// are distinct (the caller's precondition), so a far tile never holds an i
// of the block.
// ------------------------------------------------------------------------
struct IL {
    float x, y, z;  // position relative to the i-box centre
};

template <int NNM>
__device__ __forceinline__ void pair_far(const IL &L, const IF &I, AccF &a, float4 ja, float4 jb,
                                         float eps2)
{
    const float dx = ja.x - L.x, dy = ja.y - L.y, dz = ja.z - L.z;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float s = r2 + eps2;
    const float y = __builtin_amdgcn_rsqf(s);
    if (NNM == 2) a.tbest = fminf(a.tbest, r2);
    const float y2 = y * y;
    const float mi = ja.w * y;
    const float mi3 = mi * y2;
    a.pot -= mi;
    a.ax = __builtin_fmaf(mi3, dx, a.ax);
    a.ay = __builtin_fmaf(mi3, dy, a.ay);
    a.az = __builtin_fmaf(mi3, dz, a.az);
    const float dvx = jb.x - I.vx, dvy = jb.y - I.vy, dvz = jb.z - I.vz;
    const float rv = __builtin_fmaf(dz, dvz, __builtin_fmaf(dy, dvy, dx * dvx));
    const float c = (-3.0f * rv) * y2;
    a.jx = __builtin_fmaf(mi3, __builtin_fmaf(c, dx, dvx), a.jx);
    a.jy = __builtin_fmaf(mi3, __builtin_fmaf(c, dy, dvy), a.jy);
    a.jz = __builtin_fmaf(mi3, __builtin_fmaf(c, dz, dvz), a.jz);
}

// FP32 position of slot j relative to the centre (cx, cy, cz), as staged.
__device__ __forceinline__ float3 local_pos(const FParams &p, int j, double cx, double cy,
                                            double cz)
{
    return make_float3(float((double(p.hx[j]) + double(p.lx[j])) - cx),
                       float((double(p.hy[j]) + double(p.ly[j])) - cy),
                       float((double(p.hz[j]) + double(p.lz[j])) - cz));
}

template <int NJS, int IPL, bool NN, bool LONG>
__device__ __forceinline__ void force_mx_item_sp(const FParams &p, TileF &tile, int bx, int by,
                                                 int nchunk)
{
    constexpr int NNM = NN ? (LONG ? 2 : 1) : 0;
    constexpr int IBL = kWG / NJS;
    constexpr int IB = IBL * IPL;

    const int t = threadIdx.x;
    const int js = t % NJS;
    const int il = t / NJS;
    const int ibase = bx * IB;

    // the i-block's box: union of the boxes of the tiles holding its rows
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    const int last = min(p.ni, ibase + IB) - 1;
    for (int tt = ibase / kTJ; tt <= last / kTJ; tt++) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], p.tbox[6 * tt + k]);
            hi[k] = fmaxf(hi[k], p.tbox[6 * tt + 3 + k]);
        }
    }
    const float cxf = 0.5f * (lo[0] + hi[0]), cyf = 0.5f * (lo[1] + hi[1]),
                czf = 0.5f * (lo[2] + hi[2]);
    const double cx = cxf, cy = cyf, cz = czf;
    float r2box = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float h = 0.5f * (hi[k] - lo[k]);
        r2box = __builtin_fmaf(h, h, r2box);
    }

    IF I[IPL];
    IL L[IPL];
    AccF A[IPL];
    DSAcc D[IPL];
    int bfar[IPL];  // NNM 2: the best block was in a far tile
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = min(ibase + il + q * IBL, p.ni - 1);
        const double x = p.ix[i], y = p.iy[i], z = p.iz[i];
        I[q].hx = float(x); I[q].lx = float(x - double(I[q].hx));
        I[q].hy = float(y); I[q].ly = float(y - double(I[q].hy));
        I[q].hz = float(z); I[q].lz = float(z - double(I[q].hz));
        I[q].vx = float(p.ivx[i]); I[q].vy = float(p.ivy[i]); I[q].vz = float(p.ivz[i]);
        I[q].id = p.iid[i];
        // i-block rows are the slots' own double-single positions summed,
        // so this equals local_pos of the same slot
        L[q].x = float(x - cx); L[q].y = float(y - cy); L[q].z = float(z - cz);
        A[q] = AccF{0, 0, 0, 0, 0, 0, 0, __builtin_inff(), -1, __builtin_inff(), 0, 0};
        bfar[q] = 0;
#pragma unroll
        for (int k = 0; k < 7; k++) { D[q].h[k] = 0.0f; D[q].l[k] = 0.0f; }
    }
    const float eps2 = float(p.eps2);
    const int j0 = by * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);

    for (int tb = j0; tb < j1; tb += kTJ) {
        const int jn = min(kTJ, j1 - tb);
        // box distance: workgroup-uniform
        const int tt = tb / kTJ;
        float d2 = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float g = fmaxf(0.0f, fmaxf(p.tbox[6 * tt + k] - hi[k], lo[k] - p.tbox[6 * tt + 3 + k]));
            d2 = __builtin_fmaf(g, g, d2);
        }
        const bool far = d2 > 0.0f && d2 >= r2box;
        __syncthreads();
        if (t < jn) {
            const int j = tb + t;
            if (far) {
                const float3 f = local_pos(p, j, cx, cy, cz);
                tile.a[t] = make_float4(f.x, f.y, f.z, p.fm[j]);
                tile.b[t] = make_float4(p.fvx[j], p.fvy[j], p.fvz[j], 0.0f);
            } else {
                tile.a[t] = make_float4(p.hx[j], p.hy[j], p.hz[j], p.fm[j]);
                tile.b[t] = make_float4(p.lx[j], p.ly[j], p.lz[j], p.fvx[j]);
                tile.c[t] = make_float4(p.fvy[j], p.fvz[j], __int_as_float(p.jid[j]), 0.0f);
            }
        }
        __syncthreads();
        for (int jb = js; jb < jn; jb += NJS * kFlush) {
            const int je = min(jn, jb + NJS * kFlush);
            if (far) {
#pragma unroll G6_SP_UNROLL
                for (int jj = jb; jj < je; jj += NJS) {
                    const float4 a = tile.a[jj], b = tile.b[jj];
#pragma unroll
                    for (int q = 0; q < IPL; q++) {
                        if (NNM == 1) {
                            // compare on the same r^2 the far pair forms
                            const float dx = a.x - L[q].x, dy = a.y - L[q].y, dz = a.z - L[q].z;
                            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                            const bool better = r2 < A[q].best;
                            A[q].best = better ? r2 : A[q].best;
                            A[q].addr = better ? tb + jj : A[q].addr;
                        }
                        pair_far<NNM == 2 ? 2 : 0>(L[q], I[q], A[q], a, b, eps2);
                    }
                }
            } else {
#pragma unroll G6_SP_UNROLL
                for (int jj = jb; jj < je; jj += NJS) {
                    const float4 a = tile.a[jj], b = tile.b[jj], c = tile.c[jj];
#pragma unroll
                    for (int q = 0; q < IPL; q++)
                        pair_mx<NNM, true>(I[q], A[q], a, b, c, tb + jj, eps2);
                }
            }
#pragma unroll
            for (int q = 0; q < IPL; q++) flush_mx(A[q], D[q]);
            if (NNM == 2) {
#pragma unroll
                for (int q = 0; q < IPL; q++) {
                    const bool imp = A[q].tbest < A[q].best;
                    A[q].best = imp ? A[q].tbest : A[q].best;
                    A[q].blo = imp ? tb + jb : A[q].blo;
                    A[q].bhi = imp ? tb + je : A[q].bhi;
                    bfar[q] = imp ? int(far) : bfar[q];
                    A[q].tbest = __builtin_inff();
                }
            }
        }
    }
    if (NNM == 2) {
        // the slot, as in force_mx_item, with r^2 recomputed the way its
        // tile formed it (relative to c for a far tile)
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int q = 0; q < IPL; q++) {
            for (int Ln = 0; Ln < 64; Ln++) {
                const float best = __shfl(A[q].best, Ln);
                if (!(best < __builtin_inff())) continue;  // wave-uniform
                const int blo = __shfl(A[q].blo, Ln), bhi = __shfl(A[q].bhi, Ln);
                const int isfar = __shfl(bfar[q], Ln);
                const int j = blo + lane * NJS;
                bool hit = false;
                if (isfar) {
                    const float lx = __shfl(L[q].x, Ln), ly = __shfl(L[q].y, Ln),
                                lz = __shfl(L[q].z, Ln);
                    if (j < bhi) {
                        const float3 f = local_pos(p, j, cx, cy, cz);
                        const float dx = f.x - lx, dy = f.y - ly, dz = f.z - lz;
                        hit = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) == best;
                    }
                } else {
                    IF ii{};
                    ii.hx = __shfl(I[q].hx, Ln); ii.hy = __shfl(I[q].hy, Ln);
                    ii.hz = __shfl(I[q].hz, Ln); ii.lx = __shfl(I[q].lx, Ln);
                    ii.ly = __shfl(I[q].ly, Ln); ii.lz = __shfl(I[q].lz, Ln);
                    ii.id = __shfl(I[q].id, Ln);
                    if (j < bhi) {
                        const float4 a = make_float4(p.hx[j], p.hy[j], p.hz[j], 0.0f);
                        const float4 b = make_float4(p.lx[j], p.ly[j], p.lz[j], 0.0f);
                        const float4 c = make_float4(0.0f, 0.0f, __int_as_float(p.jid[j]), 0.0f);
                        float dx, dy, dz;
                        hit = pair_r2_mx<true>(ii, a, b, c, dx, dy, dz) == best;
                    }
                }
                const unsigned long long mask = __ballot(hit);
                if (lane == Ln && mask) A[q].addr = blo + (__ffsll((long long)mask) - 1) * NJS;
            }
        }
    }
    mx_item_epilogue<NJS, IPL, NN>(p, tile, I, A, D, ibase, il, js, by, nchunk);
}

template <int NJS, int IPL, bool NN, bool LONG>
__global__ void __launch_bounds__(kWG) k_force_mx_sp(FParams p)
{
    __shared__ TileF tile;
    force_mx_item_sp<NJS, IPL, NN, LONG>(p, tile, blockIdx.x, blockIdx.y, gridDim.y);
}

template <bool NN>
__global__ void __launch_bounds__(kWG) k_force_mx_p(FParams p)
{
    __shared__ TileF tile;
    force_mx_item<64, 1, NN, false, true>(p, tile, blockIdx.x, 0, 1);
}

template <int NJS, int IPL, bool NN, bool LONG>
__global__ void __launch_bounds__(kWG) k_force_mx(FParams p)
{
    __shared__ TileF tile;
    if (needs_id_check(p, blockIdx.x * (kWG / NJS) * IPL, (kWG / NJS) * IPL, blockIdx.y))
        force_mx_item<NJS, IPL, NN, LONG>(p, tile, blockIdx.x, blockIdx.y, gridDim.y);
    else
        force_mx_item<NJS, IPL, NN, LONG, false, false>(p, tile, blockIdx.x, blockIdx.y,
                                                        gridDim.y);
}

// Items whose lanes see >= kLongItem j keep the nearest neighbour as a
// running minimum (the LONG forms).
__host__ __device__ inline bool long_items(const FParams &p, int njs, bool nn)
{
    return nn && (p.jchunk + njs - 1) / njs >= kLongItem;
}

// ------------------------------------------------------------------------
// Scalar-load variants (one lane per i, NJS = 1): every lane of a wave needs
// the same j, so the j record is read with wave-uniform addresses, which the
// compiler turns into scalar (SMEM) loads into SGPRs that the VALU reads
// directly.  No LDS traffic, no barriers in the hot loop; the workgroup's
// four waves each hold their own 64 (or 128) i-particles.
// ------------------------------------------------------------------------
template <int IPL, bool NN>
__global__ void __launch_bounds__(kWG) k_force_f64_s(FParams p)
{
    constexpr int IB = kWG * IPL;
    const int t = threadIdx.x;
    const int ibase = blockIdx.x * IB;
    ID I[IPL];
    AccD A[IPL];
    DDAcc D[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = min(ibase + t + q * kWG, p.ni - 1);
        I[q].x = p.ix[i]; I[q].y = p.iy[i]; I[q].z = p.iz[i];
        I[q].vx = p.ivx[i]; I[q].vy = p.ivy[i]; I[q].vz = p.ivz[i];
        I[q].id = p.iid[i];
        A[q] = AccD{0, 0, 0, 0, 0, 0, 0, __builtin_inf(), -1, __builtin_inf(), 0, 0};
#pragma unroll
        for (int k = 0; k < 7; k++) { D[q].h[k] = 0.0; D[q].l[k] = 0.0; }
    }
    const double eps2 = p.eps2;
    const int j0 = blockIdx.y * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);
    for (int jb = j0; jb < j1; jb += kTJ) {
        const int je = min(j1, jb + kTJ);
#pragma unroll 4
        for (int j = jb; j < je; j++) {
            const double xj = p.jx[j], yj = p.jy[j], zj = p.jz[j], mj = p.jm[j];
            const double vxj = p.jvx[j], vyj = p.jvy[j], vzj = p.jvz[j];
            const int idj = p.jid[j];
#pragma unroll
            for (int q = 0; q < IPL; q++)
                pair_f64<NN>(I[q], A[q], xj, yj, zj, mj, vxj, vyj, vzj, idj, j, eps2);
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) flush_f64(A[q], D[q]);
    }
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = ibase + t + q * kWG;
        if (i >= p.ni) continue;
        const int S = p.ostride;
        if (gridDim.y == 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) p.out[k * S + i] = D[q].h[k] + D[q].l[k];
            if (NN) p.nn[i] = A[q].addr >= 0 ? p.jid[A[q].addr] : -1;
        } else {
            const int PS = p.pstride;
            double *r = p.part + size_t(blockIdx.y) * kPartFieldsD * PS;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                r[(2 * k) * PS + i] = D[q].h[k];
                r[(2 * k + 1) * PS + i] = D[q].l[k];
            }
            r[14 * PS + i] = NN ? A[q].best : __builtin_inf();
            r[15 * PS + i] = NN ? double(A[q].addr) : -1.0;
        }
    }
}

template <int IPL, bool NN>
__global__ void __launch_bounds__(kWG) k_force_mx_s(FParams p)
{
    constexpr int IB = kWG * IPL;
    const int t = threadIdx.x;
    const int ibase = blockIdx.x * IB;
    IF I[IPL];
    AccF A[IPL];
    DSAcc D[IPL];
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = min(ibase + t + q * kWG, p.ni - 1);
        const double x = p.ix[i], y = p.iy[i], z = p.iz[i];
        I[q].hx = float(x); I[q].lx = float(x - double(I[q].hx));
        I[q].hy = float(y); I[q].ly = float(y - double(I[q].hy));
        I[q].hz = float(z); I[q].lz = float(z - double(I[q].hz));
        I[q].vx = float(p.ivx[i]); I[q].vy = float(p.ivy[i]); I[q].vz = float(p.ivz[i]);
        I[q].id = p.iid[i];
        A[q] = AccF{0, 0, 0, 0, 0, 0, 0, __builtin_inff(), -1, __builtin_inff(), 0, 0};
#pragma unroll
        for (int k = 0; k < 7; k++) { D[q].h[k] = 0.0f; D[q].l[k] = 0.0f; }
    }
    const float eps2 = float(p.eps2);
    const int j0 = blockIdx.y * p.jchunk;
    const int j1 = min(p.nj, j0 + p.jchunk);
    for (int jb = j0; jb < j1; jb += kFlush) {
        const int je = min(j1, jb + kFlush);
#pragma unroll 4
        for (int j = jb; j < je; j++) {
            const float4 a = make_float4(p.hx[j], p.hy[j], p.hz[j], p.fm[j]);
            const float4 b = make_float4(p.lx[j], p.ly[j], p.lz[j], p.fvx[j]);
            const float4 c = make_float4(p.fvy[j], p.fvz[j], __int_as_float(p.jid[j]), 0.0f);
#pragma unroll
            for (int q = 0; q < IPL; q++) pair_mx<NN ? 1 : 0>(I[q], A[q], a, b, c, j, eps2);
        }
#pragma unroll
        for (int q = 0; q < IPL; q++) flush_mx(A[q], D[q]);
    }
#pragma unroll
    for (int q = 0; q < IPL; q++) {
        const int i = ibase + t + q * kWG;
        if (i >= p.ni) continue;
        const int S = p.ostride;
        if (gridDim.y == 1) {
#pragma unroll
            for (int k = 0; k < 7; k++) p.out[k * S + i] = double(D[q].h[k]) + double(D[q].l[k]);
            if (NN) p.nn[i] = A[q].addr >= 0 ? p.jid[A[q].addr] : -1;
        } else {
            const int PS = p.pstride;
            float2 *r = reinterpret_cast<float2 *>(p.part + size_t(blockIdx.y) * kPartFields * PS);
#pragma unroll
            for (int k = 0; k < 7; k++) r[k * PS + i] = make_float2(D[q].h[k], D[q].l[k]);
            r[7 * PS + i] = make_float2(NN ? A[q].best : __builtin_inff(), 0.0f);
            r[8 * PS + i] = make_float2(__int_as_float(NN ? A[q].addr : -1), 0.0f);
        }
    }
}

template <class T>
__device__ __forceinline__ void store_reduced(const T *h, const T *l, int addr, int S, int i,
                                              const int *__restrict__ jid,
                                              double *__restrict__ out, int *__restrict__ nn,
                                              int want_nn)
{
#pragma unroll
    for (int k = 0; k < 7; k++) out[k * S + i] = double(h[k]) + double(l[k]);
    if (want_nn) nn[i] = addr >= 0 ? jid[addr] : -1;
}

template <bool MX>
__device__ __forceinline__ void reduce_one(const double *__restrict__ part, int nchunk, int PS,
                                           int S, int i, bool wave, const int *__restrict__ jid,
                                           double *__restrict__ out, int *__restrict__ nn,
                                           int want_nn)
{
    const int lane = threadIdx.x & 63;
    const int first = wave ? lane : 0, step = wave ? 64 : 1;
    int addr;
    if constexpr (MX) {
        float h[7], l[7], best;
        reduce_i_mx(part, nchunk, PS, i, first, step, h, l, best, addr, want_nn);
        if (wave) butterfly_mx(h, l, best, addr);
        if (!wave || lane == 0) store_reduced(h, l, addr, S, i, jid, out, nn, want_nn);
    } else {
        double h[7], l[7], best;
        reduce_i_f64(part, nchunk, PS, i, first, step, h, l, best, addr, want_nn);
        if (wave) butterfly_f64(h, l, best, addr);
        if (!wave || lane == 0) store_reduced(h, l, addr, S, i, jid, out, nn, want_nn);
    }
}

// Row form for many chunks and few rows: one row per workgroup, the 256
// lanes take every 256th chunk, each wave's sums meet in its lane 0 by the
// butterfly and the four waves' in wave order through LDS (a fixed order, so
// runs repeat bit for bit).  The wave form puts one wave per row on ni / 4
// workgroups, too few CUs to stream the many chunk records of a small
// block.
template <bool MX>
__global__ void __launch_bounds__(256) k_reduce_row(const double *__restrict__ part, int nchunk,
                                                    int PS, int S, const int *__restrict__ jid,
                                                    double *__restrict__ out,
                                                    int *__restrict__ nn, int want_nn)
{
    const int i = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int addr;
    if constexpr (MX) {
        float h[7], l[7], best;
        reduce_i_mx(part, nchunk, PS, i, t, 256, h, l, best, addr, want_nn);
        butterfly_mx(h, l, best, addr);
        __shared__ float sh[4][7], sl[4][7], sb[4];
        __shared__ int sa[4];
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 7; k++) { sh[wave][k] = h[k]; sl[wave][k] = l[k]; }
            sb[wave] = best;
            sa[wave] = addr;
        }
        __syncthreads();
        if (t != 0) return;
        for (int w = 1; w < 4; w++) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                two_sum_add(h[k], l[k], sh[w][k]);
                l[k] += sl[w][k];
            }
            nn_merge_f(best, addr, sb[w], sa[w]);
        }
        store_reduced(h, l, addr, S, i, jid, out, nn, want_nn);
    } else {
        double h[7], l[7], best;
        reduce_i_f64(part, nchunk, PS, i, t, 256, h, l, best, addr, want_nn);
        butterfly_f64(h, l, best, addr);
        __shared__ double sh[4][7], sl[4][7], sb[4];
        __shared__ int sa[4];
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 7; k++) { sh[wave][k] = h[k]; sl[wave][k] = l[k]; }
            sb[wave] = best;
            sa[wave] = addr;
        }
        __syncthreads();
        if (t != 0) return;
        for (int w = 1; w < 4; w++) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                two_sum_add_d(h[k], l[k], sh[w][k]);
                l[k] += sl[w][k];
            }
            nn_merge(best, addr, sb[w], sa[w]);
        }
        store_reduced(h, l, addr, S, i, jid, out, nn, want_nn);
    }
}

template <bool MX, bool WAVE>
__global__ void __launch_bounds__(256) k_reduce(const double *__restrict__ part, int nchunk,
                                                int ni, int PS, int S,
                                                const int *__restrict__ jid,
                                                double *__restrict__ out, int *__restrict__ nn,
                                                int want_nn)
{
    // lane form: one i per thread; wave form: one i per wave, four per workgroup
    const int i = WAVE ? int(blockIdx.x * 4 + (threadIdx.x >> 6))
                       : int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= ni) return;  // wave form: whole waves leave together
    reduce_one<MX>(part, nchunk, PS, S, i, WAVE, jid, out, nn, want_nn);
}

// ------------------------------------------------------------------------
// Device-decided block steps: the kernels read the active count of the step
// from the control block and plan their own work (plan_step), so a captured
// sequence of them serves steps of any size.  Grids are fixed; workgroups
// loop over work items and leave when there are none (ni = 0 or done).
// ------------------------------------------------------------------------
__device__ __forceinline__ double step_time(const StepCtl *c, int at_system)
{
    return at_system ? c->system_time : c->tnext;
}

__global__ void __launch_bounds__(256) kd_predict(JStore s, int nj, const StepCtl *c,
                                                  int at_system, int prec, int compat)
{
    if (c->done) return;
    const double ti = step_time(c, at_system);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x) {
        double x[3], v[3], a2[3], j6[3];
        const double k18[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            x[k] = s.x[k][j]; v[k] = s.v[k][j]; a2[k] = s.a2[k][j]; j6[k] = s.j6[k][j];
        }
        const Pred pr = predict_slot(ti, s.tj[j], x, v, a2, j6, k18, compat, false);
        write_pred(s, j, prec, s.id[j], pr);
    }
}

template <bool MX>
__global__ void __launch_bounds__(kWG) kd_force(FParams p, StepCtl *c, int ni_slot, int tmin_slot)
{
    if (c->done) return;
    // the block is selected: its time slot is free for the step after next
    if (tmin_slot >= 0 && blockIdx.x == 0 && threadIdx.x == 0)
        c->tmin[tmin_slot] = 0x7ff0000000000000ull;
    const int ni = c->ni[ni_slot];
    if (ni <= 0) return;
    const StepPlan sp = plan_step(ni, p.nj);
    p.ni = ni;
    p.jchunk = sp.jchunk;
    p.pstride = ni;
    const int items = sp.iblocks * sp.nchunk;
    if constexpr (MX) {
        __shared__ TileF tile;
        for (int w = blockIdx.x; w < items; w += gridDim.x) {
            const int bx = w % sp.iblocks, by = w / sp.iblocks;
            switch (sp.shape) {
            case 0: force_mx_item<1, 2, false, false>(p, tile, bx, by, sp.nchunk); break;
            case 1: force_mx_item<4, 1, false, false>(p, tile, bx, by, sp.nchunk); break;
            case 2: force_mx_item<16, 1, false, false>(p, tile, bx, by, sp.nchunk); break;
            default: force_mx_item<64, 1, false, false>(p, tile, bx, by, sp.nchunk); break;
            }
        }
    } else {
        __shared__ TileD tile;
        for (int w = blockIdx.x; w < items; w += gridDim.x) {
            const int bx = w % sp.iblocks, by = w / sp.iblocks;
            switch (sp.shape) {
            case 0: force_f64_item<1, 2, false>(p, tile, bx, by, sp.nchunk); break;
            case 1: force_f64_item<4, 1, false>(p, tile, bx, by, sp.nchunk); break;
            case 2: force_f64_item<16, 1, false>(p, tile, bx, by, sp.nchunk); break;
            default: force_f64_item<64, 1, false>(p, tile, bx, by, sp.nchunk); break;
            }
        }
    }
}

template <bool MX>
__global__ void __launch_bounds__(256) kd_reduce(const double *__restrict__ part,
                                                 const StepCtl *c, int nj, int S,
                                                 const int *__restrict__ jid,
                                                 double *__restrict__ out, int ni_slot)
{
    if (c->done) return;
    const int ni = c->ni[ni_slot];
    if (ni <= 0) return;
    const StepPlan sp = plan_step(ni, nj);
    if (sp.nchunk <= 1) return;
    if (sp.nchunk >= kReduceWaveMin && ni <= kReduceWaveMaxNi) {
        // one wave per i; the loop bound is wave-uniform
        for (int i = blockIdx.x * 4 + int(threadIdx.x >> 6); i < ni; i += gridDim.x * 4)
            reduce_one<MX>(part, sp.nchunk, ni, S, i, true, jid, out, nullptr, 0);
    } else {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ni; i += gridDim.x * blockDim.x)
            reduce_one<MX>(part, sp.nchunk, ni, S, i, false, jid, out, nullptr, 0);
    }
}

// ------------------------------------------------------------------------
// Neighbour query (lib/sapporo_light/dev_evaluate_gravity.cu:60-67 criterion
// r^2 <= h2 with self excluded; empty slots excluded as in g6lib.c:138-140).
// One lane per (i, j) pair inside a 64 x 4 tile: the i-index is fast across
// the wave so the j data is a broadcast; r^2 is rounded exactly like the
// reference's so boundary cases agree with it.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ngb(JStore s, IBlock b, int ni, int nj, int prec,
                                             int jper, int *__restrict__ counts,
                                             int *__restrict__ lists)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    if (i >= ni) return;
    const int jsub = threadIdx.x >> 6;
    const double xi = b.x[0][i], yi = b.x[1][i], zi = b.x[2][i], h2 = b.h2[i];
    const int idi = b.id[i];
    const int j0 = blockIdx.y * jper;
    const int j1 = min(nj, j0 + jper);
    for (int j = j0 + jsub; j < j1; j += 4) {
        const int idj = s.pid[j];
        if (idj == -1 || idj == idi) continue;
        double xj, yj, zj;
        if (prec == kF64) {
            xj = s.px[0][j]; yj = s.px[1][j]; zj = s.px[2][j];
        } else {
            xj = double(s.hx[0][j]) + double(s.lx[0][j]);
            yj = double(s.hx[1][j]) + double(s.lx[1][j]);
            zj = double(s.hx[2][j]) + double(s.lx[2][j]);
        }
        const double r2 = exact_r2(xj - xi, yj - yi, zj - zi);
        if (r2 <= h2) {
            const int slot = atomicAdd(&counts[i], 1);
            if (slot < kNgbCap) lists[size_t(i) * kNgbCap + slot] = idj;
        }
    }
}

// ------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------
static inline int nblocks(int n, int b) { return (n + b - 1) / b; }

hipError_t launch_reset(const JStore &s, int lo, int hi, hipStream_t st)
{
    if (hi <= lo) return hipSuccess;
    hipLaunchKernelGGL(k_reset, dim3(nblocks(hi - lo, 256)), dim3(256), 0, st, s, lo, hi);
    return hipGetLastError();
}

hipError_t launch_copy_store(const JStore &dst, const JStore &src, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_store, dim3(nblocks(n, 256)), dim3(256), 0, st, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_scatter(const JStore &s, const JRecord *rec, int n, int prec, int pred_valid,
                          double ti, int compat, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(nblocks(n, 256)), dim3(256), 0, st, s, rec, n, prec,
                       pred_valid, ti, compat);
    return hipGetLastError();
}

hipError_t launch_predict(const JStore &s, int nj, double ti, int prec, int compat, int use_k18,
                          hipStream_t st)
{
    if (nj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict, dim3(nblocks(nj, 256)), dim3(256), 0, st, s, nj, ti, prec,
                       compat, use_k18);
    return hipGetLastError();
}

hipError_t launch_iblock_self(const JStore &s, const IBlock &b, int i0, int ni, int prec,
                              hipStream_t st)
{
    if (ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iblock_self, dim3(nblocks(ni, 256)), dim3(256), 0, st, s, b, i0, ni,
                       prec);
    return hipGetLastError();
}

hipError_t launch_rows_to_aos(const IBlock &b, int ni, double *acc, double *jerk, double *pot,
                              int *nn, hipStream_t st)
{
    if (ni <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rows_to_aos, dim3(nblocks(ni, 256)), dim3(256), 0, st, b.out, b.nn,
                       b.cap, ni, acc, jerk, pot, nn);
    return hipGetLastError();
}

hipError_t launch_ngb(const JStore &s, const IBlock &b, int ni, int nj, int prec, int *counts,
                      int *lists, hipStream_t st)
{
    if (ni <= 0 || nj <= 0) return hipSuccess;
    const int iblocks = nblocks(ni, 64);
    int nchunk = std::max(1, std::min(nblocks(nj, 1024), nblocks(4096, iblocks)));
    const int jper = nblocks(nj, nchunk);
    nchunk = nblocks(nj, jper);
    hipLaunchKernelGGL(k_ngb, dim3(iblocks, nchunk), dim3(256), 0, st, s, b, ni, nj, prec, jper,
                       counts, lists);
    return hipGetLastError();
}

bool force_cfg_supported(int njs, int ipl)
{
    if (ipl == 1) return njs == 1 || njs == 2 || njs == 4 || njs == 8 || njs == 16 ||
                         njs == 32 || njs == 64 || njs == 128 || njs == 256;
    if (ipl == 2) return njs == 1 || njs == 2 || njs == 4 || njs == 8;
    if (ipl == 4) return njs == 1 || njs == 2;
    return false;
}

template <int IPL>
static void launch_force_s(const FParams &fp, int prec, bool nn, dim3 grid, hipStream_t st)
{
    if (prec == kF64) {
        if (nn) hipLaunchKernelGGL((k_force_f64_s<IPL, true>), grid, dim3(kWG), 0, st, fp);
        else hipLaunchKernelGGL((k_force_f64_s<IPL, false>), grid, dim3(kWG), 0, st, fp);
    } else {
        if (nn) hipLaunchKernelGGL((k_force_mx_s<IPL, true>), grid, dim3(kWG), 0, st, fp);
        else hipLaunchKernelGGL((k_force_mx_s<IPL, false>), grid, dim3(kWG), 0, st, fp);
    }
}

template <int NJS, int IPL>
static void launch_force_t(const FParams &fp, int prec, bool nn, dim3 grid, hipStream_t st)
{
    const bool lng = long_items(fp, NJS, nn);
    if (prec == kF64) {
        if (lng) hipLaunchKernelGGL((k_force_f64<NJS, IPL, true, true>), grid, dim3(kWG), 0, st, fp);
        else if (nn) hipLaunchKernelGGL((k_force_f64<NJS, IPL, true, false>), grid, dim3(kWG), 0, st, fp);
        else hipLaunchKernelGGL((k_force_f64<NJS, IPL, false, false>), grid, dim3(kWG), 0, st, fp);
    } else {
        if (lng) hipLaunchKernelGGL((k_force_mx<NJS, IPL, true, true>), grid, dim3(kWG), 0, st, fp);
        else if (nn) hipLaunchKernelGGL((k_force_mx<NJS, IPL, true, false>), grid, dim3(kWG), 0, st, fp);
        else hipLaunchKernelGGL((k_force_mx<NJS, IPL, false, false>), grid, dim3(kWG), 0, st, fp);
    }
}

bool spatial_cfg_supported(const ForceCfg &c)
{
    return c.njs == 1 && (c.ipl == 1 || c.ipl == 2) && !c.smem;
}

template <int IPL>
static void launch_force_sp(const FParams &fp, bool nn, dim3 grid, hipStream_t st)
{
    if (long_items(fp, 1, nn))
        hipLaunchKernelGGL((k_force_mx_sp<1, IPL, true, true>), grid, dim3(kWG), 0, st, fp);
    else if (nn)
        hipLaunchKernelGGL((k_force_mx_sp<1, IPL, true, false>), grid, dim3(kWG), 0, st, fp);
    else
        hipLaunchKernelGGL((k_force_mx_sp<1, IPL, false, false>), grid, dim3(kWG), 0, st, fp);
}

hipError_t launch_force(const JStore &s, const IBlock &b, int ni, int nj, double eps2, int prec,
                        bool want_nn, const ForceCfg &cfg, hipStream_t st, bool pred_j,
                        double ti, int compat, int self_i0, const float *tbox)
{
    if (ni <= 0) return hipSuccess;
    FParams fp{};
    fp.self_i0 = pred_j ? -1 : self_i0;
    fp.tbox = tbox;
    fp.ix = b.x[0]; fp.iy = b.x[1]; fp.iz = b.x[2];
    fp.ivx = b.v[0]; fp.ivy = b.v[1]; fp.ivz = b.v[2];
    fp.iid = b.id;
    fp.ni = ni;
    fp.jx = s.px[0]; fp.jy = s.px[1]; fp.jz = s.px[2]; fp.jm = s.pm;
    fp.jvx = s.pv[0]; fp.jvy = s.pv[1]; fp.jvz = s.pv[2];
    fp.hx = s.hx[0]; fp.hy = s.hx[1]; fp.hz = s.hx[2];
    fp.lx = s.lx[0]; fp.ly = s.lx[1]; fp.lz = s.lx[2];
    fp.fm = s.fm; fp.fvx = s.fv[0]; fp.fvy = s.fv[1]; fp.fvz = s.fv[2];
    fp.jid = s.pid;
    fp.nj = nj;
    fp.jchunk = cfg.jchunk;
    fp.eps2 = eps2;
    fp.out = b.out;
    fp.nn = b.nn;
    fp.ostride = b.cap;
    fp.pstride = ni;  // partials packed: nchunk * ni records
    fp.part = b.part;
    const int ib = (kWG / cfg.njs) * cfg.ipl;
    dim3 grid(nblocks(ni, ib), cfg.nchunk);
    if (pred_j) {
        if (!pred_force_cfg(cfg, nj)) return hipErrorInvalidValue;
        for (int k = 0; k < 3; k++) {
            fp.bx[k] = s.x[k]; fp.bv[k] = s.v[k]; fp.ba2[k] = s.a2[k]; fp.bj6[k] = s.j6[k];
        }
        fp.btj = s.tj;
        fp.tpred = ti;
        fp.compat = compat;
        if (prec == kF64) {
            if (want_nn) hipLaunchKernelGGL((k_force_f64_p<true>), grid, dim3(kWG), 0, st, fp);
            else hipLaunchKernelGGL((k_force_f64_p<false>), grid, dim3(kWG), 0, st, fp);
        } else {
            if (want_nn) hipLaunchKernelGGL((k_force_mx_p<true>), grid, dim3(kWG), 0, st, fp);
            else hipLaunchKernelGGL((k_force_mx_p<false>), grid, dim3(kWG), 0, st, fp);
        }
        return hipGetLastError();
    }
    if (tbox) {
        // spatially ordered full pass, mixed path (the caller checked the shape)
        if (prec != kMixed || !spatial_cfg_supported(cfg)) return hipErrorInvalidValue;
        if (cfg.ipl == 1) launch_force_sp<1>(fp, want_nn, grid, st);
        else launch_force_sp<2>(fp, want_nn, grid, st);
        return hipGetLastError();
    }
    if (cfg.smem) {
        if (cfg.njs != 1) return hipErrorInvalidValue;
        if (cfg.ipl == 1) launch_force_s<1>(fp, prec, want_nn, grid, st);
        else if (cfg.ipl == 2) launch_force_s<2>(fp, prec, want_nn, grid, st);
        else return hipErrorInvalidValue;
        return hipGetLastError();
    }
    const int key = cfg.njs * 10 + cfg.ipl;
    switch (key) {
    case 11: launch_force_t<1, 1>(fp, prec, want_nn, grid, st); break;
    case 21: launch_force_t<2, 1>(fp, prec, want_nn, grid, st); break;
    case 41: launch_force_t<4, 1>(fp, prec, want_nn, grid, st); break;
    case 81: launch_force_t<8, 1>(fp, prec, want_nn, grid, st); break;
    case 161: launch_force_t<16, 1>(fp, prec, want_nn, grid, st); break;
    case 321: launch_force_t<32, 1>(fp, prec, want_nn, grid, st); break;
    case 641: launch_force_t<64, 1>(fp, prec, want_nn, grid, st); break;
    case 1281: launch_force_t<128, 1>(fp, prec, want_nn, grid, st); break;
    case 2561: launch_force_t<256, 1>(fp, prec, want_nn, grid, st); break;
    case 12: launch_force_t<1, 2>(fp, prec, want_nn, grid, st); break;
    case 22: launch_force_t<2, 2>(fp, prec, want_nn, grid, st); break;
    case 42: launch_force_t<4, 2>(fp, prec, want_nn, grid, st); break;
    case 82: launch_force_t<8, 2>(fp, prec, want_nn, grid, st); break;
    case 14: launch_force_t<1, 4>(fp, prec, want_nn, grid, st); break;
    case 24: launch_force_t<2, 4>(fp, prec, want_nn, grid, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reduce(const JStore &s, const IBlock &b, int ni, int nchunk, int prec,
                         bool want_nn, hipStream_t st)
{
    if (ni <= 0 || nchunk <= 1) return hipSuccess;
    const int S = b.cap;
    if (nchunk >= kReduceRowMin && ni <= kReduceRowMaxNi) {
        if (prec == kF64)
            hipLaunchKernelGGL((k_reduce_row<false>), dim3(ni), dim3(256), 0, st, b.part, nchunk,
                               ni, S, s.pid, b.out, b.nn, int(want_nn));
        else
            hipLaunchKernelGGL((k_reduce_row<true>), dim3(ni), dim3(256), 0, st, b.part, nchunk,
                               ni, S, s.pid, b.out, b.nn, int(want_nn));
        return hipGetLastError();
    }
    if (nchunk >= kReduceWaveMin && ni <= kReduceWaveMaxNi) {
        if (prec == kF64)
            hipLaunchKernelGGL((k_reduce<false, true>), dim3(nblocks(ni, 4)), dim3(256), 0, st,
                               b.part, nchunk, ni, ni, S, s.pid, b.out, b.nn, int(want_nn));
        else
            hipLaunchKernelGGL((k_reduce<true, true>), dim3(nblocks(ni, 4)), dim3(256), 0, st,
                               b.part, nchunk, ni, ni, S, s.pid, b.out, b.nn, int(want_nn));
        return hipGetLastError();
    }
    if (prec == kF64)
        hipLaunchKernelGGL((k_reduce<false, false>), dim3(nblocks(ni, 256)), dim3(256), 0, st,
                           b.part, nchunk, ni, ni, S, s.pid, b.out, b.nn, int(want_nn));
    else
        hipLaunchKernelGGL((k_reduce<true, false>), dim3(nblocks(ni, 256)), dim3(256), 0, st,
                           b.part, nchunk, ni, ni, S, s.pid, b.out, b.nn, int(want_nn));
    return hipGetLastError();
}

hipError_t launch_step_predict(const JStore &s, int nj, const StepCtl *c, int at_system,
                               int prec, int compat, hipStream_t st)
{
    if (nj <= 0) return hipSuccess;
    const int blocks = std::min(nblocks(nj, 256), 4096);
    hipLaunchKernelGGL(kd_predict, dim3(blocks), dim3(256), 0, st, s, nj, c, at_system, prec,
                       compat);
    return hipGetLastError();
}

hipError_t launch_step_force(const JStore &s, const IBlock &b, double *part, int nj,
                             double eps2, int prec, StepCtl *c, int ni_slot, int tmin_slot,
                             bool with_reduce, hipStream_t st)
{
    if (nj <= 0) return hipSuccess;
    FParams fp{};
    fp.ix = b.x[0]; fp.iy = b.x[1]; fp.iz = b.x[2];
    fp.ivx = b.v[0]; fp.ivy = b.v[1]; fp.ivz = b.v[2];
    fp.iid = b.id;
    fp.jx = s.px[0]; fp.jy = s.px[1]; fp.jz = s.px[2]; fp.jm = s.pm;
    fp.jvx = s.pv[0]; fp.jvy = s.pv[1]; fp.jvz = s.pv[2];
    fp.hx = s.hx[0]; fp.hy = s.hx[1]; fp.hz = s.hx[2];
    fp.lx = s.lx[0]; fp.ly = s.lx[1]; fp.lz = s.lx[2];
    fp.fm = s.fm; fp.fvx = s.fv[0]; fp.fvy = s.fv[1]; fp.fvz = s.fv[2];
    fp.jid = s.pid;
    fp.nj = nj;
    fp.eps2 = eps2;
    fp.out = b.out;
    fp.nn = b.nn;
    fp.ostride = b.cap;
    fp.part = part;
    // Grids sized for the largest step of this system (ni <= nj): the
    // narrowest shape puts 4 i in a workgroup, so nj / 4 workgroups cover a
    // whole-system step; surplus workgroups would only cost dispatch time on
    // small systems, and larger plans loop over their items.
    const int gf = std::min(kStepGrid, std::max(64, nblocks(nj, 4)));
    const int gr = std::min(kStepGrid, std::max(1, nblocks(nj, 64)));
    if (prec == kF64) {
        hipLaunchKernelGGL((kd_force<false>), dim3(gf), dim3(kWG), 0, st, fp, c, ni_slot,
                           tmin_slot);
        if (with_reduce)
            hipLaunchKernelGGL((kd_reduce<false>), dim3(gr), dim3(256), 0, st, part, c, nj,
                               b.cap, s.pid, b.out, ni_slot);
    } else {
        hipLaunchKernelGGL((kd_force<true>), dim3(gf), dim3(kWG), 0, st, fp, c, ni_slot,
                           tmin_slot);
        if (with_reduce)
            hipLaunchKernelGGL((kd_reduce<true>), dim3(gr), dim3(256), 0, st, part, c, nj,
                               b.cap, s.pid, b.out, ni_slot);
    }
    return hipGetLastError();
}

}  // namespace g6k

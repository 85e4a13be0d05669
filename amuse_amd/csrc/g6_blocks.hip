// Device-resident block-step Hermite integrator: ph4's integration loop with
// the scheduler, the i-particle predictor, the force call and the corrector
// all on the GPU, so a block step needs no host round trip.
//
//   scheduler          next block time = min(t + dt), the block = particles
//                      with t + dt equal to it (src/amuse_ph4/src/
//                      scheduler.cc:66-140, jdata::advance jdata.cc:752-793)
//   i-predictor        idata::predict (idata.cc:341-360), h4::predict_ph4
//   j-predictor        the GRAPE-6 predictor (lib/g6lib/g6lib.c:64-92) over
//                      the j-store, as the library does at each set_ti
//   forces             the library's force kernels, launch shape decided on
//                      the device from the block size (g6k::plan_step)
//   corrector          idata::correct (idata.cc:431-512), h4::correct_ph4,
//                      then the corrected particles reloaded into the j-store
//                      with a2 = acc/2, j6 = jerk/6 (gpu.cc:163-230)
//
// Initialisation follows the host driver (amuse_amd/csrc/hermite4.cpp):
// j-store loaded with acc = jerk = 0 (ph4's idata::setup does not reload it
// after the first force pass), full-N force, first steps with the median
// limit (jdata.cc:503-564) computed on the host from the downloaded
// acc/jerk, the same arithmetic as the host driver (h4_rules.h).
//
// One step is three kernels: selection (end test, block time, the block
// with its i-prediction, the j-prediction, and the next block time's
// minimum over the particles not in the block), force, and correction (the
// fold of the force partials, the corrector, the new t + dt into the next
// minimum); the control-block protocol is described with g6k::StepCtl.
// g6d_run captures `steps_per_graph` of them (rounded up to even: the step
// parity is a launch argument) into a hipGraph and replays it, reading the
// control block (done flag, counters) back after each replay.  Kernels of a
// finished run return at once, so a replay takes no step after the end
// condition (system time >= t_end, or the block budget spent).
#include "../../include/g6x.h"
#include "g6_kernels.h"
#include "g6_predict.h"
#include "g6_reduce.h"
#include "h4_rules.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#define G6D_API extern "C" __attribute__((visibility("default")))

namespace g6b {

using g6k::IBlock;
using g6k::JStore;
using g6k::StepCtl;

struct State {
    int n = 0;
    int *id = nullptr, *list = nullptr;
    double *m = nullptr, *pot = nullptr, *t = nullptr, *dt = nullptr;
    double *x[3] = {}, *v[3] = {}, *a[3] = {}, *j[3] = {};
};

struct Driver {
    State st;
    JStore js;
    IBlock ib;
    double *part = nullptr;
    void *state_base = nullptr;
    StepCtl *ctl = nullptr;    // device
    StepCtl *h_ctl = nullptr;  // pinned host mirror
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0, graph_prec = -1;
    int prec = 0, compat = 0;
    double eps2 = 0.0, eta = 0.01, dt_fac = 0.0625, dt_limit = 0.03125, dt_limitm = 8.0;
    bool ph4_initial_load = true;
    std::string err;
};

#define DCK(expr)                                                       \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) {                                         \
            D->err = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return G6X_ERR_HIP;                                         \
        }                                                               \
    } while (0)

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// ---------------------------------------------------------------------------
// step kernels
// ---------------------------------------------------------------------------
constexpr unsigned long long kInfBits = 0x7ff0000000000000ull;  // +inf

// Folds a lane's candidate next block time into *slot: wave shuffles, then
// one 64-bit atomicMin per wave on the order-preserving bits of the
// (positive) times.
__device__ __forceinline__ void fold_min(double lo, unsigned long long *slot)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) lo = fmin(lo, __shfl_xor(lo, off));
    if ((threadIdx.x & 63) == 0 && lo < __builtin_inf())
        atomicMin(slot, static_cast<unsigned long long>(__double_as_longlong(lo)));
}

// Control block for a fresh run of steps: no block time yet, counts zero.
__global__ void kb_ctl_reset(StepCtl *c)
{
    c->tmin[0] = c->tmin[1] = kInfBits;
    c->ni[0] = c->ni[1] = c->ni[2] = 0;
    c->cur = 0;
    c->done = 0;
}

// First block time after the initial steps: min over particles of t + dt.
__global__ void __launch_bounds__(256) kb_first_min(const double *__restrict__ t,
                                                    const double *__restrict__ dt, int n,
                                                    StepCtl *c)
{
    double lo = __builtin_inf();
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        lo = fmin(lo, t[j] + dt[j]);
    fold_min(lo, &c->tmin[0]);
}

// Step k, parity par = k & 1.  Ends the run when the system time has reached
// t_end (ph4's evolve_model loop runs while system_time < t_end,
// src/amuse_ph4/interface.cc:673-674) or the block budget is spent; every
// thread reads the same system_time, t_end and budget, which no kernel of
// this launch writes.  Otherwise the block: particles with t + dt equal to
// the block time tmin[par], appended to ni[par] in any order (each i's
// force and correction are independent of the order) and predicted to it
// into the i-block rows; every j-slot is predicted to it as well (the
// library's predictor at set_ti); the other particles' t + dt, unchanged by
// this step, go into the next step's minimum tmin[par ^ 1].
//
// sync = 1 (g6d_synchronize): ph4's synchronize_all (jdata.cc:874-900 with
// idata::set_list_sync, idata.cc:773-785) -- the block is every particle
// behind the system time, stepped to it; t_end and the budget do not apply.
__global__ void __launch_bounds__(256) kb_select(State s, JStore js, IBlock b, StepCtl *c,
                                                 int prec, int compat, int par, int sync)
{
    if (c->done) return;
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (!sync && !(c->system_time < c->t_end)) {
        if (first) c->done = 1;  // t_end reached
        return;
    }
    if (!sync && c->budget == 0) {
        if (first) c->done = 2;  // block budget spent
        return;
    }
    const double tnext = sync ? c->system_time
                              : __longlong_as_double(static_cast<long long>(c->tmin[par]));
    if (first) c->tnext = tnext;
    double lo = __builtin_inf();
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < s.n; j += gridDim.x * blockDim.x) {
        {
            double x[3], v[3], a2[3], j6[3];
            const double k18[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                x[k] = js.x[k][j]; v[k] = js.v[k][j]; a2[k] = js.a2[k][j]; j6[k] = js.j6[k][j];
            }
            const g6k::Pred pr = g6k::predict_slot(tnext, js.tj[j], x, v, a2, j6, k18, compat,
                                                   false);
            g6k::write_pred(js, j, prec, js.id[j], pr);
        }
        const double tj = s.t[j] + s.dt[j];
        if (sync ? !(s.t[j] < tnext) : tj != tnext) {
            lo = fmin(lo, tj);
            continue;
        }
        const int q = atomicAdd(&c->ni[par], 1);
        s.list[q] = j;
        double x[3], v[3], a[3], jk[3], xp[3], vp[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            x[k] = s.x[k][j]; v[k] = s.v[k][j]; a[k] = s.a[k][j]; jk[k] = s.j[k][j];
        }
        h4::predict_ph4(x, v, a, jk, tnext - s.t[j], xp, vp);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            b.x[k][q] = xp[k];
            b.v[k][q] = vp[k];
        }
        b.id[q] = s.id[j];
    }
    fold_min(lo, &c->tmin[par ^ 1]);
}

// Every particle as the i-block, predicted to the system time (energy and
// initial force passes).
__global__ void __launch_bounds__(256) kb_select_all(State s, IBlock b, StepCtl *c)
{
    const double t0 = c->system_time;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < s.n; j += gridDim.x * blockDim.x) {
        if (j == 0) {
            c->ni[2] = s.n;
            c->done = 0;
        }
        s.list[j] = j;
        double x[3], v[3], a[3], jk[3], xp[3], vp[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            x[k] = s.x[k][j]; v[k] = s.v[k][j]; a[k] = s.a[k][j]; jk[k] = s.j[k][j];
        }
        h4::predict_ph4(x, v, a, jk, t0 - s.t[j], xp, vp);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            b.x[k][j] = xp[k];
            b.v[k][j] = vp[k];
        }
        b.id[j] = s.id[j];
    }
}

// j-store base record of particle j from its state (a2 = acc/2, j6 = jerk/6,
// k18 = 0; gpu.cc:163-230); zero_deriv keeps acc = jerk = 0 (the initial load).
__device__ __forceinline__ void load_j(const State &s, const JStore &js, int j, int zero_deriv)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        js.x[k][j] = s.x[k][j];
        js.v[k][j] = s.v[k][j];
        js.a2[k][j] = zero_deriv ? 0.0 : s.a[k][j] / 2;
        js.j6[k][j] = zero_deriv ? 0.0 : s.j[k][j] / 6;
        js.k18[k][j] = 0.0;
    }
    js.tj[j] = s.t[j];
    js.dtj[j] = s.dt[j];
    js.m[j] = s.m[j];
    js.id[j] = s.id[j];
    // the force kernels' mass and id (an id of -1 marks an empty slot)
    const double mp = s.id[j] == -1 ? 0.0 : s.m[j];
    js.pm[j] = mp;
    js.fm[j] = float(mp);
    js.pid[j] = s.id[j];
}

__global__ void __launch_bounds__(256) kb_load_all(State s, JStore js, int zero_deriv)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < s.n; j += gridDim.x * blockDim.x)
        load_j(s, js, j, zero_deriv);
}

// Initial pass: the forces of every particle become its acc, jerk, pot.
__global__ void __launch_bounds__(256) kb_take_all(State s, IBlock b)
{
    const int S = b.cap;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < s.n; j += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s.a[k][j] = b.out[k * S + j];
            s.j[k][j] = b.out[(3 + k) * S + j];
        }
        s.pot[j] = b.out[6 * S + j];
    }
}

// Corrects particle list[q] with the forces f (acc3 jerk3 pot), reloads it
// into the j-store and returns its new t + dt.
__device__ __forceinline__ double correct_one(const State &s, const JStore &js, const IBlock &b,
                                              int q, double tnext, double eta, const double *f)
{
    const int j = s.list[q];
    double oa[3], oj[3], ia[3], ij[3], ip[3], iv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        oa[k] = s.a[k][j];
        oj[k] = s.j[k][j];
        ia[k] = f[k];
        ij[k] = f[3 + k];
        ip[k] = b.x[k][q];
        iv[k] = b.v[k][q];
    }
    const double newstep = h4::correct_ph4(tnext, s.t[j], s.dt[j], eta, oa, oj, ia, ij, ip, iv);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s.x[k][j] = ip[k];
        s.v[k][j] = iv[k];
        s.a[k][j] = ia[k];
        s.j[k][j] = ij[k];
    }
    s.pot[j] = f[6];
    s.t[j] = tnext;
    s.dt[j] = newstep;
    load_j(s, js, j, 0);
    return s.t[j] + s.dt[j];
}

// Forces of block row q folded from the j-chunk partials (the order and
// arithmetic of the library's reduce kernel: lane form, or lane `lane` of the
// wave form followed by the butterfly; the result is valid in lane 0).
template <bool MX>
__device__ __forceinline__ void fold_forces(const double *part, int nchunk, int ni, int q,
                                            bool wave, double *f)
{
    const int lane = threadIdx.x & 63;
    const int first = wave ? lane : 0, step = wave ? 64 : 1;
    int addr;
    if constexpr (MX) {
        float h[7], l[7], best;
        g6k::reduce_i_mx(part, nchunk, ni, q, first, step, h, l, best, addr, 0);
        if (wave) g6k::butterfly_mx(h, l, best, addr);
#pragma unroll
        for (int k = 0; k < 7; k++) f[k] = double(h[k]) + double(l[k]);
    } else {
        double h[7], l[7], best;
        g6k::reduce_i_f64(part, nchunk, ni, q, first, step, h, l, best, addr, 0);
        if (wave) g6k::butterfly_f64(h, l, best, addr);
#pragma unroll
        for (int k = 0; k < 7; k++) f[k] = h[k] + l[k];
    }
}

// Step k, parity par: the block's forces (from b.out, or folded here from
// the partials when the plan split j over workgroups), ph4's corrector, the
// new t + dt into tmin[par ^ 1]; the first thread advances the counters and
// clears ni[par ^ 1] for the next step's selection.
template <bool MX>
__global__ void __launch_bounds__(256) kb_correct(State s, JStore js, IBlock b,
                                                  const double *__restrict__ part, StepCtl *c,
                                                  double eta, int par)
{
    if (c->done) return;
    const double tnext = c->tnext;
    const int ni = c->ni[par];
    const int S = b.cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->blocks += 1;
        c->psteps += ni;
        c->system_time = tnext;
        if (c->budget > 0) c->budget -= 1;
        c->ni[par ^ 1] = 0;
        c->cur = par ^ 1;
    }
    const g6k::StepPlan sp = g6k::plan_step(ni, s.n);
    double lo = __builtin_inf();
    double f[7];
    if (sp.nchunk > 1 && sp.nchunk >= g6k::kReduceWaveMin && ni <= g6k::kReduceWaveMaxNi) {
        // one wave per row; the loop bound is wave-uniform
        for (int q = blockIdx.x * 4 + int(threadIdx.x >> 6); q < ni; q += gridDim.x * 4) {
            fold_forces<MX>(part, sp.nchunk, ni, q, true, f);
            if ((threadIdx.x & 63) == 0) lo = fmin(lo, correct_one(s, js, b, q, tnext, eta, f));
        }
    } else {
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ni; q += gridDim.x * blockDim.x) {
            if (sp.nchunk > 1) {
                fold_forces<MX>(part, sp.nchunk, ni, q, false, f);
            } else {
#pragma unroll
                for (int k = 0; k < 7; k++) f[k] = b.out[k * S + q];
            }
            lo = fmin(lo, correct_one(s, js, b, q, tnext, eta, f));
        }
    }
    fold_min(lo, &c->tmin[par ^ 1]);
}

// Arms a run: end time and block budget; the pending block time moves to
// slot 0, where the captured sequence (which starts at parity 0) reads it.
__global__ void kb_arm(StepCtl *c, double t_end, long long budget)
{
    c->t_end = t_end;
    c->budget = budget;
    c->done = 0;
    if (c->cur == 1) {
        c->tmin[0] = c->tmin[1];
        c->tmin[1] = kInfBits;
        c->ni[0] = 0;
        c->cur = 0;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr int kGrid = 1024;  // workgroups of the per-particle kernels

int enqueue_step(Driver *D, int par, int sync = 0)
{
    const int n = D->st.n;
    const int g = std::min(nblk(n, 256), kGrid);
    hipLaunchKernelGGL(kb_select, dim3(g), dim3(256), 0, D->stream, D->st, D->js, D->ib, D->ctl,
                       D->prec, D->compat, par, sync);
    DCK(g6k::launch_step_force(D->js, D->ib, D->part, n, D->eps2, D->prec, D->ctl, par, par,
                               false, D->stream));
    // enough waves for one block row each when the partials are folded per
    // wave (up to n rows, four waves per workgroup)
    const int gc = std::min(g6k::kStepGrid, std::max(g, nblk(n, 4)));
    if (D->prec == G6X_MIXED)
        hipLaunchKernelGGL(kb_correct<true>, dim3(gc), dim3(256), 0, D->stream, D->st, D->js,
                           D->ib, D->part, D->ctl, D->eta, par);
    else
        hipLaunchKernelGGL(kb_correct<false>, dim3(gc), dim3(256), 0, D->stream, D->st, D->js,
                           D->ib, D->part, D->ctl, D->eta, par);
    DCK(hipGetLastError());
    return 0;
}

// Forces on every particle at the system time into ib.out (precision prec).
int full_pass(Driver *D, int prec)
{
    const int n = D->st.n;
    const int g = std::min(nblk(n, 256), kGrid);
    hipLaunchKernelGGL(kb_select_all, dim3(g), dim3(256), 0, D->stream, D->st, D->ib, D->ctl);
    DCK(g6k::launch_step_predict(D->js, n, D->ctl, 1, prec, D->compat, D->stream));
    DCK(g6k::launch_step_force(D->js, D->ib, D->part, n, D->eps2, prec, D->ctl, 2, -1, true,
                               D->stream));
    DCK(hipGetLastError());
    return 0;
}

int read_ctl(Driver *D)
{
    DCK(hipMemcpyAsync(D->h_ctl, D->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, D->stream));
    DCK(hipStreamSynchronize(D->stream));
    return 0;
}

void release(Driver *D)
{
    if (D->stream) (void)hipStreamSynchronize(D->stream);
    if (D->graph) (void)hipGraphExecDestroy(D->graph);
    if (D->state_base) (void)hipFree(D->state_base);
    if (D->js.base) (void)hipFree(D->js.base);
    if (D->ib.base) (void)hipFree(D->ib.base);
    if (D->part) (void)hipFree(D->part);
    if (D->ctl) (void)hipFree(D->ctl);
    if (D->h_ctl) (void)hipHostFree(D->h_ctl);
    if (D->stream) (void)hipStreamDestroy(D->stream);
    delete D;
}

int setup(Driver *D, int n, const int *ids, const double *m, const double *x, const double *v)
{
    DCK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
    // state: id, list (ints) and m pot t dt x3 v3 a3 j3 (doubles), each
    // array starting on a 256-byte boundary (up to 255 bytes of padding
    // after each of the 18)
    const size_t nd = size_t(n);
    const size_t bytes = nd * sizeof(double) * 16 + nd * sizeof(int) * 2 + 19 * 256;
    DCK(hipMalloc(&D->state_base, bytes));
    char *p = static_cast<char *>(D->state_base);
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 255) / 256 * 256;
        return r;
    };
    State &s = D->st;
    s.n = n;
    s.m = reinterpret_cast<double *>(take(nd * 8));
    s.pot = reinterpret_cast<double *>(take(nd * 8));
    s.t = reinterpret_cast<double *>(take(nd * 8));
    s.dt = reinterpret_cast<double *>(take(nd * 8));
    for (int k = 0; k < 3; k++) s.x[k] = reinterpret_cast<double *>(take(nd * 8));
    for (int k = 0; k < 3; k++) s.v[k] = reinterpret_cast<double *>(take(nd * 8));
    for (int k = 0; k < 3; k++) s.a[k] = reinterpret_cast<double *>(take(nd * 8));
    for (int k = 0; k < 3; k++) s.j[k] = reinterpret_cast<double *>(take(nd * 8));
    s.id = reinterpret_cast<int *>(take(nd * 4));
    s.list = reinterpret_cast<int *>(take(nd * 4));
    if (size_t(p - static_cast<char *>(D->state_base)) > bytes) {
        D->err = "state layout";
        return G6X_ERR_HIP;
    }
    // j-store, i-block (capacity n, rounded to a power of two >= 1024), partials
    int jcap = 1024;
    while (jcap < n) jcap *= 2;
    void *jb = nullptr, *ibase = nullptr;
    DCK(hipMalloc(&jb, g6k::jstore_bytes(jcap)));
    g6k::jstore_carve(D->js, jb, jcap);
    DCK(g6k::launch_reset(D->js, 0, jcap, D->stream));
    DCK(hipMalloc(&ibase, g6k::iblock_bytes(jcap)));
    g6k::iblock_carve(D->ib, ibase, jcap);
    const long long max_chunks = std::max(1, nblk(n, g6k::kTJ));
    const long long part_entries = std::min(g6k::kStepPartBudget, max_chunks * (long long)n);
    DCK(hipMalloc(reinterpret_cast<void **>(&D->part),
                  sizeof(double) * g6k::kPartFieldsD * size_t(part_entries)));
    DCK(hipMalloc(reinterpret_cast<void **>(&D->ctl), sizeof(StepCtl)));
    DCK(hipHostMalloc(reinterpret_cast<void **>(&D->h_ctl), sizeof(StepCtl)));
    // host state in, times zero
    std::vector<double> zeros(nd, 0.0), tmp(nd);
    DCK(hipMemcpyAsync(s.id, ids, nd * 4, hipMemcpyHostToDevice, D->stream));
    DCK(hipMemcpyAsync(s.m, m, nd * 8, hipMemcpyHostToDevice, D->stream));
    for (int k = 0; k < 3; k++) {
        for (int i = 0; i < n; i++) tmp[i] = x[3 * i + k];
        DCK(hipMemcpyAsync(s.x[k], tmp.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
        DCK(hipStreamSynchronize(D->stream));
        for (int i = 0; i < n; i++) tmp[i] = v[3 * i + k];
        DCK(hipMemcpyAsync(s.v[k], tmp.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
        DCK(hipStreamSynchronize(D->stream));
        DCK(hipMemcpyAsync(s.a[k], zeros.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
        DCK(hipMemcpyAsync(s.j[k], zeros.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
    }
    DCK(hipMemcpyAsync(s.t, zeros.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
    DCK(hipMemcpyAsync(s.dt, zeros.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
    DCK(hipMemcpyAsync(s.pot, zeros.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
    StepCtl c0{};
    c0.tmin[0] = c0.tmin[1] = g6b::kInfBits;
    c0.budget = -1;
    c0.t_end = 0.0;
    c0.system_time = 0.0;
    DCK(hipMemcpyAsync(D->ctl, &c0, sizeof(StepCtl), hipMemcpyHostToDevice, D->stream));
    DCK(hipStreamSynchronize(D->stream));
    return 0;
}

// Full-N force at t = 0, first steps (host, as the host driver), j-store load.
int commit(Driver *D)
{
    State &s = D->st;
    const int n = s.n;
    const int g = std::min(nblk(n, 256), kGrid);
    // j-set with acc = jerk = 0 (the state's derivatives are still zero)
    hipLaunchKernelGGL(kb_load_all, dim3(g), dim3(256), 0, D->stream, s, D->js, 1);
    if (full_pass(D, D->prec)) return G6X_ERR_HIP;
    hipLaunchKernelGGL(kb_take_all, dim3(g), dim3(256), 0, D->stream, s, D->ib);
    DCK(hipGetLastError());
    // first steps on the host from acc and jerk (jdata.cc:503-564)
    const size_t nd = size_t(n);
    std::vector<double> acc(3 * nd), jerk(3 * nd), col(nd), step(nd);
    for (int k = 0; k < 3; k++) {
        DCK(hipMemcpyAsync(col.data(), s.a[k], nd * 8, hipMemcpyDeviceToHost, D->stream));
        DCK(hipStreamSynchronize(D->stream));
        for (int i = 0; i < n; i++) acc[3 * i + k] = col[i];
        DCK(hipMemcpyAsync(col.data(), s.j[k], nd * 8, hipMemcpyDeviceToHost, D->stream));
        DCK(hipStreamSynchronize(D->stream));
        for (int i = 0; i < n; i++) jerk[3 * i + k] = col[i];
    }
    for (int i = 0; i < n; i++)
        step[i] = h4::first_step_ph4(&acc[3 * i], &jerk[3 * i], D->eta, D->dt_fac, D->dt_limit,
                                     0.0);
    if (D->dt_limitm > 0 && n > 0) {
        std::vector<double> srt(step);
        std::sort(srt.begin(), srt.end());
        const double dtmax = D->dt_limitm * srt[n / 2];
        for (int i = 0; i < n; i++)
            for (int it = 0; it < 2100 && step[i] > dtmax; it++) step[i] /= 2;
    }
    DCK(hipMemcpyAsync(s.dt, step.data(), nd * 8, hipMemcpyHostToDevice, D->stream));
    // dtj of the j-store follows; acc/jerk stay zero there unless reloaded
    hipLaunchKernelGGL(kb_load_all, dim3(g), dim3(256), 0, D->stream, s, D->js,
                       D->ph4_initial_load ? 1 : 0);
    // the first block time
    hipLaunchKernelGGL(kb_ctl_reset, dim3(1), dim3(1), 0, D->stream, D->ctl);
    hipLaunchKernelGGL(kb_first_min, dim3(g), dim3(256), 0, D->stream, s.t, s.dt, n, D->ctl);
    DCK(hipGetLastError());
    DCK(hipStreamSynchronize(D->stream));
    return 0;
}

int capture(Driver *D, int steps)
{
    if (D->graph) (void)hipGraphExecDestroy(D->graph);
    D->graph = nullptr;
    hipGraph_t g;
    DCK(hipStreamBeginCapture(D->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < steps; k++)
        if (enqueue_step(D, k & 1)) {
            (void)hipStreamEndCapture(D->stream, &g);
            return G6X_ERR_HIP;
        }
    DCK(hipStreamEndCapture(D->stream, &g));
    DCK(hipGraphInstantiate(&D->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    D->graph_steps = steps;
    D->graph_prec = D->prec;
    return 0;
}

}  // namespace g6b

using g6b::Driver;

G6D_API void *g6d_create(int n, const int *ids, const double *m, const double (*x)[3],
                         const double (*v)[3], double eps2, double eta, double dt_fac,
                         double dt_limit, double dt_limitm, int precision, int ph4_initial_load,
                         int device)
{
    if (n < 1 || n > (1 << 27) || !m || !x || !v) return nullptr;
    if (precision != G6X_FP64 && precision != G6X_MIXED) return nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || device < 0 || device >= ndev)
        return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    Driver *D = new Driver;
    D->prec = precision;
    D->eps2 = eps2;
    D->eta = eta;
    D->dt_fac = dt_fac;
    D->dt_limit = dt_limit;
    D->dt_limitm = dt_limitm;
    D->ph4_initial_load = ph4_initial_load != 0;
    std::vector<int> own;
    if (!ids) {
        own.resize(n);
        for (int i = 0; i < n; i++) own[i] = i + 1;
        ids = own.data();
    }
    if (g6b::setup(D, n, ids, m, &x[0][0], &v[0][0]) || g6b::commit(D)) {
        fprintf(stderr, "g6d_create: %s\n", D->err.c_str());
        g6b::release(D);
        return nullptr;
    }
    return D;
}

G6D_API long long g6d_run(void *h, double t_end, long long max_blocks, int steps_per_graph,
                          int *reached)
{
    if (!h) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    // an even number of steps per graph: a replay starts at parity 0
    if (steps_per_graph < 2) steps_per_graph = 2;
    steps_per_graph += steps_per_graph & 1;
    if (!D->graph || D->graph_steps != steps_per_graph || D->graph_prec != D->prec)
        if (g6b::capture(D, steps_per_graph)) return G6X_ERR_HIP;
    if (g6b::read_ctl(D)) return G6X_ERR_HIP;
    const long long b0 = D->h_ctl->blocks;
    hipLaunchKernelGGL(g6b::kb_arm, dim3(1), dim3(1), 0, D->stream, D->ctl, t_end,
                       max_blocks > 0 ? max_blocks : -1);
    for (long long rounds = 0;; rounds++) {
        hipError_t e = hipGraphLaunch(D->graph, D->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(D->h_ctl, D->ctl, sizeof(g6k::StepCtl),
                                                hipMemcpyDeviceToHost, D->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(D->stream);
        if (e != hipSuccess) {
            D->err = std::string("block steps: ") + hipGetErrorString(e);
            return G6X_ERR_HIP;
        }
        if (D->h_ctl->done) break;
        // a run of the graph takes at least one step unless done, so this
        // bound is never reached; it keeps the loop finite regardless
        if (rounds > (1ll << 40)) break;
    }
    if (reached) *reached = D->h_ctl->done == 1;
    return D->h_ctl->blocks - b0;
}

// Every particle behind the system time takes one step to it (ph4's
// synchronize_all, src/amuse_ph4/src/jdata.cc:874-900): a block step whose
// block is that set (counted as a block step, as ph4 counts it), with the
// corrector and the next step as in any block step.  Afterwards every
// particle's time is the system time.
G6D_API int g6d_synchronize(void *h)
{
    if (!h) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    // the pending block time to slot 0 (parity 0 below), no end time
    hipLaunchKernelGGL(g6b::kb_arm, dim3(1), dim3(1), 0, D->stream, D->ctl,
                       __builtin_inf(), -1ll);
    if (g6b::enqueue_step(D, 0, 1)) return G6X_ERR_HIP;
    if (g6b::read_ctl(D)) return G6X_ERR_HIP;
    return 0;
}

// Kinetic and potential energy at the system time (a full-N FP64 pass with
// every particle predicted to it, summed on the host in particle order).
G6D_API int g6d_energy(void *h, double *kin, double *pot)
{
    if (!h || !kin || !pot) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    g6b::State &s = D->st;
    const int n = s.n;
    if (g6b::full_pass(D, G6X_FP64)) return G6X_ERR_HIP;
    const size_t nd = size_t(n), C = size_t(D->ib.cap);
    std::vector<double> vrow(3 * nd), prow(nd), m(nd);
    // predicted velocities (i-block rows) and potentials (out row 6)
    for (int k = 0; k < 3; k++)
        if (hipMemcpyAsync(vrow.data() + k * nd, D->ib.v[k], nd * 8, hipMemcpyDeviceToHost,
                           D->stream) != hipSuccess)
            return G6X_ERR_HIP;
    if (hipMemcpyAsync(prow.data(), D->ib.out + 6 * C, nd * 8, hipMemcpyDeviceToHost,
                       D->stream) != hipSuccess ||
        hipMemcpyAsync(m.data(), s.m, nd * 8, hipMemcpyDeviceToHost, D->stream) != hipSuccess ||
        hipStreamSynchronize(D->stream) != hipSuccess)
        return G6X_ERR_HIP;
    // the step kernels leave on `done`; the next run re-arms it
    double K = 0, U = 0;
    for (int i = 0; i < n; i++) {
        const double vx = vrow[i], vy = vrow[nd + i], vz = vrow[2 * nd + i];
        K += 0.5 * m[i] * (vx * vx + vy * vy + vz * vz);
        U += 0.5 * m[i] * prow[i];
    }
    *kin = K;
    *pot = U;
    return 0;
}

// Positions and velocities as last corrected, each at its particle's own
// time ([n][3] each; ph4's get_state, src/amuse_ph4/interface.cc:485-502).
G6D_API int g6d_state_raw(void *h, double (*x)[3], double (*v)[3])
{
    if (!h || !x || !v) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    g6b::State &s = D->st;
    const size_t nd = size_t(s.n);
    if (g6b::read_ctl(D)) return G6X_ERR_HIP;
    std::vector<double> col(6 * nd);
    double *rows[6] = {s.x[0], s.x[1], s.x[2], s.v[0], s.v[1], s.v[2]};
    for (int r = 0; r < 6; r++)
        if (hipMemcpyAsync(col.data() + r * nd, rows[r], nd * 8, hipMemcpyDeviceToHost,
                           D->stream) != hipSuccess)
            return G6X_ERR_HIP;
    if (hipStreamSynchronize(D->stream) != hipSuccess) return G6X_ERR_HIP;
    for (size_t i = 0; i < nd; i++)
        for (int k = 0; k < 3; k++) {
            x[i][k] = col[k * nd + i];
            v[i][k] = col[(3 + k) * nd + i];
        }
    return 0;
}

// Positions and velocities predicted to the system time ([n][3] each).
G6D_API int g6d_state(void *h, double (*x)[3], double (*v)[3])
{
    if (!h || !x || !v) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    g6b::State &s = D->st;
    const int n = s.n;
    const size_t nd = size_t(n);
    if (g6b::read_ctl(D)) return G6X_ERR_HIP;
    std::vector<double> col(12 * nd), t(nd);
    double *rows[12] = {s.x[0], s.x[1], s.x[2], s.v[0], s.v[1], s.v[2],
                        s.a[0], s.a[1], s.a[2], s.j[0], s.j[1], s.j[2]};
    for (int r = 0; r < 12; r++)
        if (hipMemcpyAsync(col.data() + r * nd, rows[r], nd * 8, hipMemcpyDeviceToHost,
                           D->stream) != hipSuccess)
            return G6X_ERR_HIP;
    if (hipMemcpyAsync(t.data(), s.t, nd * 8, hipMemcpyDeviceToHost, D->stream) != hipSuccess ||
        hipStreamSynchronize(D->stream) != hipSuccess)
        return G6X_ERR_HIP;
    const double ts = D->h_ctl->system_time;
    for (int i = 0; i < n; i++) {
        double xi[3], vi[3], ai[3], ji[3];
        for (int k = 0; k < 3; k++) {
            xi[k] = col[k * nd + i];
            vi[k] = col[(3 + k) * nd + i];
            ai[k] = col[(6 + k) * nd + i];
            ji[k] = col[(9 + k) * nd + i];
        }
        h4::predict_ph4(xi, vi, ai, ji, ts - t[i], x[i], v[i]);
    }
    return 0;
}

// Per-particle times and steps.
G6D_API int g6d_steps(void *h, double *time, double *step)
{
    if (!h || !time || !step) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    const size_t nd = size_t(D->st.n);
    if (hipMemcpyAsync(time, D->st.t, nd * 8, hipMemcpyDeviceToHost, D->stream) != hipSuccess ||
        hipMemcpyAsync(step, D->st.dt, nd * 8, hipMemcpyDeviceToHost, D->stream) != hipSuccess ||
        hipStreamSynchronize(D->stream) != hipSuccess)
        return G6X_ERR_HIP;
    return 0;
}

G6D_API int g6d_counters(void *h, long long *blocks, long long *psteps, double *system_time)
{
    if (!h || !blocks || !psteps || !system_time) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    if (g6b::read_ctl(D)) return G6X_ERR_HIP;
    *blocks = D->h_ctl->blocks;
    *psteps = D->h_ctl->psteps;
    *system_time = D->h_ctl->system_time;
    return 0;
}

G6D_API int g6d_set_precision(void *h, int precision)
{
    if (!h) return G6X_ERR_ARG;
    Driver *D = static_cast<Driver *>(h);
    if (precision != G6X_FP64 && precision != G6X_MIXED) return G6X_ERR_ARG;
    D->prec = precision;
    return 0;
}

G6D_API const char *g6d_last_error(void *h)
{
    return h ? static_cast<Driver *>(h)->err.c_str() : "null driver handle";
}

G6D_API void g6d_free(void *h)
{
    if (h) g6b::release(static_cast<Driver *>(h));
}

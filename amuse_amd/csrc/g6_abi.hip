// GRAPE-6 C ABI on MI355X: the unit state machine behind g6_open ...
// g6calc_lasthalf2 (reference: lib/g6lib/g6lib.c:190-481 for semantics,
// lib/sapporo_light/sapporo.cpp:19-272 for the firsthalf/lasthalf split).
//
// One global unit per process, like both reference libraries; not
// thread-safe, like them.  All device work goes to one private stream:
//
//   set_j_particle   host staging (deduplicated per slot, last write wins)
//   firsthalf        staging -> pinned -> HBM + scatter kernel; predictor
//                    pass if ti moved; i-block -> pinned -> HBM; with
//                    G6_EAGER=1 the force kernel is launched here too
//   lasthalf[2]      force kernel (+ nn if lasthalf2) + chunk reduction,
//                    one D2H of the packed results, stream sync, unpack.
//
// Errors are returned as negative codes (g6x.h); the library never exits.
#include "g6_kernels.h"
#include "../../include/g6lib.h"
#include "../../include/g6x.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#define G6_API extern "C" __attribute__((visibility("default")))
#define G6_WEAK extern "C" __attribute__((visibility("default"), weak))

namespace {

using g6k::ForceCfg;
using g6k::IBlock;
using g6k::JRecord;
using g6k::JStore;

constexpr int kMaxPipes = 1 << 22;
constexpr int kMaxSlots = 1 << 27;
constexpr long long kPartBudget = 1ll << 24;  // nchunk * ni limit for partials (<= 2 GB)
// i-blocks and j-stages up to this many rows go between host and kernels
// through pinned host memory that the kernels read and write themselves (no
// copy launches: at the block sizes of a block-step integrator the copies'
// latency, not their bytes, is the cost); larger ones are copied.
constexpr int kDirectRows = 4096;

struct TuneEntry {
    int ni_lo, ni_hi;  // bucket [lo, hi]
    int njs, ipl, nchunk, smem;
};

struct Unit {
    bool open = false;
    bool have_device = false;
    int device = -1;
    hipStream_t stream = nullptr;

    JStore js;
    int nj_hw = 0;  // one past the highest slot ever written since reset
    IBlock ib;

    std::vector<JRecord> stage;
    std::vector<int> stage_pos;  // slot -> index in stage, -1 if none
    // Ids of the slots as the set_j_particle calls so far leave them (-1:
    // empty) and how many slots hold each id.  g6x_calc_self leaves out the
    // per-pair id compare in work items whose slots cannot meet only while
    // every slot below nj holds an id and no id is held twice.
    std::vector<int> slot_id;
    std::unordered_map<int, int> id_count;
    long long id_repeats = 0;  // sum over ids of (slots holding it - 1)
    int n_filled = 0;          // slots holding an id other than -1

    // Full passes of the mixed path in spatial order (g6x_calc_self over
    // every slot): the sorted copy of the predicted store, the permutation,
    // the tile boxes, result rows in sorted order, the sort's buffers.
    bool spatial = true;
    JStore sj;
    int sp_cap = 0;
    int *sp_perm = nullptr;
    float *sp_tbox = nullptr;
    double *sp_rows = nullptr;  // [8][ib.cap]
    int sp_rows_cap = 0;
    void *sp_base = nullptr;
    g6k::SpatialBuf sp_buf;
    JRecord *h_stage = nullptr;
    JRecord *d_hstage = nullptr;  // the device's address of h_stage
    size_t h_stage_cap = 0;
    JRecord *d_stage = nullptr;
    size_t d_stage_cap = 0;
    hipEvent_t ev_stage = nullptr;

    // pinned mirrors of the device i-block rows (same row pitch): [8][cap]
    // x y z vx vy vz h2 id and [8][cap] acc3 jerk3 pot nn; id and nn are
    // ints in the first half of their row.  Blocks of at most kDirectRows
    // are read and written there by the kernels themselves (d_in, d_out:
    // the device's addresses of the same memory); larger ones are copied.
    double *h_in = nullptr;
    double *h_out = nullptr;
    double *d_in = nullptr;
    double *d_out = nullptr;
    int h_cap = 0;
    hipEvent_t ev_in = nullptr;  // after the last use of h_in by the device
    bool in_direct = false;      // the latched i-block is read from h_in
    bool out_direct = false;     // the last launch wrote its results into h_out

    double ti = 0.0;
    bool pred_valid = false;
    double pred_ti = 0.0;
    int pred_prec = -1;
    int pred_compat = -1;
    bool k18_nonzero = false;

    int prec = g6k::kF64;
    int compat = 0;
    int npipes = 16384;
    bool eager = false;

    bool latched = false;
    int first_rc = 0;  // error of the last firsthalf (it returns void)
    int l_ni = 0, l_nj = 0;
    double l_eps2 = 0.0;
    std::vector<int> l_id;
    bool launched = false;
    bool launched_nn = false;
    bool l_pred = false;  // the latched call's force kernel predicts its j-slots
    double l_ti = 0.0;    // ti of the latched call
    int last_ni = 0;  // rows valid in ib.out

    int tile_njs = 0, tile_ipl = 0, tile_nchunk = 0, tile_smem = 0;
    std::vector<TuneEntry> tune;

    // neighbour lists of the last calc
    int *d_ngb_counts = nullptr, *d_ngb_lists = nullptr;
    int ngb_cap_i = 0;
    std::vector<int> h_ngb_counts, h_ngb_lists;
    int ngb_ni = 0;

    bool timing = false;
    struct Ev {
        hipEvent_t a, b;
        int kind;  // 0 force, 1 reduce, 2 predict
        double pairs;
    };
    std::vector<Ev> evs;
    std::vector<hipEvent_t> ev_pool;
    double t_force = 0, t_reduce = 0, t_pred = 0, t_pairs = 0;
    long long n_force = 0;

    std::string err;
};

Unit U;

int fail(int code, const char *what, hipError_t e = hipSuccess)
{
    char buf[256];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    U.err = buf;
    return code;
}

#define HIPCHK(expr, what)                                        \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return fail(G6X_ERR_HIP, what, e_); \
    } while (0)

int env_int(const char *name, int dflt)
{
    const char *s = getenv(name);
    if (!s || !*s) return dflt;
    return atoi(s);
}

hipEvent_t pool_event()
{
    if (!U.ev_pool.empty()) {
        hipEvent_t e = U.ev_pool.back();
        U.ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timed_begin(Unit::Ev &ev, int kind, double pairs)
{
    ev.a = ev.b = nullptr;
    ev.kind = kind;
    ev.pairs = pairs;
    if (!U.timing) return;
    ev.a = pool_event();
    ev.b = pool_event();
    if (ev.a) (void)hipEventRecord(ev.a, U.stream);
}

void timed_end(Unit::Ev &ev)
{
    if (!U.timing || !ev.a || !ev.b) return;
    (void)hipEventRecord(ev.b, U.stream);
    U.evs.push_back(ev);
}

void harvest_timing()
{
    for (auto &ev : U.evs) {
        (void)hipEventSynchronize(ev.b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev.a, ev.b);
        if (ev.kind == 0) {
            U.t_force += ms;
            U.t_pairs += ev.pairs;
            U.n_force++;
        } else if (ev.kind == 1) {
            U.t_reduce += ms;
        } else {
            U.t_pred += ms;
        }
        U.ev_pool.push_back(ev.a);
        U.ev_pool.push_back(ev.b);
    }
    U.evs.clear();
}

// ---------------------------------------------------------------------------
// Device memory
// ---------------------------------------------------------------------------
int ensure_device()
{
    if (U.have_device) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(G6X_ERR_NODEVICE, "no GPU visible");
    U.have_device = true;
    return 0;
}

int grow_jstore(int need)
{
    if (need <= U.js.cap) return 0;
    if (need > kMaxSlots) return fail(G6X_ERR_ARG, "j-slot address beyond 2^27");
    int cap = std::max(U.js.cap ? U.js.cap : 1, 1 << 16);
    while (cap < need) cap *= 2;
    void *base = nullptr;
    hipError_t e = hipMalloc(&base, g6k::jstore_bytes(cap));
    if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc j-store", e);
    JStore ns;
    g6k::jstore_carve(ns, base, cap);
    const int old = U.js.cap;
    if (old > 0) {
        HIPCHK(g6k::launch_copy_store(ns, U.js, old, U.stream), "copy j-store");
        HIPCHK(g6k::launch_reset(ns, old, cap, U.stream), "reset j-store");
        HIPCHK(hipStreamSynchronize(U.stream), "grow sync");
        (void)hipFree(U.js.base);
    } else {
        HIPCHK(g6k::launch_reset(ns, 0, cap, U.stream), "reset j-store");
    }
    U.js = ns;
    U.stage_pos.resize(cap, -1);
    U.slot_id.resize(cap, -1);
    return 0;
}

// All slots empty (the device store is reset alongside).
void clear_ids()
{
    std::fill(U.slot_id.begin(), U.slot_id.end(), -1);
    U.id_count.clear();
    U.id_repeats = 0;
    U.n_filled = 0;
}

void note_id(int address, int id)
{
    const int old = U.slot_id[address];
    if (old == id) return;
    if (old != -1) {
        auto it = U.id_count.find(old);
        if (it->second > 1) U.id_repeats--;
        if (--it->second == 0) U.id_count.erase(it);
        U.n_filled--;
    }
    if (id != -1) {
        int &c = U.id_count[id];
        if (c > 0) U.id_repeats++;
        c++;
        U.n_filled++;
    }
    U.slot_id[address] = id;
}

int grow_iblock(int need)
{
    if (need <= U.ib.cap && need <= U.h_cap) return 0;
    if (need > kMaxPipes) return fail(G6X_ERR_ARG, "i-block larger than 2^22");
    int cap = std::max(U.ib.cap, 1024);
    while (cap < need) cap *= 2;
    HIPCHK(hipStreamSynchronize(U.stream), "iblock sync");
    void *base = nullptr;
    hipError_t e = hipMalloc(&base, g6k::iblock_bytes(cap));
    if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc i-block", e);
    if (U.ib.base) (void)hipFree(U.ib.base);
    if (U.ib.part_base) (void)hipFree(U.ib.part_base);
    IBlock nb;
    g6k::iblock_carve(nb, base, cap);
    U.ib = nb;
    if (U.h_in) { (void)hipHostFree(U.h_in); (void)hipHostFree(U.h_out); }
    U.h_in = nullptr; U.h_out = nullptr;
    e = hipHostMalloc(reinterpret_cast<void **>(&U.h_in), sizeof(double) * 8 * size_t(cap));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&U.h_out), sizeof(double) * 8 * size_t(cap));
    if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipHostMalloc i/o", e);
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&U.d_in), U.h_in, 0), "i/o map");
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&U.d_out), U.h_out, 0), "i/o map");
    U.h_cap = cap;
    return 0;
}

// Partials of a launch are packed [nchunk][fields][ni]; the buffer grows to
// the largest nchunk * ni seen (at most kPartBudget records).
int ensure_partials(int nchunk, int ni)
{
    const long long need = (long long)nchunk * ni;
    if (nchunk <= 1 || need <= U.ib.part_entries) return 0;
    HIPCHK(hipStreamSynchronize(U.stream), "partials sync");
    if (U.ib.part_base) (void)hipFree(U.ib.part_base);
    U.ib.part_base = nullptr;
    U.ib.part = nullptr;
    U.ib.part_entries = 0;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(double) * g6k::kPartFieldsD * size_t(need));
    if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc partials", e);
    U.ib.part_base = p;
    U.ib.part = static_cast<double *>(p);
    U.ib.part_entries = need;
    return 0;
}

// ---------------------------------------------------------------------------
// j staging and prediction
// ---------------------------------------------------------------------------
// may_allocate = false (the allocation-free entry point): pending records
// that do not fit the staging buffers already there are an error, and
// without the staging event the stream is synchronised instead of recorded.
int flush_stage(bool may_allocate = true)
{
    const size_t n = U.stage.size();
    if (n == 0) return 0;
    if (!may_allocate && (n > U.h_stage_cap || (n > size_t(kDirectRows) && n > U.d_stage_cap)))
        return fail(G6X_ERR_NOMEM, "pending set_j_particle records exceed the staging buffers "
                                   "(the allocation-free call cannot grow them)");
    if (U.ev_stage) HIPCHK(hipEventSynchronize(U.ev_stage), "stage event");
    if (n > U.h_stage_cap) {
        if (U.h_stage) (void)hipHostFree(U.h_stage);
        size_t cap = std::max<size_t>(n, 4096);
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&U.h_stage), sizeof(JRecord) * cap);
        if (e != hipSuccess) { U.h_stage = nullptr; U.h_stage_cap = 0; return fail(G6X_ERR_NOMEM, "hipHostMalloc stage", e); }
        U.h_stage_cap = cap;
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&U.d_hstage), U.h_stage, 0), "stage map");
    }
    if (n > size_t(kDirectRows) && n > U.d_stage_cap) {
        HIPCHK(hipStreamSynchronize(U.stream), "stage sync");
        if (U.d_stage) (void)hipFree(U.d_stage);
        size_t cap = std::max<size_t>(n, 4096);
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&U.d_stage), sizeof(JRecord) * cap);
        if (e != hipSuccess) { U.d_stage = nullptr; U.d_stage_cap = 0; return fail(G6X_ERR_NOMEM, "hipMalloc stage", e); }
        U.d_stage_cap = cap;
    }
    memcpy(U.h_stage, U.stage.data(), sizeof(JRecord) * n);
    // small stages are read by the scatter kernel where they lie (pinned
    // host memory); large ones are copied to the device first
    const JRecord *rec = U.d_hstage;
    if (n > size_t(kDirectRows)) {
        HIPCHK(hipMemcpyAsync(U.d_stage, U.h_stage, sizeof(JRecord) * n, hipMemcpyHostToDevice, U.stream), "stage H2D");
        rec = U.d_stage;
    }
    const bool pv = U.pred_valid && U.pred_prec == U.prec && U.pred_compat == U.compat;
    HIPCHK(g6k::launch_scatter(U.js, rec, int(n), U.prec, pv ? 1 : 0, U.pred_ti, U.compat, U.stream), "scatter");
    if (!U.ev_stage && !may_allocate) {
        HIPCHK(hipStreamSynchronize(U.stream), "stage sync");
    } else {
        if (!U.ev_stage) HIPCHK(hipEventCreateWithFlags(&U.ev_stage, hipEventDisableTiming), "event");
        HIPCHK(hipEventRecord(U.ev_stage, U.stream), "stage record");
    }
    for (const auto &r : U.stage) U.stage_pos[r.addr] = -1;
    U.stage.clear();
    if (!pv) U.pred_valid = false;
    return 0;
}

// The predicted store holds the j-slots at time t for the current precision
// and predictor mode.  A latched calc and its neighbour query use the ti of
// its firsthalf (l_ti), whatever set_ti was called with since.
bool pred_current(double t)
{
    return U.pred_valid && U.pred_ti == t && U.pred_prec == U.prec && U.pred_compat == U.compat;
}

// timed = false: no timing events (they may be created on first use).
int ensure_predicted(double t, bool timed = true)
{
    if (pred_current(t)) return 0;
    Unit::Ev ev;
    if (timed) timed_begin(ev, 2, 0.0);
    HIPCHK(g6k::launch_predict(U.js, U.nj_hw, t, U.prec, U.compat, U.k18_nonzero ? 1 : 0, U.stream), "predict");
    if (timed) timed_end(ev);
    U.pred_valid = true;
    U.pred_ti = t;
    U.pred_prec = U.prec;
    U.pred_compat = U.compat;
    return 0;
}

// ---------------------------------------------------------------------------
// Tile choice
// ---------------------------------------------------------------------------
// j-chunks of the automatic choice for `iblocks` i-blocks of an ni x nj
// launch: ~8 workgroups per CU at least; for long j-ranges up to ~128, so a
// launch is many rounds of workgroups and the last round's imbalance is
// small (at ni = nj = 2^20, two rounds lost 5 % against 32), with every
// chunk still >= 16 tiles; at most one chunk per tile, and within the
// partials budget.
int auto_chunks(int iblocks, int ni, int nj)
{
    const int max_chunks = std::max(1, (nj + g6k::kTJ - 1) / g6k::kTJ);
    const int base = (2048 + iblocks - 1) / iblocks;
    const int deep = std::min((32768 + iblocks - 1) / iblocks, nj / (16 * g6k::kTJ));
    int nchunk = std::min(std::max(1, std::max(base, deep)), max_chunks);
    while (nchunk > 1 && (long long)nchunk * ni > kPartBudget) nchunk /= 2;
    return nchunk;
}

ForceCfg with_chunks(ForceCfg c, int nchunk, int ni, int nj, long long max_records);

ForceCfg choose_cfg(int ni, int nj)
{
    ForceCfg c;
    int njs = 0, ipl = 0, nchunk = 0, smem = 0;
    if (U.tile_njs > 0) {
        njs = U.tile_njs; ipl = U.tile_ipl; nchunk = U.tile_nchunk; smem = U.tile_smem;
    } else {
        for (const auto &t : U.tune)
            if (ni >= t.ni_lo && ni <= t.ni_hi) {
                njs = t.njs; ipl = t.ipl; nchunk = t.nchunk; smem = t.smem;
                break;
            }
    }
    if (njs == 0 && nj <= 8 * g6k::kTJ) {
        // Short j-ranges: one i per wave over all j, no j-split and so no
        // reduce kernel (the small-system plan of the block-step driver);
        // a split j-range gives a few workgroups a tile each and costs a
        // reduce launch that the few tiles saved do not pay for.
        njs = 64; ipl = 1; nchunk = 1; smem = 0;
    }
    if (njs == 0) {
        int p2 = 1;
        while (p2 < ni) p2 *= 2;
        if ((long long)ni * nj <= (1ll << 29)) {
            // latency-bound launches: the widest i-block of at most 64 i (4
            // lanes per i) that still gives >= 256 workgroups with the
            // j-chunks the range allows, else the one with the most (best
            // or within a few % of the best of eleven shapes at nj = 16384
            // and 65536 for every ni = 2^k below 2^29 / nj, both
            // precisions: wider i-blocks leave too few workgroups once the
            // j-range is split as far as it goes, narrower ones pay for
            // their cross-lane sums and dispatch)
            long long most = -1;
            for (int ib = std::min(64, p2); ib >= 1; ib /= 2) {
                const int iblocks = (ni + ib - 1) / ib;
                const long long items = (long long)iblocks * auto_chunks(iblocks, ni, nj);
                if (items > most) { most = items; njs = g6k::kWG / ib; }
                if (items >= 256) break;
            }
            ipl = 1;
        } else {
            // throughput-bound: the widest i-block that does not leave lanes
            // idle, the j-range split into enough chunks below
            static const int kOrder[][2] = {{1, 2}, {1, 1}, {2, 1}, {4, 1}, {8, 1}, {16, 1},
                                            {32, 1}, {64, 1}, {128, 1}, {256, 1}};
            for (const auto &o : kOrder) {
                const int ib = (g6k::kWG / o[0]) * o[1];
                if (ib <= p2) { njs = o[0]; ipl = o[1]; break; }
            }
            if (njs == 0) { njs = 256; ipl = 1; }
        }
        nchunk = 0;
        smem = 0;  // LDS tiles measured faster than scalar loads at every size tried
    }
    const int ib = (g6k::kWG / njs) * ipl;
    const int iblocks = (ni + ib - 1) / ib;
    if (nchunk <= 0) nchunk = auto_chunks(iblocks, ni, nj);
    c.njs = njs;
    c.ipl = ipl;
    c.smem = smem != 0 && njs == 1;
    return with_chunks(c, nchunk, ni, nj, kPartBudget);
}

// cfg with at most `nchunk` j-chunks (whole tiles each) and at most
// `max_records` (chunk, i) partial records.
ForceCfg with_chunks(ForceCfg c, int nchunk, int ni, int nj, long long max_records)
{
    const int max_chunks = std::max(1, (nj + g6k::kTJ - 1) / g6k::kTJ);
    nchunk = std::max(1, std::min(nchunk, max_chunks));
    while (nchunk > 1 && (long long)nchunk * ni > max_records) nchunk /= 2;
    int jchunk = (nj + nchunk - 1) / nchunk;
    jchunk = ((jchunk + g6k::kTJ - 1) / g6k::kTJ) * g6k::kTJ;
    if (jchunk <= 0) jchunk = g6k::kTJ;
    c.nchunk = std::max(1, (nj + jchunk - 1) / jchunk);
    c.jchunk = jchunk;
    return c;
}

// Records ev_in: h_in may be rewritten once the work queued so far is done.
hipError_t record_in()
{
    if (!U.ev_in) {
        hipError_t e = hipEventCreateWithFlags(&U.ev_in, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipEventRecord(U.ev_in, U.stream);
}

// The i-block as the kernels see it: the device rows, or for a latched
// block read in place the pinned input rows; the result rows in pinned
// memory for blocks of at most kDirectRows (same row pitch either way).
IBlock io_view(int ni)
{
    IBlock b = U.ib;
    const size_t C = size_t(U.ib.cap);
    if (U.in_direct) {
        for (int k = 0; k < 3; k++) {
            b.x[k] = U.d_in + k * C;
            b.v[k] = U.d_in + (3 + k) * C;
        }
        b.h2 = U.d_in + 6 * C;
        b.id = reinterpret_cast<int *>(U.d_in + 7 * C);
    }
    if (ni <= kDirectRows) {
        b.out = U.d_out;
        b.nn = reinterpret_cast<int *>(U.d_out + 7 * C);
    }
    return b;
}

// self_i0: as g6k::launch_force (>= 0 only for g6x_calc_self).
int run_force(int ni, int nj, double eps2, bool want_nn, int self_i0 = -1)
{
    if (ni <= 0) { U.last_ni = 0; return 0; }
    ForceCfg cfg = choose_cfg(ni, std::max(nj, 1));
    int rc = ensure_partials(cfg.nchunk, ni);
    if (rc) return rc;
    // the latched call's j-slots predicted by the force kernel itself (see
    // do_firsthalf), unless the shape changed since: then by the predictor
    bool pred = false;
    if (U.latched && U.l_pred) {
        pred = g6k::pred_force_cfg(cfg, nj);
        if (!pred) {
            rc = ensure_predicted(U.l_ti);
            if (rc) return rc;
        }
    }
    const IBlock b = io_view(ni);
    Unit::Ev ev;
    timed_begin(ev, 0, double(ni) * double(nj));
    HIPCHK(g6k::launch_force(U.js, b, ni, nj, eps2, U.prec, want_nn, cfg, U.stream, pred, U.l_ti,
                             U.compat, self_i0),
           "force kernel");
    timed_end(ev);
    if (cfg.nchunk > 1) {
        Unit::Ev ev2;
        timed_begin(ev2, 1, 0.0);
        HIPCHK(g6k::launch_reduce(U.js, b, ni, cfg.nchunk, U.prec, want_nn, U.stream), "reduce kernel");
        timed_end(ev2);
    }
    U.out_direct = b.out != U.ib.out;
    if (U.in_direct) HIPCHK(record_in(), "i record");
    U.launched = true;
    U.launched_nn = want_nn;
    U.last_ni = ni;
    return 0;
}

int check_open()
{
    if (!U.open) return fail(G6X_ERR_NOT_OPEN, "g6 unit not open");
    return 0;
}

// ---------------------------------------------------------------------------
// Full pass without allocation (g6x_calc_self_noalloc)
// ---------------------------------------------------------------------------
// The i-block rows, result rows and j-chunk partials of this entry point live
// in one static device array per GPU (kept until the process ends).  Host
// code reaches it only through the address hipGetSymbolAddress gives with
// that GPU current (scratch_ptr); the array's own name is a host-side
// stand-in.  Calls share the array: they run in order on the unit's stream,
// and calls that overlapped on one GPU (another thread, another stream)
// would race.  Nothing in it needs clearing: every row a call reads is
// written by that call first.
constexpr size_t kScratchDoubles = size_t(G6X_NOALLOC_SCRATCH_BYTES) / sizeof(double);
static_assert(size_t(G6X_NOALLOC_MAX_NI) * 16 <= kScratchDoubles, "rows must fit the scratch");
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
double *scratch_ptr[kScratchDevices] = {};

// Once per device, with it current (at g6_open, else at the first call).
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= kScratchDevices) return fail(G6X_ERR_NODEVICE, "device id beyond 63");
    if (scratch_ptr[dev]) return 0;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    scratch_ptr[dev] = static_cast<double *>(p);
    return 0;
}

int do_open(int cluster)
{
    int rc = ensure_device();
    if (rc) return rc;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev), "device count");
    int dev = env_int("G6_DEVICE", cluster);
    if (dev < 0 || dev >= ndev) return fail(G6X_ERR_NODEVICE, "device id out of range");
    if (U.stream && dev != U.device) {
        // another device: drop the cached allocations of the old one
        (void)hipSetDevice(U.device);
        (void)hipStreamSynchronize(U.stream);
        if (U.js.base) (void)hipFree(U.js.base);
        if (U.ib.base) (void)hipFree(U.ib.base);
        if (U.ib.part_base) (void)hipFree(U.ib.part_base);
        if (U.d_stage) (void)hipFree(U.d_stage);
        if (U.sp_base) (void)hipFree(U.sp_base);
        if (U.sp_rows) (void)hipFree(U.sp_rows);
        if (U.sp_buf.base) (void)hipFree(U.sp_buf.base);
        if (U.sp_buf.tmp) (void)hipFree(U.sp_buf.tmp);
        U.sp_base = nullptr;
        U.sp_rows = nullptr;
        U.sp_cap = U.sp_rows_cap = 0;
        U.sp_buf = g6k::SpatialBuf();
        (void)hipStreamDestroy(U.stream);
        U.js = JStore();
        U.ib = IBlock();
        U.d_stage = nullptr;
        U.d_stage_cap = 0;
        U.stream = nullptr;
        U.ev_stage = nullptr;
        U.ev_pool.clear();
        U.evs.clear();
    }
    HIPCHK(hipSetDevice(dev), "hipSetDevice");
    U.device = dev;
    (void)scratch_setup(dev);  // retried by g6x_calc_self_noalloc if it failed
    if (!U.stream) HIPCHK(hipStreamCreateWithFlags(&U.stream, hipStreamNonBlocking), "stream");

    const char *p = getenv("G6_PRECISION");
    U.prec = (p && (!strcmp(p, "mixed") || !strcmp(p, "ds"))) ? g6k::kMixed : g6k::kF64;
    const char *c = getenv("G6_PREDICTOR_COMPAT");
    U.compat = (c && !strcmp(c, "g6lib")) ? 1 : 0;
    U.npipes = std::min(std::max(env_int("G6_NPIPES", 16384), 1), kMaxPipes);
    U.eager = env_int("G6_EAGER", 0) != 0;

    rc = grow_jstore(std::max(env_int("G6_NJ_MAX", 1 << 16), 1));
    if (rc) return rc;
    rc = grow_iblock(std::min(U.npipes, 1 << 16));
    if (rc) return rc;
    HIPCHK(g6k::launch_reset(U.js, 0, U.js.cap, U.stream), "reset");
    HIPCHK(hipStreamSynchronize(U.stream), "open sync");
    U.stage.clear();
    std::fill(U.stage_pos.begin(), U.stage_pos.end(), -1);
    clear_ids();
    U.nj_hw = 0;
    U.pred_valid = false;
    U.k18_nonzero = false;
    U.ti = 0.0;
    U.latched = false;
    U.launched = false;
    U.open = true;
    U.err.clear();
    return 0;
}

int do_set_j(int address, int index, double tj, double dtj, double mass, const double *k18,
             const double *j6, const double *a2, const double *v, const double *x)
{
    int rc = check_open();
    if (rc) return rc;
    if (address < 0) return fail(G6X_ERR_ARG, "negative j address");
    if (address >= U.js.cap) {
        rc = flush_stage();
        if (rc) return rc;
        rc = grow_jstore(address + 1);
        if (rc) return rc;
    }
    JRecord r;
    r.tj = tj; r.dtj = dtj; r.m = mass;
    for (int k = 0; k < 3; k++) {
        r.x[k] = x[k]; r.v[k] = v[k]; r.a2[k] = a2[k]; r.j6[k] = j6[k]; r.k18[k] = k18[k];
        if (k18[k] != 0.0) U.k18_nonzero = true;
    }
    r.addr = address;
    r.id = index;
    int &pos = U.stage_pos[address];
    if (pos >= 0) U.stage[pos] = r;
    else { pos = int(U.stage.size()); U.stage.push_back(r); }
    note_id(address, index);
    if (address + 1 > U.nj_hw) U.nj_hw = address + 1;
    return 0;
}

int do_firsthalf(int nj, int ni, const int *index, const double (*xi)[3], const double (*vi)[3],
                 double eps2, const double *h2)
{
    U.latched = false;
    U.launched = false;
    int rc = check_open();
    if (rc) return rc;
    if (ni < 0 || nj < 0) return fail(G6X_ERR_ARG, "negative ni or nj");
    if (ni > kMaxPipes) return fail(G6X_ERR_ARG, "ni beyond 2^22");
    if (nj > U.js.cap) {
        rc = flush_stage();
        if (rc) return rc;
        rc = grow_jstore(nj);
        if (rc) return rc;
    }
    if (nj > U.nj_hw) U.nj_hw = nj;  // slots never written are empty (id -1)
    rc = grow_iblock(std::max(ni, 1));
    if (rc) return rc;
    rc = flush_stage();
    if (rc) return rc;
    const ForceCfg cfg = choose_cfg(std::max(ni, 1), std::max(nj, 1));
    // a short j-range's force kernel predicts the j-slots as it loads them
    // (no predictor launch); otherwise the store is predicted to ti here
    U.l_pred = ni > 0 && g6k::pred_force_cfg(cfg, nj) && !U.k18_nonzero && !pred_current(U.ti);
    U.l_ti = U.ti;
    if (!U.l_pred) {
        rc = ensure_predicted(U.l_ti);
        if (rc) return rc;
    }
    if (ni > 0) {
        if (U.ev_in) HIPCHK(hipEventSynchronize(U.ev_in), "i event");
        const size_t C = size_t(U.ib.cap);
        int *h_id = reinterpret_cast<int *>(U.h_in + 7 * C);
        for (int i = 0; i < ni; i++) {
            for (int k = 0; k < 3; k++) {
                U.h_in[k * C + i] = xi[i][k];
                U.h_in[(3 + k) * C + i] = vi[i][k];
            }
            U.h_in[6 * C + i] = h2 ? h2[i] : 0.0;
            h_id[i] = index[i];
        }
        // read in place when the force launch reads each row a few times at
        // most (once per j-chunk of its i-block), else copied once
        U.in_direct = (long long)cfg.nchunk * ni <= kDirectRows;
        if (!U.in_direct) {
            // all eight rows in one pitched copy (the id row's second half
            // is copied along, unused)
            HIPCHK(hipMemcpy2DAsync(U.ib.x[0], sizeof(double) * C, U.h_in, sizeof(double) * C,
                                    sizeof(double) * ni, 8, hipMemcpyHostToDevice, U.stream),
                   "i H2D");
            HIPCHK(record_in(), "i record");
        }
    }
    U.l_ni = ni;
    U.l_nj = nj;
    U.l_eps2 = eps2;
    U.l_id.assign(index, index + ni);
    U.latched = true;
    if (U.eager) {
        rc = run_force(ni, nj, eps2, true);
        if (rc) return rc;
    }
    return 0;
}

int fetch_results(int ni, double (*acc)[3], double (*jerk)[3], double *pot, int *nn, bool want_nn)
{
    if (ni <= 0) return 0;
    const size_t C = size_t(U.ib.cap);
    // result rows acc3 jerk3 pot (nn) follow each other at pitch C: written
    // there by the kernels, or one copy
    if (!U.out_direct)
        HIPCHK(hipMemcpy2DAsync(U.h_out, sizeof(double) * C, U.ib.out, sizeof(double) * C,
                                sizeof(double) * ni, want_nn ? 8 : 7, hipMemcpyDeviceToHost,
                                U.stream),
               "out D2H");
    const int *h_nn = reinterpret_cast<const int *>(U.h_out + 7 * C);
    HIPCHK(hipStreamSynchronize(U.stream), "lasthalf sync");
    for (int i = 0; i < ni; i++) {
        for (int k = 0; k < 3; k++) {
            acc[i][k] = U.h_out[k * C + i];
            jerk[i][k] = U.h_out[(3 + k) * C + i];
        }
        pot[i] = U.h_out[6 * C + i];
        if (want_nn) nn[i] = h_nn[i];
    }
    return 0;
}

int do_lasthalf(int *index, double (*acc)[3], double (*jerk)[3], double *pot, int *nn)
{
    int rc = check_open();
    if (rc) return rc;
    if (!U.latched) {
        if (U.first_rc < 0) return U.first_rc;  // message of the failed firsthalf kept
        return fail(G6X_ERR_SEQUENCE, "lasthalf without firsthalf");
    }
    const bool want_nn = nn != nullptr;
    if (!U.launched || (want_nn && !U.launched_nn)) {
        rc = run_force(U.l_ni, U.l_nj, U.l_eps2, want_nn);
        if (rc) return rc;
    }
    rc = fetch_results(U.l_ni, acc, jerk, pot, nn, want_nn);
    if (rc) return rc;
    for (int i = 0; i < U.l_ni; i++) index[i] = U.l_id[i];  // g6lib.c:397,427
    return 0;
}

}  // namespace

// ===========================================================================
// Fortran / by-pointer ABI
// ===========================================================================
G6_API int g6_open_(int *id) { return do_open(id ? *id : 0); }

G6_API int g6_close_(int *)
{
    if (U.open && U.stream) (void)hipStreamSynchronize(U.stream);
    U.open = false;
    U.latched = false;
    return 0;
}

G6_API int g6_npipes_(void) { return U.npipes; }
G6_API int g6_set_tunit_(void *) { return 0; }
G6_API int g6_set_xunit_(void *) { return 0; }

G6_API int g6_set_ti_(int *, double *ti)
{
    int rc = check_open();
    if (rc) return rc;
    U.ti = *ti;
    return 0;
}

G6_API int g6_set_j_particle_(int *, int *address, int *index, double *tj, double *dtj,
                              double *mass, double k18[3], double j6[3], double a2[3],
                              double v[3], double x[3])
{
    return do_set_j(*address, *index, *tj, *dtj, *mass, k18, j6, a2, v, x);
}

G6_API void g6calc_firsthalf_(int *, int *nj, int *ni, int index[], double xi[][3],
                              double vi[][3], double (*)[3], double (*)[3], double *,
                              double *eps2, double h2[])
{
    U.first_rc = do_firsthalf(*nj, *ni, index, xi, vi, *eps2, h2);
}

G6_API int g6calc_lasthalf_(int *, int *, int *, int index[], double (*)[3], double (*)[3],
                            double *, double *, double acc[][3], double jerk[][3], double pot[])
{
    return do_lasthalf(index, acc, jerk, pot, nullptr);
}

G6_API int g6calc_lasthalf2_(int *, int *, int *, int index[], double (*)[3], double (*)[3],
                             double *, double *, double acc[][3], double jerk[][3],
                             double pot[], int nnbindex[])
{
    return do_lasthalf(index, acc, jerk, pot, nnbindex);
}

G6_API int g6_initialize_jp_buffer_(int *, int *) { return 0; }
// A no-op in both references (g6lib.c:447-450, sapporoG6lib.cpp:54); here it
// applies the staged set_j_particle records to the device store now rather
// than at the next calc (same results), e.g. before a stretch of
// g6x_calc_self_noalloc calls, which cannot grow the staging buffers.
G6_API int g6_flush_jp_buffer_(int *)
{
    if (!U.open) return 0;
    return flush_stage();
}

G6_API int g6_reset_(int *)
{
    int rc = check_open();
    if (rc) return rc;
    U.stage.clear();
    std::fill(U.stage_pos.begin(), U.stage_pos.end(), -1);
    clear_ids();
    HIPCHK(g6k::launch_reset(U.js, 0, U.js.cap, U.stream), "reset");
    HIPCHK(hipStreamSynchronize(U.stream), "reset sync");
    U.nj_hw = 0;
    U.pred_valid = false;
    U.k18_nonzero = false;
    return 0;
}

G6_API int g6_reset_fofpga_(int *) { return 0; }

// Neighbour lists (sapporo.cpp:234-272): computed on demand for the i-block
// of the last firsthalf against the same predicted j-set.  Returns 1 when
// some i has more than 256 neighbours (its list then holds 256 of them), 0
// otherwise, negative on error.
G6_API int g6_read_neighbour_list_(int *)
{
    int rc = check_open();
    if (rc) return rc;
    if (!U.latched) return fail(G6X_ERR_SEQUENCE, "read_neighbour_list without a calc");
    const int ni = U.l_ni;
    U.ngb_ni = 0;
    if (ni <= 0) return 0;
    if (ni > U.ngb_cap_i) {
        HIPCHK(hipStreamSynchronize(U.stream), "ngb sync");
        if (U.d_ngb_counts) (void)hipFree(U.d_ngb_counts);
        if (U.d_ngb_lists) (void)hipFree(U.d_ngb_lists);
        U.d_ngb_counts = nullptr;
        U.d_ngb_lists = nullptr;
        U.ngb_cap_i = 0;
        int cap = std::max(ni, 1024);
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&U.d_ngb_counts), sizeof(int) * size_t(cap));
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&U.d_ngb_lists), sizeof(int) * size_t(cap) * g6k::kNgbCap);
        if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc neighbour lists", e);
        U.ngb_cap_i = cap;
    }
    HIPCHK(hipMemsetAsync(U.d_ngb_counts, 0, sizeof(int) * size_t(ni), U.stream), "ngb clear");
    rc = ensure_predicted(U.l_ti);  // the force kernel may have predicted on the fly
    if (rc) return rc;
    if (U.in_direct) {
        // the query reads each i once per j-tile: the rows go to the device
        const size_t C = size_t(U.ib.cap);
        HIPCHK(hipMemcpy2DAsync(U.ib.x[0], sizeof(double) * C, U.h_in, sizeof(double) * C,
                                sizeof(double) * ni, 8, hipMemcpyHostToDevice, U.stream), "i H2D");
        HIPCHK(record_in(), "i record");
        U.in_direct = false;
    }
    HIPCHK(g6k::launch_ngb(U.js, U.ib, ni, U.l_nj, U.prec, U.d_ngb_counts, U.d_ngb_lists, U.stream), "ngb kernel");
    U.h_ngb_counts.resize(ni);
    U.h_ngb_lists.resize(size_t(ni) * g6k::kNgbCap);
    HIPCHK(hipMemcpyAsync(U.h_ngb_counts.data(), U.d_ngb_counts, sizeof(int) * size_t(ni), hipMemcpyDeviceToHost, U.stream), "ngb D2H");
    HIPCHK(hipMemcpyAsync(U.h_ngb_lists.data(), U.d_ngb_lists, sizeof(int) * size_t(ni) * g6k::kNgbCap, hipMemcpyDeviceToHost, U.stream), "ngb D2H");
    HIPCHK(hipStreamSynchronize(U.stream), "ngb sync");
    U.ngb_ni = ni;
    int overflow = 0;
    for (int i = 0; i < ni; i++) {
        const int c = U.h_ngb_counts[i];
        int *row = &U.h_ngb_lists[size_t(i) * g6k::kNgbCap];
        std::sort(row, row + std::min(c, g6k::kNgbCap));
        if (c > g6k::kNgbCap) overflow = 1;
    }
    return overflow;
}

// Copies the sorted neighbour ids of i-pipe *ipipe (at most *maxlength) and
// sets *n_neighbours to the full count; returns 1 when ids were left out
// (more than *maxlength, or more than the 256 kept), else 0.  The reference
// (sapporo.cpp:248-272) exits on a bad ipipe; this returns G6X_ERR_ARG.
G6_API int g6_get_neighbour_list_(int *, int *ipipe, int *maxlength, int *n_neighbours,
                                  int list[])
{
    int rc = check_open();
    if (rc) return rc;
    if (*ipipe < 0 || *ipipe >= U.ngb_ni) return fail(G6X_ERR_ARG, "ipipe outside the last neighbour read");
    if (*maxlength < 0) return fail(G6X_ERR_ARG, "negative maxlength");
    const int c = U.h_ngb_counts[*ipipe];
    const int kept = std::min(c, g6k::kNgbCap);
    const int ncopy = std::min(kept, *maxlength);
    const int *row = &U.h_ngb_lists[size_t(*ipipe) * g6k::kNgbCap];
    std::copy(row, row + ncopy, list);
    *n_neighbours = c;
    return (c > *maxlength || c > g6k::kNgbCap) ? 1 : 0;
}

// ===========================================================================
// C by-value ABI (weak: ph4's grape.cc defines the same wrappers)
// ===========================================================================
G6_WEAK int g6_open(int clusterid) { return g6_open_(&clusterid); }
G6_WEAK int g6_close(int clusterid) { return g6_close_(&clusterid); }
G6_WEAK int g6_npipes(void) { return g6_npipes_(); }
G6_WEAK int g6_set_tunit(double) { return 0; }
G6_WEAK int g6_set_xunit(double) { return 0; }
G6_WEAK int g6_set_ti(int clusterid, double ti) { return g6_set_ti_(&clusterid, &ti); }
G6_WEAK int g6_set_j_particle(int clusterid, int address, int index, double tj, double dtj,
                              double mass, double a2by18[3], double a1by6[3], double aby2[3],
                              double v[3], double x[3])
{
    return g6_set_j_particle_(&clusterid, &address, &index, &tj, &dtj, &mass, a2by18, a1by6, aby2,
                              v, x);
}
G6_WEAK void g6calc_firsthalf(int clusterid, int nj, int ni, int index[], double xi[][3],
                              double vi[][3], double fold[][3], double jold[][3], double phiold[],
                              double eps2, double h2[])
{
    g6calc_firsthalf_(&clusterid, &nj, &ni, index, xi, vi, fold, jold, phiold, &eps2, h2);
}
G6_WEAK int g6calc_lasthalf(int clusterid, int nj, int ni, int index[], double xi[][3],
                            double vi[][3], double eps2, double h2[], double acc[][3],
                            double jerk[][3], double pot[])
{
    return g6calc_lasthalf_(&clusterid, &nj, &ni, index, xi, vi, &eps2, h2, acc, jerk, pot);
}
G6_WEAK int g6calc_lasthalf2(int clusterid, int nj, int ni, int index[], double xi[][3],
                             double vi[][3], double eps2, double h2[], double acc[][3],
                             double jerk[][3], double pot[], int nnbindex[])
{
    return g6calc_lasthalf2_(&clusterid, &nj, &ni, index, xi, vi, &eps2, h2, acc, jerk, pot,
                             nnbindex);
}
G6_WEAK int g6_initialize_jp_buffer(int clusterid, int buf_size)
{
    return g6_initialize_jp_buffer_(&clusterid, &buf_size);
}
G6_WEAK int g6_flush_jp_buffer(int clusterid) { return g6_flush_jp_buffer_(&clusterid); }
G6_WEAK int g6_reset(int clusterid) { return g6_reset_(&clusterid); }
G6_WEAK int g6_reset_fofpga(int clusterid) { return g6_reset_fofpga_(&clusterid); }
G6_WEAK int g6_read_neighbour_list(int clusterid) { return g6_read_neighbour_list_(&clusterid); }
G6_WEAK int g6_get_neighbour_list(int clusterid, int ipipe, int maxlength, int *n_neighbours,
                                  int neighbour_list[])
{
    return g6_get_neighbour_list_(&clusterid, &ipipe, &maxlength, n_neighbours, neighbour_list);
}
G6_WEAK void g6_set_neighbour_list_sort_mode(int) {}
G6_WEAK int g6_get_neighbour_list_sort_mode(void) { return 1; }
G6_WEAK int g6_set_overflow_flag_test_mode(int, int, int) { return 0; }
G6_WEAK int g6_get_number_of_pipelines(void) { return g6_npipes_(); }
G6_WEAK void g6_reinitialize(int clusterid) { g6_reset_(&clusterid); }

G6_WEAK int get_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
G6_WEAK void force_j_particle_send(void) {}
G6_WEAK int get_j_part_data(int, int, double *, double *, double *, double *, double *, double *)
{
    return 0;
}

// ===========================================================================
// Extensions (g6x.h)
// ===========================================================================
G6_API int g6x_set_precision(int mode)
{
    if (mode != G6X_FP64 && mode != G6X_MIXED) return fail(G6X_ERR_ARG, "precision mode");
    U.prec = mode;
    return 0;
}
G6_API int g6x_get_precision(void) { return U.prec; }
G6_API int g6x_set_predictor_compat(int on)
{
    U.compat = on ? 1 : 0;
    return 0;
}
G6_API int g6x_set_npipes(int npipes)
{
    if (npipes < 1 || npipes > kMaxPipes) return fail(G6X_ERR_ARG, "npipes out of range");
    U.npipes = npipes;
    return 0;
}
G6_API const char *g6x_last_error(void) { return U.err.c_str(); }

// Full passes at least this long take the spatial order (mixed path).
constexpr int kSpatialMin = 1 << 17;

int ensure_spatial(int nj)
{
    if (nj > U.sp_cap) {
        HIPCHK(hipStreamSynchronize(U.stream), "spatial sync");
        if (U.sp_base) (void)hipFree(U.sp_base);
        U.sp_base = nullptr;
        U.sp_cap = 0;
        int cap = 1 << 16;
        while (cap < nj) cap *= 2;
        const size_t js = g6k::jstore_bytes(cap);
        const size_t tiles = size_t(cap / g6k::kTJ + 1);
        void *b = nullptr;
        hipError_t e = hipMalloc(&b, js + sizeof(int) * size_t(cap) + sizeof(float) * 6 * tiles + 512);
        if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc spatial store", e);
        U.sp_base = b;
        g6k::jstore_carve(U.sj, b, cap);
        char *q = static_cast<char *>(b) + ((js + 255) & ~size_t(255));
        U.sp_perm = reinterpret_cast<int *>(q);
        q += (sizeof(int) * size_t(cap) + 255) & ~size_t(255);
        U.sp_tbox = reinterpret_cast<float *>(q);
        U.sp_cap = cap;
    }
    if (U.ib.cap > U.sp_rows_cap) {
        HIPCHK(hipStreamSynchronize(U.stream), "spatial sync");
        if (U.sp_rows) (void)hipFree(U.sp_rows);
        U.sp_rows = nullptr;
        U.sp_rows_cap = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&U.sp_rows),
                                 sizeof(double) * 8 * size_t(U.ib.cap));
        if (e != hipSuccess) return fail(G6X_ERR_NOMEM, "hipMalloc spatial rows", e);
        U.sp_rows_cap = U.ib.cap;
    }
    return 0;
}

// Full pass (i = every slot) of the mixed path in spatial order: the slots
// sorted along a Morton curve (g6k::spatial_order), the i-block taken from
// the sorted copy, the force kernel for spatially ordered slots, the fold,
// and the result rows put back in slot order.  False with nothing done when
// the shape has no spatial kernel.
int run_force_spatial(int nj, double eps2, bool want_nn, bool &done)
{
    done = false;
    const ForceCfg cfg = choose_cfg(nj, nj);
    if (!g6k::spatial_cfg_supported(cfg)) return 0;
    int rc = ensure_spatial(nj);
    if (rc) return rc;
    rc = ensure_partials(cfg.nchunk, nj);
    if (rc) return rc;
    Unit::Ev ev0;
    timed_begin(ev0, 2, 0.0);
    HIPCHK(g6k::spatial_order(U.js, nj, U.sj, U.sp_perm, U.sp_tbox, U.sp_buf, U.stream),
           "spatial order");
    HIPCHK(g6k::launch_iblock_self(U.sj, U.ib, 0, nj, g6k::kMixed, U.stream), "iblock sorted");
    timed_end(ev0);
    IBlock b = U.ib;
    b.out = U.sp_rows;
    b.nn = reinterpret_cast<int *>(U.sp_rows + 7 * size_t(U.ib.cap));
    Unit::Ev ev;
    timed_begin(ev, 0, double(nj) * double(nj));
    HIPCHK(g6k::launch_force(U.sj, b, nj, nj, eps2, g6k::kMixed, want_nn, cfg, U.stream, false,
                             0.0, U.compat, -1, U.sp_tbox),
           "force kernel (spatial)");
    timed_end(ev);
    Unit::Ev ev2;
    timed_begin(ev2, 1, 0.0);
    if (cfg.nchunk > 1)
        HIPCHK(g6k::launch_reduce(U.sj, b, nj, cfg.nchunk, g6k::kMixed, want_nn, U.stream),
               "reduce kernel");
    HIPCHK(g6k::launch_unsort_rows(b.out, b.nn, U.ib.cap, U.sp_perm, nj, U.ib.out, U.ib.nn,
                                   U.ib.cap, want_nn, U.stream),
           "unsort rows");
    timed_end(ev2);
    U.out_direct = false;
    U.launched = true;
    U.launched_nn = want_nn;
    U.last_ni = nj;
    done = true;
    return 0;
}

G6_API int g6x_set_spatial(int on)
{
    U.spatial = on != 0;
    return 0;
}

G6_API int g6x_calc_self(int nj, int i0, int ni, double eps2, int want_nn)
{
    int rc = check_open();
    if (rc) return rc;
    if (ni < 0 || i0 < 0 || nj < 0) return fail(G6X_ERR_ARG, "negative size");
    if (nj > U.js.cap) {
        rc = flush_stage();
        if (rc) return rc;
        rc = grow_jstore(nj);
        if (rc) return rc;
    }
    if (nj > U.nj_hw) U.nj_hw = nj;
    if ((long long)i0 + ni > U.js.cap) return fail(G6X_ERR_ARG, "i-range beyond j-store");
    rc = grow_iblock(std::max(ni, 1));
    if (rc) return rc;
    rc = flush_stage();
    if (rc) return rc;
    rc = ensure_predicted(U.ti);
    if (rc) return rc;
    U.latched = false;
    U.in_direct = false;  // the i-block is in the device rows
    const bool distinct = U.id_repeats == 0 && U.n_filled == nj && U.nj_hw == nj;
    if (U.spatial && distinct && U.prec == g6k::kMixed && i0 == 0 && ni == nj &&
        nj >= kSpatialMin) {
        bool done = false;
        rc = run_force_spatial(nj, eps2, want_nn != 0, done);
        if (rc || done) return rc;
    }
    HIPCHK(g6k::launch_iblock_self(U.js, U.ib, i0, ni, U.prec, U.stream), "iblock self");
    return run_force(ni, nj, eps2, want_nn != 0, distinct ? i0 : -1);
}

// ---------------------------------------------------------------------------
// Full pass without allocation (g6x_calc_self_noalloc; scratch_setup and the
// array are defined before g6_open)
// ---------------------------------------------------------------------------
G6_API int g6x_calc_self_noalloc(int nj, int i0, int ni, double eps2, double acc[][3],
                                 double jerk[][3], double pot[], int nn[])
{
    int rc = check_open();
    if (rc) return rc;
    if (ni < 0 || i0 < 0 || nj < 0) return fail(G6X_ERR_ARG, "negative size");
    if (ni > G6X_NOALLOC_MAX_NI)
        return fail(G6X_ERR_ARG, "ni beyond G6X_NOALLOC_MAX_NI (the static scratch)");
    if (nj > U.js.cap || (long long)i0 + ni > U.js.cap)
        return fail(G6X_ERR_ARG, "slots beyond the j-store's capacity (the allocation-free "
                                 "call cannot grow it)");
    U.latched = false;
    U.launched = false;
    U.last_ni = 0;  // the results go to the caller's arrays, not to g6x_fetch
    if (ni == 0) return 0;
    if (!acc || !jerk || !pot) return fail(G6X_ERR_ARG, "null result array");
    rc = scratch_setup(U.device);
    if (rc) return rc;
    double *base = scratch_ptr[U.device];
    if (nj > U.nj_hw) U.nj_hw = nj;
    rc = flush_stage(false);
    if (rc) return rc;
    rc = ensure_predicted(U.ti, false);
    if (rc) return rc;
    // rows at pitch ni: x3 v3 h2 id | acc3 jerk3 pot nn | partials
    IBlock b;
    const size_t S = size_t(ni);
    b.cap = ni;
    for (int k = 0; k < 3; k++) {
        b.x[k] = base + k * S;
        b.v[k] = base + (3 + k) * S;
    }
    b.h2 = base + 6 * S;
    b.id = reinterpret_cast<int *>(base + 7 * S);
    b.out = base + 8 * S;
    b.nn = reinterpret_cast<int *>(base + 15 * S);
    b.part = base + 16 * S;
    const long long fields = U.prec == g6k::kF64 ? g6k::kPartFieldsD : g6k::kPartFields;
    const long long room = (long long)(kScratchDoubles - 16 * S) / fields;
    b.part_entries = room;
    ForceCfg cfg = choose_cfg(ni, std::max(nj, 1));
    if (cfg.nchunk > 1 && (long long)cfg.nchunk * ni > room)
        cfg = with_chunks(cfg, cfg.nchunk, ni, std::max(nj, 1), room);
    const bool want_nn = nn != nullptr;
    HIPCHK(g6k::launch_iblock_self(U.js, b, i0, ni, U.prec, U.stream), "iblock self");
    const bool distinct = U.id_repeats == 0 && U.n_filled == nj && U.nj_hw == nj;
    HIPCHK(g6k::launch_force(U.js, b, ni, std::max(nj, 1), eps2, U.prec, want_nn, cfg, U.stream,
                             false, 0.0, U.compat, distinct ? i0 : -1),
           "force kernel");
    if (cfg.nchunk > 1)
        HIPCHK(g6k::launch_reduce(U.js, b, ni, cfg.nchunk, U.prec, want_nn, U.stream),
               "reduce kernel");
    HIPCHK(g6k::launch_rows_to_aos(b, ni, &acc[0][0], &jerk[0][0], pot, nn, U.stream),
           "result rows");
    return 0;
}

G6_API int g6x_fetch(int ni, double acc[][3], double jerk[][3], double pot[], int nn[])
{
    int rc = check_open();
    if (rc) return rc;
    if (ni < 0 || ni > U.last_ni) return fail(G6X_ERR_ARG, "fetch beyond last result");
    return fetch_results(ni, acc, jerk, pot, nn, nn != nullptr && U.launched_nn);
}

G6_API int g6x_sync(void)
{
    int rc = check_open();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(U.stream), "sync");
    return 0;
}

G6_API int g6x_timing(int enable)
{
    if (U.stream) (void)hipStreamSynchronize(U.stream);
    harvest_timing();
    U.timing = enable != 0;
    U.t_force = U.t_reduce = U.t_pred = U.t_pairs = 0.0;
    U.n_force = 0;
    return 0;
}

G6_API int g6x_get_timing(double *force_ms, double *reduce_ms, double *predict_ms,
                          long long *force_launches, double *pairs)
{
    if (U.stream) (void)hipStreamSynchronize(U.stream);
    harvest_timing();
    if (force_ms) *force_ms = U.t_force;
    if (reduce_ms) *reduce_ms = U.t_reduce;
    if (predict_ms) *predict_ms = U.t_pred;
    if (force_launches) *force_launches = U.n_force;
    if (pairs) *pairs = U.t_pairs;
    return 0;
}

G6_API int g6x_set_tiles(int njs, int ipl, int nchunk, int smem)
{
    if (njs == 0) {
        U.tile_njs = U.tile_ipl = U.tile_nchunk = U.tile_smem = 0;
        return 0;
    }
    if (!g6k::force_cfg_supported(njs, ipl) || nchunk < 0 || (smem && (njs != 1 || ipl > 2)))
        return fail(G6X_ERR_ARG, "tile config");
    U.tile_njs = njs;
    U.tile_ipl = ipl;
    U.tile_nchunk = nchunk;
    U.tile_smem = smem ? 1 : 0;
    return 0;
}

G6_API int g6x_set_j_particles(int n, const int address[], const int index[], const double tj[],
                               const double dtj[], const double mass[], const double k18[][3],
                               const double j6[][3], const double a2[][3], const double v[][3],
                               const double x[][3])
{
    if (n < 0) return fail(G6X_ERR_ARG, "negative count");
    for (int q = 0; q < n; q++) {
        int rc = do_set_j(address[q], index[q], tj[q], dtj[q], mass[q], k18[q], j6[q], a2[q], v[q],
                          x[q]);
        if (rc) return rc;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: for each power-of-two ni bucket up to ni_max, time every
// compiled (njs, ipl) shape with the automatic chunk count and 2x / 0.5x
// of it on the current j-store (i-block = the first ni predicted slots),
// and keep the fastest.  Timing uses events on the unit's stream.
// ---------------------------------------------------------------------------
G6_API int g6x_autotune(int nj, int ni_max, double eps2)
{
    int rc = check_open();
    if (rc) return rc;
    if (nj <= 0 || ni_max <= 0 || ni_max > nj) return fail(G6X_ERR_ARG, "autotune sizes");
    const int save_njs = U.tile_njs, save_ipl = U.tile_ipl, save_nc = U.tile_nchunk,
              save_smem = U.tile_smem;
    const bool save_timing = U.timing;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t ee = hipEventCreate(&a);
    if (ee == hipSuccess) ee = hipEventCreate(&b);
    if (ee != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        return fail(G6X_ERR_HIP, "event", ee);
    }
    U.timing = false;
    U.tune.clear();
    // {njs, ipl, smem}
    static const int kShapes[][3] = {{1, 2, 1}, {1, 1, 1}, {1, 2, 0}, {1, 1, 0}, {2, 2, 0},
                                     {2, 1, 0}, {4, 2, 0}, {4, 1, 0}, {8, 2, 0}, {8, 1, 0},
                                     {16, 1, 0}, {32, 1, 0}, {64, 1, 0}, {128, 1, 0},
                                     {256, 1, 0}};
    // a failure below still restores the unit's tile settings and timing
#define TUNECHK(expr, what)                        \
    do {                                           \
        const hipError_t e_ = (expr);              \
        if (e_ != hipSuccess) {                    \
            rc = fail(G6X_ERR_HIP, what, e_);      \
            goto done;                             \
        }                                          \
    } while (0)
    int timed = 0;
    std::vector<TuneEntry> table;
    for (int ni = 1; ni <= ni_max; ni *= 2) {
        float best_ms = 1e30f;
        TuneEntry best{ni, ni, 0, 0, 0, 0};
        for (const auto &sh : kShapes) {
            const int ib = (g6k::kWG / sh[0]) * sh[1];
            if (ib > 4 * ni && ib > 1) continue;  // mostly idle lanes
            U.tile_njs = sh[0];
            U.tile_ipl = sh[1];
            U.tile_smem = sh[2];
            U.tile_nchunk = 0;
            const int auto_nc = choose_cfg(ni, nj).nchunk;
            const int ncs[3] = {auto_nc, std::max(1, auto_nc / 2), auto_nc * 2};
            for (int v = 0; v < 3; v++) {
                if (v > 0 && ncs[v] == auto_nc) continue;
                U.tile_nchunk = ncs[v];
                rc = g6x_calc_self(nj, 0, ni, eps2, 1);  // warm-up / allocation
                if (rc) goto done;
                TUNECHK(hipEventRecord(a, U.stream), "record");
                const int reps = 3;
                for (int r = 0; r < reps; r++) {
                    rc = run_force(ni, nj, eps2, true);
                    if (rc) goto done;
                }
                TUNECHK(hipEventRecord(b, U.stream), "record");
                TUNECHK(hipEventSynchronize(b), "sync");
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, a, b);
                ms /= reps;
                timed++;
                if (ms < best_ms) {
                    best_ms = ms;
                    best.njs = sh[0];
                    best.ipl = sh[1];
                    best.smem = sh[2];
                    best.nchunk = choose_cfg(ni, nj).nchunk;
                }
            }
        }
        table.push_back(best);
    }
    // buckets: (prev power, this power]
    for (size_t k = 0; k < table.size(); k++) {
        table[k].ni_lo = k == 0 ? 1 : table[k - 1].ni_hi + 1;
    }
    table.back().ni_hi = kMaxPipes;
done:
#undef TUNECHK
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    U.tile_njs = save_njs;
    U.tile_ipl = save_ipl;
    U.tile_nchunk = save_nc;
    U.tile_smem = save_smem;
    U.timing = save_timing;
    if (rc) return rc;
    U.tune = table;
    return timed;
}

G6_API int g6x_save_tuning(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return fail(G6X_ERR_ARG, "cannot write tuning file");
    fprintf(f, "# ni_lo ni_hi njs ipl nchunk smem\n");
    for (const auto &t : U.tune)
        fprintf(f, "%d %d %d %d %d %d\n", t.ni_lo, t.ni_hi, t.njs, t.ipl, t.nchunk, t.smem);
    fclose(f);
    return int(U.tune.size());
}

G6_API int g6x_load_tuning(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(G6X_ERR_ARG, "cannot read tuning file");
    std::vector<TuneEntry> table;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        TuneEntry t;
        if (sscanf(line, "%d %d %d %d %d %d", &t.ni_lo, &t.ni_hi, &t.njs, &t.ipl, &t.nchunk,
                   &t.smem) == 6 &&
            g6k::force_cfg_supported(t.njs, t.ipl) && t.ni_lo <= t.ni_hi && t.nchunk >= 0 &&
            (!t.smem || (t.njs == 1 && t.ipl <= 2)))
            table.push_back(t);
    }
    fclose(f);
    U.tune = table;
    return int(table.size());
}

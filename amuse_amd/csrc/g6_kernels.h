// Device data layout and kernel launchers of the GRAPE-6 force path.
//
// HBM layout (all structure-of-arrays, one array per component, indexed by
// j-slot so every kernel reads and writes contiguous, coalesced words):
//
//   base j-store (FP64, written by the scatter kernel from set_j_particle):
//     x[3] v[3] a2[3] j6[3] k18[3] tj dtj m        17 doubles / slot
//     id                                            int32
//   predicted, plain path:   px[3] pv[3] pm         7 doubles / slot
//   predicted, mixed path:   hx[3] lx[3] (double-single position, float hi
//                            + float lo), fv[3], fm 10 floats / slot
//   pid                       id, with empty slots (-1) mapped to zero mass
//                            and a far position (shared by both paths)
//   i-block (FP64):           x[3] v[3] h2, id, capacity = npipes
//   results:                  out[7][cap] (acc xyz, jerk xyz, pot), nn[cap]
//   partials (j-split over workgroups): part[chunk][16][cap] (plain path,
//                            double-double) or [chunk][9][cap] (mixed)
//
// Reference for the arithmetic: lib/g6lib/g6lib.c:64-186 (predictor, pair
// loop); the launch shapes are this library's own.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace g6k {

constexpr int kWG = 256;        // threads per workgroup (4 wave64)
constexpr int kTJ = 256;        // j per LDS tile
constexpr int kPartFields = 9;   // mixed: (hi,lo) float pairs of acc3 jerk3 pot, r2, slot
constexpr int kPartFieldsD = 16; // plain: hi, lo of acc3 jerk3 pot, r2, slot
constexpr double kFar = 1.0e150;  // position given to empty slots
// The reduce runs one wave per i from kReduceWaveMin chunks up to
// kReduceWaveMaxNi i-particles (beyond that one lane per i has enough waves
// in flight to hide the serial chunk loop).
constexpr int kReduceWaveMin = 16;
constexpr int kReduceWaveMaxNi = 8192;
// From kReduceRowMin chunks up to kReduceRowMaxNi rows the library's
// reduce kernel takes a whole workgroup per row instead (ABI path; at
// nj = 65536 force + fold -14..-20 % FP64 and -5..-11 % mixed for ni =
// 4-128, while from ni = 512 the wave form's ni / 4 workgroups were faster).
constexpr int kReduceRowMin = 128;
constexpr int kReduceRowMaxNi = 128;
constexpr int kNgbCap = 256;      // neighbour ids kept per i (sapporo NGB_PP)

struct JStore {
    int cap = 0;
    double *x[3] = {}, *v[3] = {}, *a2[3] = {}, *j6[3] = {}, *k18[3] = {};
    double *tj = nullptr, *dtj = nullptr, *m = nullptr;
    int *id = nullptr;
    double *px[3] = {}, *pv[3] = {}, *pm = nullptr;
    float *hx[3] = {}, *lx[3] = {}, *fv[3] = {}, *fm = nullptr;
    int *pid = nullptr;
    void *base = nullptr;   // single allocation holding everything above
};

// One staged set_j_particle call.
struct JRecord {
    double tj, dtj, m;
    double x[3], v[3], a2[3], j6[3], k18[3];
    int addr, id;
};

struct IBlock {
    int cap = 0;
    double *x[3] = {}, *v[3] = {}, *h2 = nullptr;
    int *id = nullptr;
    double *out = nullptr;   // [7][cap]
    int *nn = nullptr;       // [cap]
    double *part = nullptr;  // [nchunk][fields][ni] of the last launch
    long long part_entries = 0;  // (chunk, i) records the buffer holds
    void *base = nullptr;
    void *part_base = nullptr;
};

enum Prec { kF64 = 0, kMixed = 1 };

struct ForceCfg {
    int njs = 1;     // lanes sharing one i-particle (j-split inside a WG)
    int ipl = 1;     // i-particles per lane
    int nchunk = 1;  // workgroups along j
    int jchunk = 0;  // j per chunk (multiple of kTJ)
    bool smem = false;  // njs == 1 only: j read through scalar loads, no LDS
};

// ------------------------------------------------------------------------
// Block steps decided on the device (the device-resident driver,
// g6_blocks.hip).  The active count of a step exists only on the GPU, so
// the predictor, force and reduce kernels of a step read it from this
// control block and size their work themselves; one captured launch
// sequence then serves every step.
//
// Protocol of one step k (three kernels; par = k & 1 is a launch argument,
// fixed in the captured sequence, so a captured graph holds an even number
// of steps): the selection kernel ends the run (system_time >= t_end, or the
// block budget spent) or turns tmin[par] into tnext, appends the block to
// ni[par] and folds min(t + dt) of the other particles into tmin[par ^ 1];
// the force kernel resets tmin[par] (consumed); the correction kernel folds
// the force partials, corrects, folds the new t + dt into tmin[par ^ 1],
// resets ni[par ^ 1] for the next step and advances the counters and cur.
// Each field is reset by a kernel that no other kernel of the same launch
// reads it in.  ni[2] serves full passes (energies, initial forces), which
// leave the step fields alone.
struct StepCtl {
    unsigned long long tmin[2]; // bits of min(t + dt) (positive doubles order as integers)
    int ni[3];                  // active particles: steps of parity 0 and 1, full passes
    int cur;                    // parity of the next step (tmin[cur] holds its time)
    int done;                   // 1: t_end reached, 2: block budget spent
    long long blocks, psteps;   // block and particle steps taken
    long long budget;           // block steps left in the current run (< 0: no limit)
    double t_end;
    double system_time;
    double tnext;               // block time of the current step
};

struct StepPlan {
    int shape;    // 0..3: (lanes per i, i per lane) = (1,2) (4,1) (16,1) (64,1)
    int iblocks;  // i-blocks of the shape's width
    int nchunk;   // j-chunks (partials reduced when > 1)
    int jchunk;   // j per chunk (multiple of kTJ)
};

// work items a large step aims at (build-time tunable) and the nchunk * ni
// bound on its partials (4096: +2 % on N = 2^20 block steps against 2048,
// the same at N = 16384 and 65536; 8192 no better)
#ifndef G6_STEP_ITEMS
#define G6_STEP_ITEMS 4096
#endif
#ifndef G6_STEP_PART_LOG2
#define G6_STEP_PART_LOG2 22
#endif
constexpr int kStepItems = G6_STEP_ITEMS;
constexpr long long kStepPartBudget = 1ll << G6_STEP_PART_LOG2;
constexpr int kStepGrid = 2048;                   // persistent workgroups of a step

__host__ __device__ inline int step_ib(int shape)
{
    return shape == 0 ? 512 : shape == 1 ? 64 : shape == 2 ? 16 : 4;
}
__host__ __device__ inline int step_njs(int shape)
{
    return shape == 0 ? 1 : shape == 1 ? 4 : shape == 2 ? 16 : 64;
}

// Launch plan of one step.  Small systems (nj <= 2048, at most 8 tiles):
// 64 lanes per i and no j-split over workgroups, so a step is one short
// force kernel and no reduce.  Otherwise the rule of the ABI path's
// automatic choice: while ni * nj <= 2^29 (latency-bound) the widest
// i-block of at most 64 i that gives >= 256 workgroups; beyond, the widest
// i-block not wider than the next power of two of ni; then j-chunks for
// ~kStepItems work items within the partials budget.
__host__ __device__ inline StepPlan plan_step(int ni, int nj)
{
    StepPlan sp;
    if (nj <= 8 * kTJ) {
        sp.shape = 3;
        sp.iblocks = (ni + step_ib(3) - 1) / step_ib(3);
        sp.nchunk = 1;
        sp.jchunk = nj > 0 ? ((nj + kTJ - 1) / kTJ) * kTJ : kTJ;
        return sp;
    }
    int p2 = 1;
    while (p2 < ni) p2 *= 2;
    sp.shape = 3;
    const int max_chunks = (nj + kTJ - 1) / kTJ;
    if ((long long)ni * nj <= (1ll << 29)) {
        // latency-bound: the widest of this kernel's i-blocks of at most 64
        // i that gives >= 256 workgroups with the chunks the j-range allows
        // (the library's automatic choice on this kernel's shapes)
        for (int k = 1; k < 4; k++) {
            sp.shape = k;
            const int ib = step_ib(k);
            if (ib > p2 && k < 3) continue;
            const long long iblocks = (ni + ib - 1) / ib;
            const long long want = (kStepItems + iblocks - 1) / iblocks;
            if (iblocks * (want < max_chunks ? want : max_chunks) >= 256) break;
        }
    } else {
        for (int k = 0; k < 4; k++)
            if (step_ib(k) <= p2) { sp.shape = k; break; }
    }
    const int ib = step_ib(sp.shape);
    sp.iblocks = (ni + ib - 1) / ib;
    int nchunk = sp.iblocks > 0 ? (kStepItems + sp.iblocks - 1) / sp.iblocks : 1;
    if (nchunk > max_chunks) nchunk = max_chunks;
    while (nchunk > 1 && (long long)nchunk * ni > kStepPartBudget) nchunk /= 2;
    int jchunk = (nj + nchunk - 1) / nchunk;
    jchunk = ((jchunk + kTJ - 1) / kTJ) * kTJ;
    sp.jchunk = jchunk;
    sp.nchunk = (nj + jchunk - 1) / jchunk;
    return sp;
}

// Predicts every j-slot to the step's block time c->tnext (at_system = 0)
// or to c->system_time (1); no-op once the run is done.
hipError_t launch_step_predict(const JStore &s, int nj, const StepCtl *c, int at_system,
                               int prec, int compat, hipStream_t st);
// Forces (acc, jerk, pot; no nn) of the step's i-block (b rows 0..ni-1,
// ni = c->ni[ni_slot]) against j-slots [0, nj): into b.out when the plan has
// one j-chunk, else as partials in part (capacity kStepPartBudget * fields
// doubles), which a second kernel reduces into b.out when with_reduce (the
// block step's correction kernel folds them itself).  tmin_slot >= 0: the
// force kernel resets c->tmin[tmin_slot] (see StepCtl).
hipError_t launch_step_force(const JStore &s, const IBlock &b, double *part, int nj,
                             double eps2, int prec, StepCtl *c, int ni_slot, int tmin_slot,
                             bool with_reduce, hipStream_t st);

size_t jstore_bytes(int cap);
void jstore_carve(JStore &s, void *base, int cap);
size_t iblock_bytes(int cap);
void iblock_carve(IBlock &b, void *base, int cap);

hipError_t launch_reset(const JStore &s, int lo, int hi, hipStream_t st);
hipError_t launch_copy_store(const JStore &dst, const JStore &src, int n,
                             hipStream_t st);
hipError_t launch_scatter(const JStore &s, const JRecord *rec, int n,
                          int prec, int pred_valid, double ti, int compat,
                          hipStream_t st);
hipError_t launch_predict(const JStore &s, int nj, double ti, int prec,
                          int compat, int use_k18, hipStream_t st);
hipError_t launch_iblock_self(const JStore &s, const IBlock &b, int i0,
                              int ni, int prec, hipStream_t st);
// Valid (njs, ipl) pairs compiled in.
bool force_cfg_supported(int njs, int ipl);
// pred_j: the j-slots are predicted to ti (compat mode as given, k18 = 0)
// from their base records as the kernel loads them, and the predicted store
// is neither read nor written; only for pred_force_cfg launches.
// self_i0 >= 0: i row r is j-slot self_i0 + r and the nj slots hold
// distinct ids, none empty, so work items whose j-range holds none of their
// i-slots skip the per-pair id compare; -1: every pair compares ids.
// tbox != nullptr: the slots are in spatial order and tbox holds each
// kTJ-slot tile's position box (spatial_order); mixed path, shapes of
// spatial_cfg_supported only.
hipError_t launch_force(const JStore &s, const IBlock &b, int ni, int nj,
                        double eps2, int prec, bool want_nn,
                        const ForceCfg &cfg, hipStream_t st, bool pred_j = false,
                        double ti = 0.0, int compat = 0, int self_i0 = -1,
                        const float *tbox = nullptr);
bool spatial_cfg_supported(const ForceCfg &c);

// Spatial order of a full pass (g6_spatial.hip): the mixed predicted slots
// [0, nj) of `s` (all holding distinct ids) sorted along a Morton curve of
// their positions into `o` (mixed predicted arrays, id and pid), perm[k] =
// the slot at sorted position k, tbox[6 * t ...] = min xyz, max xyz of sorted
// tile t (kTJ slots).  SpatialBuf holds the keys and the sort's temporary
// storage, grown on demand (the only allocations).
struct SpatialBuf {
    void *base = nullptr;
    size_t bytes = 0;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
};
hipError_t spatial_order(const JStore &s, int nj, const JStore &o, int *perm, float *tbox,
                         SpatialBuf &buf, hipStream_t st);
// Result rows of a spatially ordered pass (pitch S_in, sorted order) to slot
// order (pitch S_out): out[k][perm[i]] = in[k][i], nn likewise.
hipError_t launch_unsort_rows(const double *in, const int *in_nn, int S_in, const int *perm,
                              int ni, double *out, int *out_nn, int S_out, bool nn,
                              hipStream_t st);
// Launches that can predict their j-slots as they load them: one i per
// wave (64 lanes), one j-chunk, at most 8 tiles.
inline bool pred_force_cfg(const ForceCfg &c, int nj)
{
    return c.njs == 64 && c.ipl == 1 && c.nchunk == 1 && !c.smem && nj <= 8 * kTJ;
}
hipError_t launch_reduce(const JStore &s, const IBlock &b, int ni,
                         int nchunk, int prec, bool want_nn, hipStream_t st);
// The result rows of b (pitch b.cap) into caller-owned device arrays
// acc[ni][3], jerk[ni][3], pot[ni] and nn[ni] (nn may be null).
hipError_t launch_rows_to_aos(const IBlock &b, int ni, double *acc, double *jerk,
                              double *pot, int *nn, hipStream_t st);
// Neighbour query over the latched i-block: ids of j with unsoftened
// r^2 <= h2[i] (self and empty slots excluded).  counts[i] is the full count
// (must be zeroed first); at most kNgbCap ids per i are stored, in no
// particular order.
hipError_t launch_ngb(const JStore &s, const IBlock &b, int ni, int nj, int prec,
                      int *counts, int *lists, hipStream_t st);

}  // namespace g6k

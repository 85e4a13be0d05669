// Two-level grid barrier of the persistent ensemble kernel (kc_run).
//
// One counter for every workgroup of the grid costs ~0.1 ms per barrier at
// ~1300 workgroups (device-scope atomics on one address serialise).  Here
// workgroups arrive in groups of kGroup on their group's counter; the last
// of a group arrives on the top counter; the last group's last workgroup
// resets both and opens the barrier by advancing the generation word, which
// the others poll.  Counters sit kStride words (256 bytes) apart.  A waiter
// gives up after kSpinLimit polls and raises the abort word, which every
// waiter polls as well, so a launch always ends.
//
// Layout of the kWords words (all zero at the start of a launch): group
// counters at g * kStride, top counter, generation word, abort word.
//
// The counting is written once, over an operations policy: the kernel uses
// HIP's device-scope atomics and fences, the CPU test
// (tests/gridbar_check.cpp) std::thread workgroups and GCC atomics.
#pragma once

#if defined(__HIPCC__)
#define G6BAR_HD __host__ __device__
#else
#define G6BAR_HD
#endif

namespace g6bar {

constexpr int kGroup = 32;
constexpr int kStride = 64;
constexpr int kMaxGroups = 128;  // grids of up to kGroup * kMaxGroups workgroups
constexpr int kWords = (kMaxGroups + 3) * kStride;
constexpr unsigned kSpinLimit = 1u << 24;

G6BAR_HD constexpr int top_word() { return kMaxGroups * kStride; }
G6BAR_HD constexpr int gen_word() { return (kMaxGroups + 1) * kStride; }
G6BAR_HD constexpr int abort_word() { return (kMaxGroups + 2) * kStride; }

// Called by one thread per workgroup, after the workgroup's own writes are
// ordered before it.  gen = barriers this workgroup passed in this launch.
// Returns false when the launch was aborted (SpinLimit polls without the
// barrier opening, here or in another workgroup).  Ops:
//   release_fence(), acquire_fence()        agent-scope fences
//   fetch_add(p, v) -> old                   acquire-release read-modify-write
//   store(p, v), store_release(p, v)
//   poll(p)                                  load of a word other workgroups write
//   pause()                                  back-off between polls
template <class Ops, unsigned SpinLimit = kSpinLimit>
G6BAR_HD inline bool arrive_and_wait(unsigned *bar, int block, int grid, unsigned gen)
{
    const int g = block / kGroup;
    const int rest = grid - g * kGroup;
    const unsigned gsize = unsigned(rest < kGroup ? rest : kGroup);
    const unsigned ngroups = unsigned((grid + kGroup - 1) / kGroup);
    unsigned *cnt = bar + g * kStride;
    unsigned *top = bar + top_word();
    unsigned *gw = bar + gen_word();
    unsigned *ab = bar + abort_word();
    Ops::release_fence();
    if (Ops::fetch_add(cnt, 1u) == gsize - 1u) {
        Ops::store(cnt, 0u);  // nobody of this group arrives again before gw moves
        if (Ops::fetch_add(top, 1u) == ngroups - 1u) {
            Ops::store(top, 0u);
            Ops::store_release(gw, gen + 1u);
        }
    }
    bool ok = true;
    for (unsigned spins = 0; Ops::poll(gw) == gen; spins++) {
        if (Ops::poll(ab) != 0u || spins >= SpinLimit) {
            Ops::store(ab, 1u);
            ok = false;
            break;
        }
        Ops::pause();
    }
    Ops::acquire_fence();
    return ok;
}

}  // namespace g6bar

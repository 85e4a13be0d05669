// CPU check of the persistent ensemble kernel's grid barrier
// (amuse_amd/csrc/g6_gridbar.h): the same counting code, one std::thread per
// workgroup.  Between barriers every workgroup writes its slot of a plain
// array; after each barrier every workgroup reads all slots and must see the
// values of this phase (the barrier publishes them), and no slot is written
// again before the next barrier.  Built with -fsanitize=thread the run also
// proves the barrier orders those plain accesses (no data race reported).
// A grid with a workgroup that never arrives checks the bounded wait: the
// others give up, raise the abort word and return false.
// Exit status: 0 all checks passed; 1 a check failed.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "../amuse_amd/csrc/g6_gridbar.h"

struct HostOps {
    // ThreadSanitizer does not model stand-alone fences, so the host policy
    // puts the ordering on the atomics themselves (acquire poll; the
    // read-modify-writes are acquire-release on both sides)
    static void release_fence() {}
    static void acquire_fence() {}
    static unsigned fetch_add(unsigned *p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_ACQ_REL); }
    static void store(unsigned *p, unsigned v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
    static void store_release(unsigned *p, unsigned v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
    static unsigned poll(unsigned *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
    static void pause() { std::this_thread::yield(); }
};

// A workgroup that never arrives: every other one gives up (after its own
// spin limit, or at once on seeing the abort word another raised), and
// returns false.  The kernel leaves its loop on the first false.
static bool abort_path()
{
    std::vector<unsigned> bar(g6bar::kWords, 0u);
    const int grid = 40;  // two groups; workgroup 39 never arrives
    std::atomic<int> aborted{0}, passed{0};
    std::vector<std::thread> th;
    for (int b = 0; b < grid - 1; b++)
        th.emplace_back([&, b] {
            if (g6bar::arrive_and_wait<HostOps, 1000u>(bar.data(), b, grid, 0u)) passed++;
            else aborted++;
        });
    for (auto &t : th) t.join();
    const bool ok = passed == 0 && aborted == grid - 1 && bar[g6bar::abort_word()] == 1u &&
                    bar[g6bar::gen_word()] == 0u;
    std::printf("abort path: %s\n", ok ? "ok" : "FAILED");
    return ok;
}

int main()
{
    int failed = abort_path() ? 0 : 1;
    std::vector<unsigned> bar(g6bar::kWords);
    for (int grid : {1, 2, 31, 32, 33, 64, 100, 161}) {
        std::fill(bar.begin(), bar.end(), 0u);
        const int phases = 300;
        std::vector<unsigned> slot(grid, 0u);
        std::atomic<int> bad{0};
        std::vector<std::thread> th;
        for (int b = 0; b < grid; b++)
            th.emplace_back([&, b] {
                unsigned gen = 0;
                for (int k = 1; k <= phases; k++) {
                    slot[b] = unsigned(k);
                    if (!g6bar::arrive_and_wait<HostOps>(bar.data(), b, grid, gen++)) {
                        bad++;
                        return;
                    }
                    for (int o = 0; o < grid; o++)
                        if (slot[o] != unsigned(k)) bad++;
                    if (!g6bar::arrive_and_wait<HostOps>(bar.data(), b, grid, gen++)) {
                        bad++;
                        return;
                    }
                }
            });
        for (auto &t : th) t.join();
        // every counter back at zero, generation = barriers passed, no abort
        bool clean = bar[g6bar::gen_word()] == unsigned(2 * phases) &&
                     bar[g6bar::abort_word()] == 0u && bar[g6bar::top_word()] == 0u;
        for (int g = 0; g < g6bar::kMaxGroups; g++) clean = clean && bar[g * g6bar::kStride] == 0u;
        std::printf("grid %3d: %s\n", grid, bad == 0 && clean ? "ok" : "FAILED");
        failed |= bad != 0 || !clean;
    }
    return failed;
}

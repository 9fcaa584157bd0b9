// Host run of the grid barrier protocol (engine_amd/csrc/gridbar.h): one
// thread per workgroup, many barriers; checks that no thread leaves barrier k
// before every thread has arrived at it, that counters return to zero, and
// that a missing participant ends in an abort rather than a hang.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "gridbar.h"

struct HostBarOps {
    static unsigned fetch_add(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_ACQ_REL); }
    static unsigned load_relaxed(unsigned* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
    static void store_relaxed(unsigned* p, unsigned v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
    static void store_release(unsigned* p, unsigned v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
    static void fence_acquire() { __atomic_thread_fence(__ATOMIC_ACQUIRE); }
    static void pause() { std::this_thread::yield(); }
};

static int run(int nblk, int rounds)
{
    std::vector<unsigned> bar(pdm::GRIDBAR_WORDS, 0u);
    std::vector<std::atomic<int>> arrived(nblk);
    for (auto& a : arrived) a.store(0);
    std::atomic<int> bad(0), aborted(0);
    std::vector<std::thread> th;
    for (int b = 0; b < nblk; ++b)
        th.emplace_back([&, b] {
            unsigned gen = 0;
            for (int k = 0; k < rounds; ++k) {
                arrived[b].store(k + 1, std::memory_order_relaxed);
                if (!pdm::gridbar_arrive_wait<HostBarOps>(bar.data(), b, nblk, gen, 1u << 30)) {
                    aborted.fetch_add(1);
                    return;
                }
                ++gen;
                for (int o = 0; o < nblk; ++o)
                    if (arrived[o].load(std::memory_order_relaxed) < k + 1) bad.fetch_add(1);
            }
        });
    for (auto& t : th) t.join();
    int rc = 0;
    if (bad.load() || aborted.load()) rc = 1;
    for (int g = 0; g < pdm::GRIDBAR_GROUPS + 1; ++g)
        if (bar[g * pdm::GRIDBAR_LINE] != 0u) rc = 1;
    if (bar[(pdm::GRIDBAR_GROUPS + 1) * pdm::GRIDBAR_LINE] != (unsigned)rounds) rc = 1;
    printf("nblk %d rounds %d: early leaves %d, aborts %d, gen %u -> %s\n", nblk, rounds, bad.load(), aborted.load(),
           bar[(pdm::GRIDBAR_GROUPS + 1) * pdm::GRIDBAR_LINE], rc ? "FAIL" : "ok");
    return rc;
}

// nblk - 1 threads of an nblk grid: every participant must abort
static int run_missing(int nblk)
{
    std::vector<unsigned> bar(pdm::GRIDBAR_WORDS, 0u);
    std::atomic<int> aborted(0);
    std::vector<std::thread> th;
    for (int b = 0; b < nblk - 1; ++b)
        th.emplace_back([&, b] {
            if (!pdm::gridbar_arrive_wait<HostBarOps>(bar.data(), b, nblk, 0u, 20000u)) aborted.fetch_add(1);
        });
    for (auto& t : th) t.join();
    const int rc = aborted.load() == nblk - 1 ? 0 : 1;
    printf("missing participant of %d: %d aborts -> %s\n", nblk, aborted.load(), rc ? "FAIL" : "ok");
    return rc;
}

int main()
{
    int rc = 0;
    const int sizes[] = {1, 2, 3, 7, 8, 9, 16, 31, 64, 100};
    for (int n : sizes) rc |= run(n, 300);
    rc |= run_missing(9);
    rc |= run_missing(2);
    return rc;
}

// gridbar.h -- two-level barrier across the workgroups of a co-resident
// (cooperative) grid, with a bounded wait.
//
// Workgroups are split into G = min(8, grid) groups by blockIdx % G.  A
// workgroup's first thread adds to its group's counter; the last arrival of
// a group resets that counter and adds to the top counter; the last group
// resets the top counter and bumps the generation word, which every waiting
// workgroup polls.  Counters sit on separate 256-byte lines, so the G
// groups arrive in parallel and at most max(grid / G, G) arrivals serialise
// on one address.
//
// Ordering: each arrival is an acquire-release read-modify-write at agent
// scope after a workgroup barrier, so the workgroup's writes before the
// barrier are visible to every workgroup that leaves it; the waiting thread
// polls with relaxed loads and issues one acquire fence before releasing its
// workgroup.  Resets happen before the release that lets anyone re-enter.
//
// A wait that exceeds `spin_limit` polls sets *abort and returns false; a
// caller that sees false (or a set *abort) leaves its loop, so a grid whose
// workgroups are not all resident ends with an error instead of hanging.
//
// The protocol is written once over an `Ops` policy: the device policy uses
// HIP's scoped atomics, the host policy (tests) GCC atomics on threads.
#pragma once

namespace pdm {

constexpr int GRIDBAR_LINE = 64;                    // unsigned words per 256-byte line
constexpr int GRIDBAR_GROUPS = 8;
constexpr int GRIDBAR_WORDS = (GRIDBAR_GROUPS + 3) * GRIDBAR_LINE;  // groups, top, generation, abort

// members of group g when nblk workgroups are dealt round-robin to ng groups
#if defined(__HIPCC__)
__host__ __device__
#endif
inline int gridbar_members(int nblk, int ng, int g) { return (nblk - g + ng - 1) / ng; }

template <class Ops>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline bool gridbar_arrive_wait(unsigned* bar, int blk, int nblk, unsigned gen_local, unsigned spin_limit)
{
    const int ng = nblk < GRIDBAR_GROUPS ? nblk : GRIDBAR_GROUPS;
    const int g = blk % ng;
    unsigned* cnt = bar + g * GRIDBAR_LINE;
    unsigned* top = bar + GRIDBAR_GROUPS * GRIDBAR_LINE;
    unsigned* gen = bar + (GRIDBAR_GROUPS + 1) * GRIDBAR_LINE;
    unsigned* abort_flag = bar + (GRIDBAR_GROUPS + 2) * GRIDBAR_LINE;
    const unsigned members = (unsigned)gridbar_members(nblk, ng, g);
    if (Ops::fetch_add(cnt, 1u) == members - 1u) {
        Ops::store_relaxed(cnt, 0u);
        if (Ops::fetch_add(top, 1u) == (unsigned)ng - 1u) {
            Ops::store_relaxed(top, 0u);
            Ops::store_release(gen, gen_local + 1u);
        }
    }
    unsigned spins = 0;
    while (Ops::load_relaxed(gen) == gen_local) {
        if (Ops::load_relaxed(abort_flag) || ++spins > spin_limit) {
            Ops::store_relaxed(abort_flag, 1u);
            Ops::fence_acquire();
            return false;
        }
        Ops::pause();
    }
    Ops::fence_acquire();
    return true;
}

#if defined(__HIPCC__)
struct DeviceBarOps {
    __device__ static unsigned fetch_add(unsigned* p, unsigned v)
    {
        return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static unsigned load_relaxed(unsigned* p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static void store_relaxed(unsigned* p, unsigned v)
    {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static void store_release(unsigned* p, unsigned v)
    {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
    __device__ static void pause() { __builtin_amdgcn_s_sleep(2); }
};

// Whole-workgroup barrier across the grid: every thread of every workgroup
// calls it; returns false (all threads of the workgroup) after an abort.
__device__ __forceinline__ bool grid_barrier(unsigned* bar, unsigned& gen_local, unsigned spin_limit)
{
    __shared__ int ok_shared;
    __syncthreads();
    if (threadIdx.x == 0)
        ok_shared = gridbar_arrive_wait<DeviceBarOps>(bar, blockIdx.x, gridDim.x, gen_local, spin_limit) ? 1 : 0;
    __syncthreads();
    ++gen_local;
    return ok_shared != 0;
}
#endif

}  // namespace pdm

// batch.hip -- many independent unit-square packings solved in one
// persistent launch.
//
// The reference solves one graph per process (core/paradmm.c:736-741; its
// single-block GPU variant d_updateG_XMZUN_single_block,
// core/paradmm-gpu.cu:177-281, is the closest relative).  Here B instances
// of uneven size share one buffer set, each in the pair-tile layout of
// pairtile.h with tile side T (32 or 64, chosen by the autotuner):
//   slots  [s0_b, s0_b + nt_b * 2T^2)   X[3][S], U[3][S], rho[S], alpha[S]
//   circles[c0_b, c0_b + N_b)           Z, Zprev, radius exponents
//   tiles  [t0_b, t0_b + nt_b)          (instance, I, J), FP64 primal and
//                                       quality partials
//   partials P + p0_b: [5][nb_b][N_b]   per-block sums for the Z consensus
//
// k_batch_solve is a cooperative (co-resident) grid that loops over
// iterations with grid-wide barriers between phases:
//   A  every tile of every active instance, strided over the workgroups:
//      the fused tile pass (U update of the previous iteration, prox,
//      partials); at check iterations also the quality partials of Z;
//   B  every circle of every active instance: Z = sum rho M / sum rho over its
//      block partials in fixed order, FP64 dual partial.
// Every `check_every` iterations, after phase A, one wavefront per active
// instance reduces its residuals (primal of this iteration from phase A, dual
// from the previous phase B) and its quality, and retires the instance when
// primal, dual, overlap and container violation are all within tolerance;
// every workgroup then rebuilds the list of active instances in LDS, so the
// tiles and circles of retired instances drop out of both phases.  All
// workgroups take the same decisions from the same global state after a grid
// barrier, so they leave the loop together.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "gridbar.h"
#include "mlp_device.h"
#include "pairtile.h"
#include "contact_device.h"
#include "contact_policy.h"
#include "paradmm_gpu.h"

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

namespace cg = cooperative_groups;

namespace pdm {

constexpr int BATCH_MAX = 1024;  // instances per batch (LDS lists)

struct BInst {
    int N, nb, nt;  // circles, blocks, tiles
    int c0, t0;     // first circle, first tile
    long s0;        // first slot
    long p0;        // first float of the partial region
    double E;       // edges: N(N-1) pair edges + 2N unary
    int uni;        // every edge of the instance has rho_u and alpha_u
    float rho_u, alpha_u;
};

struct BState {
    int active;
    int iterations;
    int converged;
    int pad;
    double primal_rms, dual_rms, max_overlap, max_box, objective;
    long long t_end;
};

struct BArgs {
    int ninst;
    int max_iters, check_every;
    double tol_p, tol_d, tol_f;
    long stride;
    float *X, *U;
    const float *rho, *alpha, *Nb;
    float *Z, *Zprev, *P;
    const float* aexp;
    const BInst* inst;
    const int4* tiles;  // {instance, I, J, local tile}
    double* Rp;         // [tiles]
    double* Rd;         // [circles]
    double* Q;          // [tiles][3]
    BState* st;
    long long* t_start;
    unsigned* bar;        // grid barrier words (zeroed before the launch)
    const int* order;     // [ninst] the order instances are handed out in
    unsigned spin_limit;  // polls before a barrier wait gives up
    int barrier;          // 0: cooperative-groups grid sync, 1: two-level barrier
    // iterations [it0, it_end) of this launch; the first pass reads N from
    // the N buffer (FIRST_NBUF), forms N = Z - U with U current
    // (FIRST_CURRENT, after a rho policy) or applies the pending U update
    // (FIRST_FUSED); finish_only: the pending U update and the check of
    // iteration it0, nothing else
    int it0, it_end, first_mode, finish_only;
    const float* inst_rho;  // [ninst][2]: common rho, alpha of each instance
    int alpha1;             // alpha = 1 on every edge: block tiles keep M = X + U between checks
};

enum { FIRST_NBUF = 0, FIRST_CURRENT = 1, FIRST_FUSED = 2 };

// Grid-wide barrier between phases; false when the two-level barrier gave up
// (every workgroup then leaves the kernel).
template <class A>
__device__ __forceinline__ bool phase_barrier(const A& a, cg::grid_group& grid, unsigned& gen)
{
    if (a.barrier == 0) {
        grid.sync();
        return true;
    }
    return grid_barrier(a.bar, gen, a.spin_limit);
}

// the active instances (at most CAP) with prefix sums, in LDS
template <int CAP>
struct BListsT {
    int n;
    int inst[CAP];
    int tpre[CAP + 1];
    int cpre[CAP + 1];
    int scan[3][256];
};
using BLists = BListsT<BATCH_MAX>;

__device__ __forceinline__ DArgs inst_view(const BArgs& a, const BInst& in, const int i)
{
    DArgs v;
    v.N = in.N;
    v.nb = in.nb;
    v.stride = a.stride;
    v.X = a.X + in.s0;
    v.U = a.U + in.s0;
    v.rho = a.rho + in.s0;
    v.alpha = a.alpha + in.s0;
    v.Nb = a.Nb ? a.Nb + in.s0 : nullptr;
    v.Mb = nullptr;
    v.Z = a.Z + 3L * in.c0;
    v.P = a.P + in.p0;
    v.Rp = nullptr;
    v.aexp = a.aexp + in.c0;
    v.tiles = nullptr;
    v.rho_u = a.inst_rho[2 * i];
    v.alpha_u = a.inst_rho[2 * i + 1];
    v.uni = nullptr;
    return v;
}

// Active instances in the launch's order with prefix sums of their work
// items (items(i): tiles, or contact blocks and tiles) and circle counts;
// every workgroup builds the same lists from st[].active.
template <class Items, class Lists>
__device__ void rebuild_lists_of(int ninst, const int* order, BState* st, const BInst* inst, Items items, Lists& L)
{
    const int tid = threadIdx.x;
    const int per = (ninst + 255) / 256;
    const int lo = min(tid * per, ninst), hi = min(lo + per, ninst);
    int na = 0, nt = 0, nc = 0;
    for (int j = lo; j < hi; ++j)
        if (const int i = order[j]; __hip_atomic_load(&st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ++na;
            nt += items(i);
            nc += inst[i].N;
        }
    L.scan[0][tid] = na;
    L.scan[1][tid] = nt;
    L.scan[2][tid] = nc;
    __syncthreads();
    if (tid < 3) {  // exclusive scans, one thread per quantity
        int acc = 0;
        for (int k = 0; k < 256; ++k) {
            const int v = L.scan[tid][k];
            L.scan[tid][k] = acc;
            acc += v;
        }
    }
    __syncthreads();
    int k = L.scan[0][tid], tp = L.scan[1][tid], cp = L.scan[2][tid];
    for (int j = lo; j < hi; ++j)
        if (const int i = order[j]; __hip_atomic_load(&st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            L.inst[k] = i;
            L.tpre[k] = tp;
            L.cpre[k] = cp;
            ++k;
            tp += items(i);
            cp += inst[i].N;
        }
    if (tid == 255) {
        L.n = k;
        L.tpre[k] = tp;
        L.cpre[k] = cp;
    }
    __syncthreads();
}

__device__ void rebuild_lists(const BArgs& a, BLists& L)
{
    const BInst* inst = a.inst;
    rebuild_lists_of(a.ninst, a.order, a.st, inst, [inst](int i) { return inst[i].nt; }, L);
}

// position of w in a prefix array: pre[k] <= w < pre[k+1]
__device__ __forceinline__ int find_prefix(const int* pre, int n, int w)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Z consensus of circle c of one instance: the block partials summed in
// fixed order; the dual-residual term rho^2 |Z - Zprev|^2 of the circle.
__device__ __forceinline__ void circle_consensus(const BArgs& a, const BInst& in, const int c)
{
    const float* P = a.P + in.p0;
    const long N = in.N;
    float s[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int K = 0; K < in.nb; ++K)
#pragma unroll
        for (int q = 0; q < 5; ++q) s[q] += P[((long)q * in.nb + K) * N + c];
    float* Z = a.Z + 3L * (in.c0 + c);
    float* Zp = a.Zprev + 3L * (in.c0 + c);
    double dz2 = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float z = s[q] / s[3];
        const float zo = Z[q];
        const double d = (double)(z - zo);
        dz2 += d * d;
        Zp[q] = zo;
        Z[q] = z;
    }
    a.Rd[in.c0 + c] = (double)s[4] * dz2;
}

// phase A for tile side 64: one workgroup per tile
template <int MODE, bool QUAL>
__device__ __forceinline__ void phase_tiles_block(const BArgs& a, const BLists& L, TileShared<64>& sh)
{
    const int ntiles = L.tpre[L.n];
    for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
        const int k = find_prefix(L.tpre, L.n, w);
        const int i = L.inst[k];
        const BInst in = a.inst[i];
        const int lt = w - L.tpre[k];
        const int t = in.t0 + lt;
        const int4 td = a.tiles[t];
        const DArgs v = inst_view(a, in, i);
        // in the steady state, instances whose edges share one rho and alpha
        // skip those streams
        if constexpr ((MODE == DM_FUSED || MODE >= DM_M_ENTER) && !QUAL) {
            if (in.uni) tile_pass<64, MODE, QUAL, true>(v, lt, td.y, td.z, sh, a.Rp + t, a.Q + 3L * t);
            else tile_pass<64, MODE, QUAL, false>(v, lt, td.y, td.z, sh, a.Rp + t, a.Q + 3L * t);
        } else {
            tile_pass<64, MODE, QUAL, false>(v, lt, td.y, td.z, sh, a.Rp + t, a.Q + 3L * t);
        }
        __syncthreads();  // the tile's shared scratch is reused by the next tile
    }
}

// phase A for tile side 32: one wavefront per tile, waves independent
template <int MODE, bool QUAL>
__device__ __forceinline__ void phase_tiles_wave(const BArgs& a, const BLists& L, WaveShared& ws)
{
    const int ntiles = L.tpre[L.n];
    const int nw = gridDim.x * 4;
    for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < ntiles; w += nw) {
        const int k = find_prefix(L.tpre, L.n, w);
        const int i = L.inst[k];
        const BInst in = a.inst[i];
        const int lt = w - L.tpre[k];
        const int t = in.t0 + lt;
        const int4 td = a.tiles[t];
        const DArgs v = inst_view(a, in, i);
        if constexpr (MODE == DM_FUSED && !QUAL) {
            if (in.uni) wave_tile_pass<MODE, QUAL, true>(v, lt, td.y, td.z, ws, a.Rp + t, a.Q + 3L * t);
            else wave_tile_pass<MODE, QUAL, false>(v, lt, td.y, td.z, ws, a.Rp + t, a.Q + 3L * t);
        } else {
            wave_tile_pass<MODE, QUAL, false>(v, lt, td.y, td.z, ws, a.Rp + t, a.Q + 3L * t);
        }
    }
}

// The steady-state pass is inlined into the kernel; the passes of the first iteration, of checks and of the final
// iteration are calls, so their registers do not weigh on the steady state
// (a call saves and restores the registers it uses, which only these rare
// passes pay for).
template <int MODE, bool QUAL>
__device__ __noinline__ void phase_tiles_wave_call(const BArgs& a, const BLists& L, WaveShared& ws)
{
    phase_tiles_wave<MODE, QUAL>(a, L, ws);
}
template <int MODE, bool QUAL>
__device__ __noinline__ void phase_tiles_block_call(const BArgs& a, const BLists& L, TileShared<64>& sh)
{
    phase_tiles_block<MODE, QUAL>(a, L, sh);
}

// `a` is the kernel's argument block (inlined use only), `sa` its copy in
// LDS for the calls: a callee given a reference to the argument block would
// make every workgroup copy the block to scratch.
template <int T, int MODE, bool QUAL>
__device__ __forceinline__ void phase_tiles(const BArgs& a, const BArgs& sa, const BLists& L, TileShared<64>& sh,
                                            WaveShared* ws)
{
    constexpr bool STEADY = (MODE == DM_FUSED || MODE == DM_M) && !QUAL;
    if constexpr (T == 32) {
        if constexpr (STEADY) phase_tiles_wave<MODE, QUAL>(a, L, ws[threadIdx.x >> 6]);
        else phase_tiles_wave_call<MODE, QUAL>(sa, L, ws[threadIdx.x >> 6]);
    } else {
        if constexpr (STEADY) phase_tiles_block<MODE, QUAL>(a, L, sh);
        else phase_tiles_block_call<MODE, QUAL>(sa, L, sh);
    }
}

// Z consensus of every active circle, one thread per circle.
__device__ void phase_circles(const BArgs& a, const BLists& L)
{
    const int ncirc = L.cpre[L.n];
    for (int w = blockIdx.x * 256 + threadIdx.x; w < ncirc; w += gridDim.x * 256) {
        const int k = find_prefix(L.cpre, L.n, w);
        circle_consensus(a, a.inst[L.inst[k]], w - L.cpre[k]);
    }
}
__device__ void phase_check(const BArgs& a, const BLists& L, int it, bool final)
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    for (int k = wave; k < L.n; k += gridDim.x * 4) {
        const int i = L.inst[k];
        const BInst in = a.inst[i];
        double p = 0.0, d = 0.0, ov = 0.0, bx = 0.0, sr = 0.0;
        for (int t = lane; t < in.nt; t += 64) {
            const long g = in.t0 + t;
            p += a.Rp[g];
            ov = fmax(ov, a.Q[3 * g]);
            bx = fmax(bx, a.Q[3 * g + 1]);
            sr += a.Q[3 * g + 2];
        }
        for (int c = lane; c < in.N; c += 64) d += a.Rd[in.c0 + c];
        p = wsum64d(p);
        d = wsum64d(d);
        sr = wsum64d(sr);
        ov = wmax64d(ov);
        bx = wmax64d(bx);
        if (lane == 0) {
            const double pr = sqrt(p / (3.0 * in.E)), du = sqrt(d / (3.0 * in.E));
            bool conv = pr <= a.tol_p && du <= a.tol_d;
            if (a.tol_f >= 0.0) conv = conv && ov <= a.tol_f && bx <= a.tol_f;
            // an instance whose state is no longer finite (a NaN or an
            // overflow anywhere reaches its residual sums and its sum of
            // radii) cannot recover: it retires now, not converged, with a
            // NaN primal residual
            const bool lost = !isfinite(pr) || !isfinite(du) || !isfinite(sr);
            BState& s = a.st[i];
            s.iterations = it;
            s.primal_rms = pr;
            s.dual_rms = du;
            s.max_overlap = ov;
            s.max_box = bx;
            s.objective = sr;
            if (lost) s.primal_rms = __builtin_nan("");
            if (conv || final || lost) {
                s.converged = conv && !lost ? 1 : 0;
                s.t_end = wall_clock64();
                s.active = 0;
            }
        }
    }
}

template <int T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_batch_solve(BArgs a)
{
    cg::grid_group grid = cg::this_grid();
    __shared__ BLists L;
    // tile scratch: one block tile (T = 64) or four wave tiles (T = 32)
    __shared__ TileShared<64> sh;
    __shared__ WaveShared ws[4];
    __shared__ BArgs sa;
    if (threadIdx.x == 0) sa = a;
    __syncthreads();
    if (a.it0 == 0 && !a.finish_only && blockIdx.x == 0 && threadIdx.x == 0) *a.t_start = wall_clock64();
    unsigned gen = 0;
    bool mst = false;  // the X array holds M (alpha = 1, block tiles)
    rebuild_lists(a, L);
    if (a.finish_only) {
        // the pending U update of iteration it0 (its primal residual) and the
        // quality of Z_it0, then the check of iteration it0
        phase_tiles<T, DM_FINISH, true>(a, sa, L, sh, ws);
        if (!phase_barrier(a, grid, gen)) return;
        phase_check(a, L, a.it0, a.it0 >= a.max_iters);
        return;
    }
    for (int it = a.it0;; ++it) {
        if (it >= a.it_end) break;  // Z of iteration it is ready, its U update pending
        const bool final = it >= a.max_iters;
        const bool first = it == a.it0;
        const bool check = it > 0 && (final || it % a.check_every == 0) && !(first && a.first_mode != FIRST_FUSED);
        // phase A: the tiles (first pass: N from the buffer, or Z - U)
        if (final) phase_tiles<T, DM_FINISH, true>(a, sa, L, sh, ws);
        else if (first && a.first_mode == FIRST_NBUF) phase_tiles<T, DM_FUSED_NBUF, false>(a, sa, L, sh, ws);
        else if (first && a.first_mode == FIRST_CURRENT)
            phase_tiles<T, DM_FUSED_CURRENT, false>(a, sa, L, sh, ws);
        else if (check) phase_tiles<T, DM_FUSED, true>(a, sa, L, sh, ws);
        else {
            bool done = false;
            if constexpr (T == 64) {
                if (a.alpha1) {
                    // between checks the tiles keep M = X + U (alpha = 1);
                    // the iteration before a check, the final pass or the
                    // launch's end stores X and U again
                    const bool next_split =
                        it + 1 >= a.it_end || it + 1 >= a.max_iters || (it + 1) % a.check_every == 0;
                    if (!mst && !next_split) {
                        phase_tiles<T, DM_M_ENTER, false>(a, sa, L, sh, ws);
                        mst = true;
                    } else if (!mst) {
                        phase_tiles<T, DM_FUSED, false>(a, sa, L, sh, ws);
                    } else if (next_split) {
                        phase_tiles<T, DM_M_EXIT, false>(a, sa, L, sh, ws);
                        mst = false;
                    } else {
                        phase_tiles<T, DM_M, false>(a, sa, L, sh, ws);
                    }
                    done = true;
                }
            }
            if (!done) phase_tiles<T, DM_FUSED, false>(a, sa, L, sh, ws);
        }
        if (!phase_barrier(a, grid, gen)) return;
        if (check) {
            phase_check(a, L, it, final);
            if (!phase_barrier(a, grid, gen)) return;
            rebuild_lists(a, L);
            if (L.n == 0) break;
        }
        // phase B: the Z consensus of the instances still active
        phase_circles(a, L);
        if (!phase_barrier(a, grid, gen)) return;
    }
}

// ---- the contact state on the batch
//
// With alpha = 1 and one rho per instance (the plain fixed-rho batch) every
// instance runs the contact state of the single-instance dense layout
// (contact_device.h): per contact tile of CS x CS circles a list of the
// pairs whose prox is active; every other pair edge carries no state (its M
// is its circle's Z).  The launch sequence of a solve:
//   1. one k_batch_solve iteration (the after-upload pass, per edge);
//   2. k_batch_contact_enter: every instance's lists from its X, U;
//   3. k_batch_contact, one cooperative launch over the remaining
//      iterations: per iteration one phase over the work items of every
//      active instance (its CS-circle blocks' box and radius edges, then its
//      contact tiles), one grid barrier; Z_{k+1} = Z_k + sum / (N + 1) from
//      the integer circle sums is applied by the next phase as it reads the
//      circles, as between the single-instance launches, so a run of
//      iterations needs one barrier each.  At a check (every check_every
//      iterations) the run ends: the per-circle consensus and residuals into
//      the canonical Z arrays, the quality of that Z per contact tile, one
//      wavefront per instance reduces and retires;
//   4. k_batch_contact_leave: every instance's X and U from its lists, at
//      the parity of its own last iteration.
// The circle sums are integers (2^-31 fixed point), so an instance's
// iterations do not depend on the tiling or on what else shares the launch:
// they equal the single-instance contact iterations bit for bit wherever the
// after-upload pass does.

struct BCInst {
    int nbs, nts;  // contact blocks of CS circles, contact tiles
    long l0;       // the instance's first contact tile
};

struct BCArgs {
    int ninst, max_iters, check_every;
    int it0, it_end;  // iterations [it0, it_end) of this launch (Z of it0 canonical)
    int par0;         // the list the last iteration wrote, at it0
    double tol_p, tol_d, tol_f;
    int T;            // dense tile side (the box and radius slots)
    long stride;
    float *X, *U, *Z, *Zprev, *Zpp, *Zb;
    float* W;                 // per instance [2][N]: sums of rho and rho^2 per circle (per-edge rho, after a policy)
    float* Xu;                // per instance [CUNARY][N]: box and radius edges inside the contact state
    const float *rho, *aexp, *inst_rho;
    const BInst* inst;
    const BCInst* cinst;
    float4* L;                // per instance [2][nts][CS*CS][2]
    unsigned* Lmask;          // per instance [2][nts][CS][CS/32]
    int* cnt;                 // per instance [2][nts]
    unsigned long long* Ac;   // per instance [3][CQ][N]
    double *Rpc, *Rd, *cq;    // per circle primal and dual residual; per contact tile max overlap
    unsigned* flags;          // per instance: a circle sum could not take a value
    int* cpar;                // per instance: the list its last iteration wrote
    BState* st;
    long long* t_start;
    unsigned* bar;
    const int* order;
    unsigned spin_limit;
    int barrier;
};

// instance i's contact state as the single-instance kernels see it
template <int CS>
__device__ __forceinline__ CArgs cview(const BCArgs& a, const BInst& in, const BCInst& ci, const int i, const int par)
{
    constexpr long CSS = (long)CS * CS;
    CArgs v;
    v.N = in.N;
    v.nbs = ci.nbs;
    v.nts = ci.nts;
    v.nb = in.nb;
    v.T = a.T;
    v.stride = a.stride;
    v.X = a.X + in.s0;
    v.U = a.U + in.s0;
    v.Z = a.Z + 3L * in.c0;
    v.Zp = a.Zprev + 3L * in.c0;
    v.Zpp = a.Zpp + 3L * in.c0;
    v.aexp = a.aexp + in.c0;
    v.uni = a.inst_rho + 2 * i;
    v.rho = a.rho + in.s0;
    v.W = a.W ? a.W + 2L * in.c0 : nullptr;
    v.par = par;
    v.L = a.L + 2 * ci.l0 * (2 * CSS);
    v.Lmask = a.Lmask + 2 * ci.l0 * (CS * (CS / 32));
    v.cnt = a.cnt + 2 * ci.l0;
    v.Ac = v.Acl = nullptr;
    v.Ap = nullptr;
    v.Zin = v.Zpin = nullptr;
    v.Zo = nullptr;
    v.Rpc = a.Rpc + in.c0;
    v.Rd = a.Rd + in.c0;
    v.flag = a.flags + i;
    v.Xu = a.Xu + (long)CUNARY * in.c0;
    return v;
}

// the sum buffers and Z ring of launch j of a run (as contact_iteration)
__device__ __forceinline__ void cview_run(CArgs& v, const BCArgs& a, const BInst& in, const int j)
{
    const long AQ = (long)CQ * in.N, ZN = 3L * in.N;
    unsigned long long* A = a.Ac + 3L * CQ * in.c0;
    float* Zb = a.Zb + 6L * in.c0;
    v.Ac = A + (j % 3) * AQ;
    v.Ap = j ? A + ((j - 1) % 3) * AQ : nullptr;
    v.Acl = A + ((j + 1) % 3) * AQ;
    v.Zin = j == 0 ? nullptr : j == 1 ? v.Z : Zb + ((j - 1) % 2) * ZN;
    v.Zo = j ? Zb + (j % 2) * ZN : nullptr;
}

// the end of a run whose last launch was j (k_contact_zreduce's arguments)
__device__ __forceinline__ void cview_end(CArgs& f, const BCArgs& a, const BInst& in, const int j)
{
    const long AQ = (long)CQ * in.N, ZN = 3L * in.N;
    unsigned long long* A = a.Ac + 3L * CQ * in.c0;
    float* Zb = a.Zb + 6L * in.c0;
    f.Ac = A + (j % 3) * AQ;
    f.Acl = A + ((j + 2) % 3) * AQ;
    f.Zin = j == 0 ? f.Z : Zb + (j % 2) * ZN;
    f.Zpin = j == 0 ? f.Zp : j == 1 ? f.Z : Zb + ((j - 1) % 2) * ZN;
}

// one iteration's work items, one wavefront each: an instance's nbs blocks
// (box and radius edges) then its nts contact tiles
template <int CS, bool RES, class Lists>
__device__ __forceinline__ void bc_phase_items(const BCArgs& a, const Lists& L, ContactWave<CS, RES ? CQ : 3>& cw,
                                               const int par, const int j)
{
    const int nitems = L.tpre[L.n];
    const int nw = gridDim.x * 4;
    // the wavefront's index is uniform: the item's instance and pointers live in scalar registers
    for (int w = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); w < nitems; w += nw) {
        const int k = find_prefix(L.tpre, L.n, w);
        const int i = L.inst[k];
        const BInst in = a.inst[i];
        const BCInst ci = a.cinst[i];
        const int li = w - L.tpre[k];
        CArgs v = cview<CS>(a, in, ci, i, par);
        cview_run(v, a, in, j);
        if (li < ci.nbs) contact_unary<CS, RES, false>(v, li);
        else contact_tile_body<CS, RES, false>(v, li - ci.nbs, cw);
    }
}

// the end of a run: every active circle's consensus into the canonical Z
// arrays (RES: with its residuals), one thread per circle
template <int CS, bool RES, class Lists>
__device__ void bc_phase_zreduce(const BCArgs& a, const Lists& L, const int j)
{
    const int nc = L.cpre[L.n];
    for (int w = blockIdx.x * 256 + threadIdx.x; w < nc; w += gridDim.x * 256) {
        const int k = find_prefix(L.cpre, L.n, w);
        const int i = L.inst[k];
        const BInst in = a.inst[i];
        CArgs f = cview<CS>(a, in, a.cinst[i], i, 0);
        cview_end(f, a, in, j);
        contact_zreduce_circle<RES, false>(f, w - L.cpre[k]);
    }
}

// the largest pair overlap r_i + r_j - |c_i - c_j| of Z over each contact
// tile's pairs (FP64, as the per-edge batch's quality partials), one
// wavefront per tile
template <int CS, class Lists>
__device__ void bc_phase_quality(const BCArgs& a, const Lists& L)
{
    const int lane = threadIdx.x & 63;
    const int nitems = L.tpre[L.n];
    const int nw = gridDim.x * 4;
    // the wavefront's index is uniform: the item's instance and pointers live in scalar registers
    for (int w = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); w < nitems; w += nw) {
        const int k = find_prefix(L.tpre, L.n, w);
        const int i = L.inst[k];
        const BCInst ci = a.cinst[i];
        const int li = w - L.tpre[k];
        if (li < ci.nbs) continue;
        const BInst in = a.inst[i];
        const int t = li - ci.nbs;
        int P, Qb;
        tri_tile(t, ci.nbs, P, Qb);
        const float* Z = a.Z + 3L * in.c0;
        double mo = 0.0;
        for (int idx = lane; idx < CS * CS; idx += 64) {
            const int il = idx / CS, jl = idx % CS;
            const int ii = P * CS + il, jj = Qb * CS + jl;
            if (ii >= in.N || jj >= in.N || (P == Qb && il >= jl)) continue;
            const double dx = (double)Z[3 * ii] - (double)Z[3 * jj], dy = (double)Z[3 * ii + 1] - (double)Z[3 * jj + 1];
            const double ov = (double)Z[3 * ii + 2] + (double)Z[3 * jj + 2] - sqrt(dx * dx + dy * dy);
            mo = ov > mo ? ov : mo;
        }
        mo = wmax64d(mo);
        if (lane == 0) a.cq[ci.l0 + t] = mo;
    }
}

// one wavefront per active instance: the residuals and quality of the
// run's last iteration (`done` iterations in all), retirement as phase_check
// one wavefront: instance i's check
__device__ __forceinline__ void bc_check_instance(const BCArgs& a, const int i, const int done, const bool final,
                                                  const int par)
{
    const int lane = threadIdx.x & 63;
    {
        const BInst in = a.inst[i];
        const BCInst ci = a.cinst[i];
        double p = 0.0, d = 0.0, ov = 0.0, bx = 0.0, sr = 0.0;
        for (int t = lane; t < ci.nts; t += 64) ov = fmax(ov, a.cq[ci.l0 + t]);
        for (int c = lane; c < in.N; c += 64) {
            const long g = in.c0 + c;
            p += a.Rpc[g];
            d += a.Rd[g];
            const double cx = a.Z[3 * g], cy = a.Z[3 * g + 1], r = a.Z[3 * g + 2];
            const double v[5] = {r - cx, cx + r - 1.0, r - cy, cy + r - 1.0, -r};
#pragma unroll
            for (int q = 0; q < 5; ++q) bx = v[q] > bx ? v[q] : bx;
            sr += r;
        }
        p = wsum64d(p);
        d = wsum64d(d);
        sr = wsum64d(sr);
        ov = wmax64d(ov);
        bx = wmax64d(bx);
        if (lane == 0) {
            const double pr = sqrt(p / (3.0 * in.E)), du = sqrt(d / (3.0 * in.E));
            bool conv = pr <= a.tol_p && du <= a.tol_d;
            if (a.tol_f >= 0.0) conv = conv && ov <= a.tol_f && bx <= a.tol_f;
            // a circle sum that could not take a value (a non-finite state or
            // an overflow) or a non-finite residual: retire now, not converged
            const bool lost = a.flags[i] != 0u || !isfinite(pr) || !isfinite(du) || !isfinite(sr);
            BState& s = a.st[i];
            s.iterations = done;
            s.primal_rms = lost ? __builtin_nan("") : pr;
            s.dual_rms = du;
            s.max_overlap = ov;
            s.max_box = bx;
            s.objective = sr;
            if (conv || final || lost) {
                s.converged = conv && !lost ? 1 : 0;
                s.t_end = wall_clock64();
                a.cpar[i] = par;
                s.active = 0;
            }
        }
    }
}

template <class Lists>
__device__ void bc_phase_check(const BCArgs& a, const Lists& L, const int done, const bool final, const int par)
{
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int k = wave; k < L.n; k += gridDim.x * 4) bc_check_instance(a, L.inst[k], done, final, par);
}

// CAP: the most instances the launch lists (256 or BATCH_MAX; fewer lists
// leave LDS for more resident workgroups)
template <int CS, int CAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CS == 32 ? 4 : 2))) void k_batch_contact(BCArgs a)
{
    cg::grid_group grid = cg::this_grid();
    __shared__ BListsT<CAP> L;
    // one contact scratch per wavefront (the iterations without residual
    // sums use the smaller layout in the same bytes)
    __shared__ __align__(16) unsigned char cw[4][sizeof(ContactWave<CS, CQ>)];
    unsigned gen = 0;
    const BInst* inst = a.inst;
    const BCInst* cinst = a.cinst;
    auto items = [cinst](int i) { return cinst[i].nbs + cinst[i].nts; };
    rebuild_lists_of(a.ninst, a.order, a.st, inst, items, L);
    unsigned char* mine = cw[threadIdx.x >> 6];
    int par = a.par0, j = 0;
    for (int it = a.it0; it < a.it_end && L.n > 0; ++it) {
        const int done = it + 1;
        const bool final = done >= a.max_iters;
        const bool check = final || done % a.check_every == 0;
        const bool end = check || done >= a.it_end;
        if (check) bc_phase_items<CS, true>(a, L, *reinterpret_cast<ContactWave<CS, CQ>*>(mine), par, j);
        else bc_phase_items<CS, false>(a, L, *reinterpret_cast<ContactWave<CS, 3>*>(mine), par, j);
        if (!phase_barrier(a, grid, gen)) return;
        par ^= 1;
        if (!end) {
            ++j;
            continue;
        }
        if (check) bc_phase_zreduce<CS, true>(a, L, j);
        else bc_phase_zreduce<CS, false>(a, L, j);
        if (!phase_barrier(a, grid, gen)) return;
        j = 0;
        if (!check) continue;
        bc_phase_quality<CS>(a, L);
        if (!phase_barrier(a, grid, gen)) return;
        bc_phase_check(a, L, done, final, par);
        if (!phase_barrier(a, grid, gen)) return;
        rebuild_lists_of(a.ninst, a.order, a.st, inst, items, L);
    }
    // instances still active at the launch's end (it_end before their last
    // check) record the parity here
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < L.n; k += 256) a.cpar[L.inst[k]] = par;
}

// one workgroup per contact tile of every instance (ctile: instance, tile)
template <int CS>
__global__ __launch_bounds__(256) void k_batch_contact_enter(BCArgs a, const int2* ctile)
{
    __shared__ ContactEnterShared<CS> sh;
    const int2 e = ctile[blockIdx.x];
    const CArgs v = cview<CS>(a, a.inst[e.x], a.cinst[e.x], e.x, 0);
    contact_enter_tile<CS>(v, e.y, sh);
}

template <int CS>
__global__ __launch_bounds__(256) void k_batch_contact_leave(BCArgs a, const int2* ctile)
{
    __shared__ ContactLeaveShared<CS> sh;
    const int2 e = ctile[blockIdx.x];
    const CArgs v = cview<CS>(a, a.inst[e.x], a.cinst[e.x], e.x, a.cpar[e.x]);
    contact_leave_tile<CS>(v, e.y, sh);
}

// ---- learned rho on the batch: one rho per instance, 10^(mean of the
// network's log10 rho over the instance's edges), between launches

struct PolicySlot {
    bool valid;
    int c, type;  // instance-local circle, PDM_EDGE_* type
};

// slot q of tile (I, J) of an instance of N circles with tile side T
template <int T>
__device__ __forceinline__ PolicySlot policy_slot(const int q, const int I, const int J, const int N)
{
    constexpr int TT = T * T;
    const int side = q / TT, rem = q % TT, il = rem / T, jl = rem % T;
    const int i = I * T + il, j = J * T + jl;
    PolicySlot s{false, 0, PDM_EDGE_PAIR};
    if (I != J) {
        s.valid = i < N && j < N;
        s.c = side ? j : i;
    } else if (il < jl) {
        s.valid = j < N;
        s.c = side ? j : i;
    } else if (il == jl) {
        s.valid = i < N;
        s.c = i;
        s.type = side ? PDM_EDGE_RADIUS : PDM_EDGE_CONTAINER;
    }
    return s;
}

// ---- the contact batch with one launch per iteration: the same work items
// and phases as k_batch_contact, each phase a kernel of its own (the
// launch boundary orders them), so the hardware keeps as many item
// wavefronts resident as fit instead of a fixed persistent grid.  Items of
// retired instances return at once.

// items: (instance, local item), every instance's blocks then its tiles
// RE: rho per edge (after a policy application)
template <int CS, bool RES, bool RE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_contact_step(
    BCArgs a, const int2* items, int par, int j)
{
    __shared__ ContactWave<CS, RES ? CQ : 3> sh;
    const int2 e = items[blockIdx.x];
    if (!__hip_atomic_load(&a.st[e.x].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const BInst in = a.inst[e.x];
    const BCInst ci = a.cinst[e.x];
    CArgs v = cview<CS>(a, in, ci, e.x, par);
    cview_run(v, a, in, j);
    if (e.y < ci.nbs) contact_unary<CS, RES, RE>(v, e.y);
    else contact_tile_body<CS, RES, RE>(v, e.y - ci.nbs, sh);
}

// the end of a run: every circle of an active instance (circ_inst: its instance)
template <int CS, bool RES, bool RE>
__global__ __launch_bounds__(256) void k_batch_contact_zreduce(BCArgs a, const int* circ_inst, int ncirc, int j)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ncirc) return;
    const int i = circ_inst[g];
    if (!__hip_atomic_load(&a.st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const BInst in = a.inst[i];
    CArgs f = cview<CS>(a, in, a.cinst[i], i, 0);
    cview_end(f, a, in, j);
    contact_zreduce_circle<RES, RE>(f, g - in.c0);
}

// ---- the learned rho policy inside the batch's contact state: each active
// instance takes the application of contact_policy.h (modes 2 and 3, as
// paradmm_gpu_solve applies it to one graph in the contact state), on its
// own regions of the policy buffers

struct BCPol {
    const float* params;  // the network
    float frac, raise_log;
    int exceptions;       // rho already per edge (an earlier application)
    int average;
    float* base;          // [circles]
    float* yc;            // per instance [3][N]
    int* rows;            // [circles]
    long long* wacc;      // per instance [2][N]
    long long* ysum;      // [instances]
    double* hist;         // [instances][6]
    float* gl;            // [instances][2]
};

template <int CS>
__device__ __forceinline__ CPolArgs cpol_view(const BCArgs& a, const BCPol& pp, const BInst& in, const int i,
                                              const int par)
{
    CPolArgs p;
    p.c = cview<CS>(a, in, a.cinst[i], i, par);
    p.params = pp.params;
    p.frac = pp.frac;
    p.raise_log = pp.raise_log;
    p.exceptions = pp.exceptions;
    p.average = pp.average;
    p.base = pp.base + in.c0;
    p.yc = pp.yc + 3L * in.c0;
    p.rows = pp.rows + in.c0;
    p.wacc = pp.wacc + 2L * in.c0;
    p.ysum = pp.ysum + i;
    p.hist = pp.hist + 6L * i;
    p.gl = pp.gl + 2L * i;
    return p;
}

// one workgroup per contact tile of an active instance
template <int CS, bool APPLY>
__global__ __launch_bounds__(256) void k_batch_cpol_tile(BCArgs a, BCPol pp, const int2* ctile, int par)
{
    __shared__ CPolShared<CS> sh;
    const int2 e = ctile[blockIdx.x];
    if (!__hip_atomic_load(&a.st[e.x].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const CPolArgs p = cpol_view<CS>(a, pp, a.inst[e.x], e.x, par);
    cpol_tile_body<CS, APPLY>(p, e.y, sh);
}

// one workgroup per active instance: its rho
template <int CS>
__global__ __launch_bounds__(256) void k_batch_cpol_final(BCArgs a, BCPol pp, int par)
{
    __shared__ long long sh[256];
    const int i = blockIdx.x;
    if (!__hip_atomic_load(&a.st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const BInst in = a.inst[i];
    const CPolArgs p = cpol_view<CS>(a, pp, in, i, par);
    cpol_final_body(p, (long)in.E, sh);
}

// one thread per circle of an active instance
template <int CS>
__global__ __launch_bounds__(256) void k_batch_cpol_zprev(BCArgs a, BCPol pp, const int* circ_inst, int ncirc, int par)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ncirc) return;
    const int i = circ_inst[g];
    if (!__hip_atomic_load(&a.st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const BInst in = a.inst[i];
    const CPolArgs p = cpol_view<CS>(a, pp, in, i, par);
    cpol_zprev_circle(p, g - in.c0);
}

// every circle's base rho: its instance's rho (one rho per instance at upload)
__global__ __launch_bounds__(256) void k_batch_cpol_base(const float* inst_rho, const int* circ_inst, int ncirc,
                                                         float* base)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g < ncirc) base[g] = inst_rho[2 * circ_inst[g]];
}

// quality partials: one wavefront per contact tile (ctile: instance, tile)
template <int CS>
__global__ __launch_bounds__(64) void k_batch_contact_quality(BCArgs a, const int2* ctile)
{
    const int2 e = ctile[blockIdx.x];
    if (!__hip_atomic_load(&a.st[e.x].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int lane = threadIdx.x;
    const BInst in = a.inst[e.x];
    const BCInst ci = a.cinst[e.x];
    int P, Qb;
    tri_tile(e.y, ci.nbs, P, Qb);
    const float* Z = a.Z + 3L * in.c0;
    double mo = 0.0;
    for (int idx = lane; idx < CS * CS; idx += 64) {
        const int il = idx / CS, jl = idx % CS;
        const int ii = P * CS + il, jj = Qb * CS + jl;
        if (ii >= in.N || jj >= in.N || (P == Qb && il >= jl)) continue;
        const double dx = (double)Z[3 * ii] - (double)Z[3 * jj], dy = (double)Z[3 * ii + 1] - (double)Z[3 * jj + 1];
        const double ov = (double)Z[3 * ii + 2] + (double)Z[3 * jj + 2] - sqrt(dx * dx + dy * dy);
        mo = ov > mo ? ov : mo;
    }
    mo = wmax64d(mo);
    if (lane == 0) a.cq[ci.l0 + e.y] = mo;
}

// the check: one wavefront per active instance
__global__ __launch_bounds__(64) void k_batch_contact_check(BCArgs a, int done, int final, int par)
{
    const int i = blockIdx.x;
    if (!__hip_atomic_load(&a.st[i].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    bc_check_instance(a, i, done, final != 0, par);
}

// The pending U update of every edge of every instance, U = (X + U) - Z
// (alpha = 1, u_update's rounding): after the contact state's exit the
// edges hold X_j and U_{j-1}, as after every steady iteration, and a solve
// hands back U_j, as the per-edge batch's last pass and the single-instance
// host sync do.  One workgroup per dense tile.
template <int T>
__global__ __launch_bounds__(256) void k_batch_finish_u(const int4* tiles, const BInst* inst, long stride,
                                                        const float* X, float* U, const float* Z)
{
    constexpr int NS = 2 * T * T;
    const int4 td = tiles[blockIdx.x];
    const BInst in = inst[td.x];
    const long base = in.s0 + (long)td.w * NS;
    for (int q = threadIdx.x; q < NS; q += 256) {
        const PolicySlot ps = policy_slot<T>(q, td.y, td.z, in.N);
        if (!ps.valid) continue;
        const long p = base + q;
        const float* z = Z + 3L * (in.c0 + ps.c);
#pragma unroll
        for (int k = 0; k < 3; ++k) U[k * stride + p] = u_update(U[k * stride + p], X[k * stride + p], z[k], 1.0f, 0.0f);
    }
}

struct BPolicyArgs {
    const int* ptiles;   // active tiles (global tile ids), instance order
    const BInst* inst;
    const int4* tiles;
    long stride;
    const float* X;
    float* U;
    float* rho;
    const float* Z;
    const float* Zprev;
    float* inst_rho;     // [ninst][2]
    float frac;
    double* part;        // [n ptiles] sum of log10 rho
    int* cnt;            // [n ptiles]
    const int2* iranges; // [n active] {instance, first ptile index}, then the end
    int nact;
    float* scale;        // [ninst] rho_old / rho_new
    float* yout;         // [slots] network output per slot (raises), or null
    int* raised;         // [ninst] edges given their own rho (raises)
    double* hist;        // [ninst][2] running {sum of means, count}, or null
};

// an edge keeps its own network rho when that exceeds its instance's rho by
// more than this factor (log10; 10^0.5 = 3.2); the same rule as the
// single-instance policy mode 3
constexpr float BATCH_RAISE_LOG10 = 0.5f;

// pass 1: the network's log10 rho of every edge of one active tile, summed
template <int T>
__global__ __launch_bounds__(256) void k_batch_policy_eval(BPolicyArgs a, const float* __restrict__ params)
{
    constexpr int NS = 2 * T * T;
    __shared__ MlpShared w;
    __shared__ TileScratch tsw[4];
    __shared__ float yv[4][64];
    __shared__ double ssum[4];
    __shared__ int scnt[4];
    stage_weights(params, w);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tg = a.ptiles[blockIdx.x];
    const int4 td = a.tiles[tg];
    const BInst in = a.inst[td.x];
    const long base = in.s0 + (long)td.w * NS;
    double acc = 0.0;
    int num = 0;
    for (int q0 = wv * 64; q0 < NS; q0 += 256) {
        const int q = q0 + l;
        const PolicySlot ps = policy_slot<T>(q, td.y, td.z, in.N);
        float r = 0.0f, sres = 0.0f, rho = 1.0f, f[MLP_NIN];
        if (ps.valid) {
            const long p = base + q;
            rho = a.rho[p];
            const int v = in.c0 + ps.c;
            float r2 = 0.0f, dz2 = 0.0f;
            for (int k = 0; k < 3; ++k) {
                const float z = a.Z[3L * v + k];
                const float d = a.X[k * a.stride + p] - z;
                const float dz = z - a.Zprev[3L * v + k];
                r2 += d * d;
                dz2 += dz * dz;
            }
            r = sqrtf(r2);
            sres = rho * sqrtf(dz2);
        }
        edge_features(r, sres, rho, ps.type, a.frac, f);
        const float y = policy_wave_eval(w, tsw[wv], yv[wv], f, l);
        if (ps.valid && a.yout) a.yout[base + q] = y;
        acc += ps.valid ? (double)y : 0.0;
        num += ps.valid ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o, 64);
        num += __shfl_xor(num, o, 64);
    }
    if (l == 0) {
        ssum[wv] = acc;
        scnt[wv] = num;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = (ssum[0] + ssum[1]) + (ssum[2] + ssum[3]);
        a.cnt[blockIdx.x] = scnt[0] + scnt[1] + scnt[2] + scnt[3];
    }
}

// pass 2: one thread per active instance, its tiles' sums in fixed order
__global__ __launch_bounds__(256) void k_batch_policy_reduce(BPolicyArgs a)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.nact) return;
    const int i = a.iranges[k].x;
    double sum = 0.0;
    long n = 0;
    for (int t = a.iranges[k].y; t < a.iranges[k + 1].y; ++t) {
        sum += a.part[t];
        n += a.cnt[t];
    }
    double m = sum / (double)(n ? n : 1);
    if (a.hist) {  // the average of this and the earlier estimates
        a.hist[2 * i] += m;
        a.hist[2 * i + 1] += 1.0;
        m = a.hist[2 * i] / a.hist[2 * i + 1];
    }
    const float rn = exp10f((float)m);
    a.scale[i] = a.inst_rho[2 * i] / rn;
    a.inst_rho[2 * i] = rn;
}

// pass 3: U_e <- U_e rho_old / rho_new, rho_e <- rho_new on every edge
template <int T>
__global__ __launch_bounds__(256) void k_batch_policy_apply(BPolicyArgs a)
{
    constexpr int NS = 2 * T * T;
    const int tg = a.ptiles[blockIdx.x];
    const int4 td = a.tiles[tg];
    const BInst in = a.inst[td.x];
    const long base = in.s0 + (long)td.w * NS;
    const float sc = a.scale[td.x], rn = a.inst_rho[2 * td.x];
    if (!a.yout) {
        for (int q = threadIdx.x; q < NS; q += 256) {
            if (!policy_slot<T>(q, td.y, td.z, in.N).valid) continue;
            const long p = base + q;
            for (int k = 0; k < 3; ++k) a.U[k * a.stride + p] *= sc;
            a.rho[p] = rn;
        }
        return;
    }
    // raises: an edge the network singles out keeps its own rho
    const float lg = log10f(rn);
    for (int q = threadIdx.x; q < NS; q += 256) {
        const bool valid = policy_slot<T>(q, td.y, td.z, in.N).valid;
        const long p = base + q;
        bool up = false;
        if (valid) {
            const float y = a.yout[p];
            up = y - lg > BATCH_RAISE_LOG10;
            const float r = up ? exp10f(y) : rn;
            const float f = a.rho[p] / r;
            for (int k = 0; k < 3; ++k) a.U[k * a.stride + p] *= f;
            a.rho[p] = r;
        }
        const unsigned long long bal = __ballot(up);
        if ((threadIdx.x & 63) == 0 && bal) atomicAdd(a.raised + td.x, (int)__popcll(bal));
    }
}

}  // namespace pdm

using namespace pdm;

struct paradmm_batch {
    std::vector<graph*> graphs;
    int ninst = 0;
    int T = 0;              // configured tile side (0: not configured)
    int blocks_per_cu = 0;  // persistent grid
    int grid = 0;
    int cus = 0;
    double clock_khz = 0.0;
    std::vector<BInst> inst;
    std::vector<std::vector<int>> pos;  // per graph: edge -> instance slot
    std::vector<float> aexp;
    long slots = 0;
    int ntiles = 0, ncirc = 0;
    long pfloats = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> allocs;
    // device
    float *X = nullptr, *U = nullptr, *rho = nullptr, *alpha = nullptr, *Nb = nullptr;
    float *Z = nullptr, *Zprev = nullptr, *P = nullptr, *aexp_d = nullptr;
    BInst* inst_d = nullptr;
    int4* tiles_d = nullptr;
    double *Rp = nullptr, *Rd = nullptr, *Q = nullptr;
    BState* st = nullptr;
    long long* t_start = nullptr;
    unsigned* bar = nullptr;  // grid barrier words
    size_t bar_words = 0;
    int* order_d = nullptr;   // instances by descending tile count
    int barrier = 1;  // two-level barrier by default
    bool solved = false;
    // learned rho: one rho per instance between launches
    paradmm_rho_mlp* policy = nullptr;
    float* inst_rho = nullptr;  // [ninst][2] device
    int* ptiles = nullptr;      // [ntiles] device
    int2* iranges = nullptr;    // [ninst + 1] device
    double* ppart = nullptr;    // [ntiles]
    int* pcnt = nullptr;        // [ntiles]
    float* pscale = nullptr;    // [ninst]
    int policy_mode = 2;        // 2: one rho per instance; 3: plus the edges the network raises
    int policy_average = 0;     // running mean of each instance's estimates
    float* yout = nullptr;      // [slots] (mode 3)
    int* raised = nullptr;      // [ninst] (mode 3)
    double* phist = nullptr;    // [ninst][2] (average)
    std::vector<int> uni0;      // per instance: one rho and alpha on every edge at upload
    int alpha1 = 0;             // alpha = 1 on every edge of every instance
    // the contact state (alpha = 1, one rho per instance, no policy)
    int ccs = 32;               // contact tile side
    int cbpc = -1, cgrid = 0;   // its persistent grid (0: a launch per iteration, the default; -1: not configured)
    int cbarrier = 1;
    bool contact_off = false;   // paradmm_batch_configure_contact(b, 0, ...)
    int ccs_alloc = 0;          // side the contact buffers are laid out for
    long cts = 0;               // contact tiles of all instances
    std::vector<BCInst> cinst;
    std::vector<void*> callocs;
    BCInst* cinst_d = nullptr;
    int2* ctile_d = nullptr;    // [cts] (instance, tile)
    int2* citems_d = nullptr;   // every instance's blocks, then every instance's tiles (one launch per iteration)
    int nitems = 0;
    int* circ_inst_d = nullptr; // [ncirc] instance of each circle
    // the policy inside the contact state
    float *cW = nullptr, *cbase = nullptr, *cyc = nullptr, *cgl = nullptr, *cXu = nullptr;
    int* crows = nullptr;
    long long *cwacc = nullptr, *cysum = nullptr;
    double* chist = nullptr;
    float4* cL = nullptr;
    unsigned* cLmask = nullptr;
    int* ccnt = nullptr;
    unsigned long long* cAc = nullptr;
    float *cZb = nullptr, *cZpp = nullptr;
    double *cRpc = nullptr, *ccq = nullptr;
    unsigned* cflags = nullptr;
    int* cpar = nullptr;
    float last_ms[4] = {0, 0, 0, 0};  // last contact solve: after-upload pass, entry, iterations, exit
};

static void batch_contact_free(paradmm_batch* b)
{
    for (void* p : b->callocs) (void)hipFree(p);
    b->callocs.clear();
    b->ccs_alloc = 0;
}

static void batch_free_dev(paradmm_batch* b)
{
    batch_contact_free(b);
    for (void* p : b->allocs) (void)hipFree(p);
    b->allocs.clear();
    b->yout = nullptr;  // allocated on first use
    b->raised = nullptr;
    b->phist = nullptr;
}

static int bmalloc(paradmm_batch* b, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        pdm_set_error(PARADMM_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return PARADMM_ERR_NOMEM;
    }
    b->allocs.push_back(*p);
    return PARADMM_OK;
}

static int grid_for(paradmm_batch* b, int T, int bpc, int* grid)
{
    int per_cu = 0;
    hipError_t e = T == 32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_solve<32>, 256, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_solve<64>, 256, 0);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor", __FILE__, __LINE__);
    if (per_cu < 1) {
        pdm_set_error(PARADMM_ERR_HIP, "batch kernel does not fit a compute unit");
        return PARADMM_ERR_HIP;
    }
    if (bpc < 1 || bpc > per_cu) {
        pdm_set_error(PARADMM_ERR_ARG, "blocks_per_cu %d outside 1..%d for tile %d", bpc, per_cu, T);
        return PARADMM_ERR_ARG;
    }
    *grid = bpc * b->cus;
    return PARADMM_OK;
}

// Builds the layout for tile side T and allocates the device buffers.
static int batch_layout(paradmm_batch* b, int T)
{
    batch_free_dev(b);
    b->inst.assign(b->ninst, BInst{});
    b->pos.assign(b->ninst, std::vector<int>());
    b->aexp.clear();
    long s0 = 0, p0 = 0;
    int t0 = 0, c0 = 0;
    std::vector<int4> tiles;
    for (int i = 0; i < b->ninst; ++i) {
        graph* G = b->graphs[i];
        std::vector<int> fk;
        RC(pdm_factor_kinds(G, fk, false));
        std::vector<float> ae;
        int nb = 0, nt = 0;
        const int rc = dense_structure(G, fk, T, b->pos[i], ae, nb, nt);
        if (rc < 0) {
            pdm_set_error(PARADMM_ERR_SHAPE, "batch instance %d is not an all-pairs unit-square packing (dim 3)", i);
            return PARADMM_ERR_SHAPE;
        }
        RC(rc);
        BInst& in = b->inst[i];
        in.N = G->num_vars;
        in.nb = nb;
        in.nt = nt;
        in.c0 = c0;
        in.t0 = t0;
        in.s0 = s0;
        in.p0 = p0;
        in.E = (double)G->num_edges;
        for (int I = 0, t = 0; I < nb; ++I)
            for (int J = I; J < nb; ++J, ++t) tiles.push_back(make_int4(i, I, J, t));
        b->aexp.insert(b->aexp.end(), ae.begin(), ae.end());
        s0 += (long)nt * 2 * T * T;
        p0 += 5L * in.N * nb;
        t0 += nt;
        c0 += in.N;
    }
    b->slots = s0;
    b->ntiles = t0;
    b->ncirc = c0;
    b->pfloats = p0;
    const size_t S = (size_t)s0;
    RC(bmalloc(b, (void**)&b->X, sizeof(float) * 3 * S));
    RC(bmalloc(b, (void**)&b->U, sizeof(float) * 3 * S));
    RC(bmalloc(b, (void**)&b->Nb, sizeof(float) * 3 * S));
    RC(bmalloc(b, (void**)&b->rho, sizeof(float) * S));
    RC(bmalloc(b, (void**)&b->alpha, sizeof(float) * S));
    RC(bmalloc(b, (void**)&b->Z, sizeof(float) * 3 * (size_t)c0));
    RC(bmalloc(b, (void**)&b->Zprev, sizeof(float) * 3 * (size_t)c0));
    RC(bmalloc(b, (void**)&b->aexp_d, sizeof(float) * (size_t)c0));
    RC(bmalloc(b, (void**)&b->P, sizeof(float) * (size_t)p0));
    RC(bmalloc(b, (void**)&b->inst_d, sizeof(BInst) * (size_t)b->ninst));
    RC(bmalloc(b, (void**)&b->tiles_d, sizeof(int4) * (size_t)t0));
    RC(bmalloc(b, (void**)&b->Rp, sizeof(double) * (size_t)t0));
    RC(bmalloc(b, (void**)&b->Q, sizeof(double) * 3 * (size_t)t0));
    RC(bmalloc(b, (void**)&b->Rd, sizeof(double) * (size_t)c0));
    RC(bmalloc(b, (void**)&b->st, sizeof(BState) * (size_t)b->ninst));
    RC(bmalloc(b, (void**)&b->t_start, sizeof(long long)));
    b->bar_words = (size_t)GRIDBAR_WORDS;
    RC(bmalloc(b, (void**)&b->bar, sizeof(unsigned) * b->bar_words));
    RC(bmalloc(b, (void**)&b->order_d, sizeof(int) * (size_t)b->ninst));
    RC(bmalloc(b, (void**)&b->inst_rho, sizeof(float) * 2 * (size_t)b->ninst));
    RC(bmalloc(b, (void**)&b->ptiles, sizeof(int) * (size_t)t0));
    RC(bmalloc(b, (void**)&b->iranges, sizeof(int2) * ((size_t)b->ninst + 1)));
    RC(bmalloc(b, (void**)&b->ppart, sizeof(double) * (size_t)t0));
    RC(bmalloc(b, (void**)&b->pcnt, sizeof(int) * (size_t)t0));
    RC(bmalloc(b, (void**)&b->pscale, sizeof(float) * (size_t)b->ninst));
    CK(hipMemcpy(b->inst_d, b->inst.data(), sizeof(BInst) * (size_t)b->ninst, hipMemcpyHostToDevice));
    CK(hipMemcpy(b->tiles_d, tiles.data(), sizeof(int4) * (size_t)t0, hipMemcpyHostToDevice));
    // larger instances first, so the tiles handed out last are small ones
    std::vector<int> order(b->ninst);
    for (int i = 0; i < b->ninst; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->inst[x].nt > b->inst[y].nt; });
    CK(hipMemcpy(b->order_d, order.data(), sizeof(int) * (size_t)b->ninst, hipMemcpyHostToDevice));
    CK(hipMemcpy(b->aexp_d, b->aexp.data(), sizeof(float) * (size_t)c0, hipMemcpyHostToDevice));
    // partials of retired or padded circles are never read, but keep them finite
    CK(hipMemset(b->P, 0, sizeof(float) * (size_t)p0));
    CK(hipMemset(b->Rp, 0, sizeof(double) * (size_t)t0));
    CK(hipMemset(b->Q, 0, sizeof(double) * 3 * (size_t)t0));
    CK(hipMemset(b->Rd, 0, sizeof(double) * (size_t)c0));
    b->T = T;
    b->solved = false;
    return PARADMM_OK;
}

// Every instance active again, results cleared.
static int batch_reset_flags(paradmm_batch* b)
{
    std::vector<BState> st(b->ninst);
    for (BState& s : st) {
        memset(&s, 0, sizeof s);
        s.active = 1;
        s.max_overlap = s.max_box = s.objective = -1.0;
    }
    CK(hipMemcpyAsync(b->st, st.data(), sizeof(BState) * st.size(), hipMemcpyHostToDevice, b->stream));
    CK(hipStreamSynchronize(b->stream));
    return PARADMM_OK;
}

// Host state of every graph -> device (fp32 SoA); padding slots zero.
static int batch_upload(paradmm_batch* b)
{
    const size_t S = (size_t)b->slots;
    std::vector<float> x(3 * S, 0.0f), u(3 * S, 0.0f), n(3 * S, 0.0f), r(S, 0.0f), al(S, 0.0f);
    std::vector<float> z(3 * (size_t)b->ncirc);
    for (int i = 0; i < b->ninst; ++i) {
        graph* G = b->graphs[i];
        const BInst& in = b->inst[i];
        if (!G->X || !G->U || !G->N || !G->Z || !G->RHOS || !G->ALPHAS) {
            pdm_set_error(PARADMM_ERR_STATE, "batch instance %d: graph arrays not initialised", i);
            return PARADMM_ERR_STATE;
        }
        const std::vector<int>& pos = b->pos[i];
        for (int e = 0; e < G->num_edges; ++e) {
            const size_t p = (size_t)in.s0 + (size_t)pos[e];
            for (int k = 0; k < 3; ++k) {
                x[k * S + p] = (float)G->X[3L * e + k];
                u[k * S + p] = (float)G->U[3L * e + k];
                n[k * S + p] = (float)G->N[3L * e + k];
            }
            r[p] = (float)G->RHOS[e];
            al[p] = (float)G->ALPHAS[e];
        }
        for (int v = 0; v < in.N; ++v)
            for (int k = 0; k < 3; ++k) z[3L * (in.c0 + v) + k] = (float)G->Z[3L * v + k];
        BInst& bi = b->inst[i];
        bi.rho_u = (float)G->RHOS[0];
        bi.alpha_u = (float)G->ALPHAS[0];
        bi.uni = 1;
        for (int e = 1; e < G->num_edges && bi.uni; ++e)
            bi.uni = (float)G->RHOS[e] == bi.rho_u && (float)G->ALPHAS[e] == bi.alpha_u;
    }
    b->uni0.resize(b->ninst);
    for (int i = 0; i < b->ninst; ++i) b->uni0[i] = b->inst[i].uni;
    b->alpha1 = 1;
    for (int i = 0; i < b->ninst && b->alpha1; ++i) {
        graph* G = b->graphs[i];
        for (int e = 0; e < G->num_edges && b->alpha1; ++e) b->alpha1 = (float)G->ALPHAS[e] == 1.0f;
    }
    CK(hipMemcpyAsync(b->inst_d, b->inst.data(), sizeof(BInst) * (size_t)b->ninst, hipMemcpyHostToDevice, b->stream));
    std::vector<float> ir(2 * (size_t)b->ninst);
    for (int i = 0; i < b->ninst; ++i) {
        ir[2 * i] = b->inst[i].rho_u;
        ir[2 * i + 1] = b->inst[i].alpha_u;
    }
    CK(hipMemcpyAsync(b->inst_rho, ir.data(), sizeof(float) * ir.size(), hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->X, x.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->U, u.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->Nb, n.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->rho, r.data(), sizeof(float) * S, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->alpha, al.data(), sizeof(float) * S, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->Z, z.data(), sizeof(float) * z.size(), hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->Zprev, z.data(), sizeof(float) * z.size(), hipMemcpyHostToDevice, b->stream));
    return batch_reset_flags(b);
}

// One cooperative launch over iterations [it0, it_end) (see BArgs).
static int batch_launch(paradmm_batch* b, const paradmm_solve_opts* o, float* ms, int it0 = 0,
                        int it_end = 0x7fffffff, int first_mode = FIRST_NBUF, int finish_only = 0)
{
    BArgs a;
    a.ninst = b->ninst;
    a.max_iters = o->max_iters;
    a.check_every = o->check_every;
    a.it0 = it0;
    a.it_end = it_end;
    a.first_mode = first_mode;
    a.finish_only = finish_only;
    a.inst_rho = b->inst_rho;
    a.alpha1 = b->alpha1;
    a.tol_p = o->tol_primal;
    a.tol_d = o->tol_dual;
    a.tol_f = o->tol_feas;
    a.stride = b->slots;
    a.X = b->X;
    a.U = b->U;
    a.rho = b->rho;
    a.alpha = b->alpha;
    a.Nb = b->Nb;
    a.Z = b->Z;
    a.Zprev = b->Zprev;
    a.P = b->P;
    a.aexp = b->aexp_d;
    a.inst = b->inst_d;
    a.tiles = b->tiles_d;
    a.Rp = b->Rp;
    a.Rd = b->Rd;
    a.Q = b->Q;
    a.st = b->st;
    a.t_start = b->t_start;
    a.bar = b->bar;
    a.order = b->order_d;
    a.spin_limit = 1u << 24;
    a.barrier = b->barrier;
    void* args[] = {&a};
    CK(hipMemsetAsync(b->bar, 0, sizeof(unsigned) * b->bar_words, b->stream));
    const void* fn = b->T == 32 ? (const void*)k_batch_solve<32> : (const void*)k_batch_solve<64>;
    CK(hipEventRecord(b->ev0, b->stream));
    CK(hipLaunchCooperativeKernel(fn, dim3(b->grid), dim3(256), args, 0, b->stream));
    CK(hipEventRecord(b->ev1, b->stream));
    CK(hipEventSynchronize(b->ev1));
    CK(hipEventElapsedTime(ms, b->ev0, b->ev1));
    unsigned aborted = 0;
    CK(hipMemcpy(&aborted, b->bar + (GRIDBAR_GROUPS + 2) * GRIDBAR_LINE, sizeof aborted, hipMemcpyDeviceToHost));
    if (aborted) {
        pdm_set_error(PARADMM_ERR_HIP, "batch: a grid barrier wait timed out (workgroups not co-resident?)");
        return PARADMM_ERR_HIP;
    }
    return PARADMM_OK;
}

// ---- the contact state on the batch: host side

// the contact kernel for contact tile side CS and a batch of ninst instances
static const void* contact_kernel(int CS, int ninst)
{
    if (ninst <= 256) return CS == 64 ? (const void*)k_batch_contact<64, 256> : (const void*)k_batch_contact<32, 256>;
    return CS == 64 ? (const void*)k_batch_contact<64, BATCH_MAX> : (const void*)k_batch_contact<32, BATCH_MAX>;
}

static int contact_occupancy(const paradmm_batch* b, int CS, int* per_cu)
{
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, contact_kernel(CS, b->ninst), 256, 0);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor", __FILE__, __LINE__);
    if (*per_cu < 1) {
        pdm_set_error(PARADMM_ERR_HIP, "batch contact kernel does not fit a compute unit");
        return PARADMM_ERR_HIP;
    }
    return PARADMM_OK;
}

// The batch runs in the contact state when alpha = 1 and one rho is on every
// edge of each instance (as uploaded), for at least two iterations, with or
// without a rho policy; PARADMM_CONTACT=0 or paradmm_batch_configure_contact
// (b, 0, ..) keep the per-edge kernel (PARADMM_CONTACT_POLICY=0: with a
// policy).
static bool batch_contact_applies(const paradmm_batch* b, const paradmm_solve_opts* o)
{
    if (b->contact_off || !b->alpha1 || o->max_iters < 2) return false;
    const char* e = getenv("PARADMM_CONTACT");
    if (e && e[0] == '0') return false;
    if (o->policy_every > 0) {  // the policy inside the contact state (PARADMM_CONTACT_POLICY=0: per edge)
        const char* p = getenv("PARADMM_CONTACT_POLICY");
        if (!b->policy || (p && p[0] == '0')) return false;
    }
    for (int i = 0; i < b->ninst; ++i)
        if (!b->uni0[i]) return false;
    return true;
}

static int cmalloc(paradmm_batch* b, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        pdm_set_error(PARADMM_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return PARADMM_ERR_NOMEM;
    }
    b->callocs.push_back(*p);
    return PARADMM_OK;
}

// Contact tiling of every instance for side CS and its buffers (the lists
// hold every pair of a tile at worst: 32 B per pair and parity).
static int batch_contact_alloc(paradmm_batch* b, int CS)
{
    if (b->ccs_alloc == CS) return PARADMM_OK;
    batch_contact_free(b);
    b->cinst.assign(b->ninst, BCInst{});
    std::vector<int2> ct;
    long l0 = 0;
    for (int i = 0; i < b->ninst; ++i) {
        const int nbs = (b->inst[i].N + CS - 1) / CS, nts = nbs * (nbs + 1) / 2;
        b->cinst[i] = BCInst{nbs, nts, l0};
        for (int t = 0; t < nts; ++t) ct.push_back(make_int2(i, t));
        l0 += nts;
    }
    b->cts = l0;
    const size_t css = (size_t)CS * CS, words = (size_t)CS * (CS / 32), nc = (size_t)b->ncirc;
    RC(cmalloc(b, (void**)&b->cL, sizeof(float4) * 2 * 2 * (size_t)l0 * css));
    RC(cmalloc(b, (void**)&b->cLmask, sizeof(unsigned) * 2 * (size_t)l0 * words));
    RC(cmalloc(b, (void**)&b->ccnt, sizeof(int) * 2 * (size_t)l0));
    RC(cmalloc(b, (void**)&b->cAc, sizeof(unsigned long long) * 3 * CQ * nc));
    RC(cmalloc(b, (void**)&b->cZb, sizeof(float) * 6 * nc));
    RC(cmalloc(b, (void**)&b->cZpp, sizeof(float) * 3 * nc));
    RC(cmalloc(b, (void**)&b->cRpc, sizeof(double) * nc));
    RC(cmalloc(b, (void**)&b->ccq, sizeof(double) * (size_t)l0));
    RC(cmalloc(b, (void**)&b->cflags, sizeof(unsigned) * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->cpar, sizeof(int) * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->cinst_d, sizeof(BCInst) * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->ctile_d, sizeof(int2) * ct.size()));
    CK(hipMemcpy(b->cinst_d, b->cinst.data(), sizeof(BCInst) * (size_t)b->ninst, hipMemcpyHostToDevice));
    CK(hipMemcpy(b->ctile_d, ct.data(), sizeof(int2) * ct.size(), hipMemcpyHostToDevice));
    // work items of the launch-per-iteration form: the blocks' box and
    // radius edges of every instance first, then the tiles, larger
    // instances first
    std::vector<int> ord(b->ninst);
    for (int i = 0; i < b->ninst; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return b->cinst[x].nts > b->cinst[y].nts; });
    std::vector<int2> items;
    for (int i : ord)
        for (int k = 0; k < b->cinst[i].nbs; ++k) items.push_back(make_int2(i, k));
    for (int i : ord)
        for (int t = 0; t < b->cinst[i].nts; ++t) items.push_back(make_int2(i, b->cinst[i].nbs + t));
    b->nitems = (int)items.size();
    std::vector<int> ci(nc);
    for (int i = 0; i < b->ninst; ++i)
        for (int c = 0; c < b->inst[i].N; ++c) ci[(size_t)b->inst[i].c0 + c] = i;
    RC(cmalloc(b, (void**)&b->cW, sizeof(float) * 2 * nc));
    RC(cmalloc(b, (void**)&b->cXu, sizeof(float) * CUNARY * nc));
    RC(cmalloc(b, (void**)&b->cbase, sizeof(float) * nc));
    RC(cmalloc(b, (void**)&b->cyc, sizeof(float) * 3 * nc));
    RC(cmalloc(b, (void**)&b->crows, sizeof(int) * nc));
    RC(cmalloc(b, (void**)&b->cwacc, sizeof(long long) * 2 * nc));
    RC(cmalloc(b, (void**)&b->cysum, sizeof(long long) * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->chist, sizeof(double) * 6 * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->cgl, sizeof(float) * 2 * (size_t)b->ninst));
    RC(cmalloc(b, (void**)&b->citems_d, sizeof(int2) * items.size()));
    RC(cmalloc(b, (void**)&b->circ_inst_d, sizeof(int) * nc));
    CK(hipMemcpy(b->citems_d, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
    CK(hipMemcpy(b->circ_inst_d, ci.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    CK(hipMemset(b->ccq, 0, sizeof(double) * (size_t)l0));
    CK(hipMemset(b->cRpc, 0, sizeof(double) * nc));
    b->ccs_alloc = CS;
    return PARADMM_OK;
}

static BCArgs batch_contact_args(paradmm_batch* b, const paradmm_solve_opts* o, int it0, int it_end, int par0)
{
    BCArgs a;
    a.ninst = b->ninst;
    a.max_iters = o->max_iters;
    a.check_every = o->check_every;
    a.it0 = it0;
    a.it_end = it_end;
    a.par0 = par0;
    a.tol_p = o->tol_primal;
    a.tol_d = o->tol_dual;
    a.tol_f = o->tol_feas;
    a.T = b->T;
    a.stride = b->slots;
    a.X = b->X;
    a.U = b->U;
    a.Z = b->Z;
    a.Zprev = b->Zprev;
    a.Zpp = b->cZpp;
    a.Zb = b->cZb;
    a.W = b->cW;
    a.Xu = b->cXu;
    a.rho = b->rho;
    a.aexp = b->aexp_d;
    a.inst_rho = b->inst_rho;
    a.inst = b->inst_d;
    a.cinst = b->cinst_d;
    a.L = b->cL;
    a.Lmask = b->cLmask;
    a.cnt = b->ccnt;
    a.Ac = b->cAc;
    a.Rpc = b->cRpc;
    a.Rd = b->Rd;
    a.cq = b->ccq;
    a.flags = b->cflags;
    a.cpar = b->cpar;
    a.st = b->st;
    a.t_start = b->t_start;
    a.bar = b->bar;
    a.order = b->order_d;
    a.spin_limit = 1u << 24;
    a.barrier = b->cbarrier;
    return a;
}

static int contact_grid(paradmm_batch* b, int CS, int bpc, int* grid)
{
    if (bpc == 0) {  // one launch per iteration, no persistent grid
        *grid = 0;
        return PARADMM_OK;
    }
    int per_cu = 0;
    RC(contact_occupancy(b, CS, &per_cu));
    if (bpc < 0 || bpc > per_cu) {
        pdm_set_error(PARADMM_ERR_ARG, "blocks_per_cu %d outside 0..%d for contact tile %d", bpc, per_cu, CS);
        return PARADMM_ERR_ARG;
    }
    *grid = bpc * b->cus;
    return PARADMM_OK;
}

// The contact iterations 1 .. max_iters with one launch per iteration: runs
// of step launches up to each check, then the run's end, the quality and
// the check kernels; with a rho policy (check_every == policy_every) every
// instance still active after a check that more iterations follow takes
// the policy inside the contact state, and the later iterations read rho per
// edge.  The host stops once no instance is active.
template <int CS>
static int batch_contact_steps(paradmm_batch* b, const paradmm_solve_opts* o, const BCArgs& a)
{
    const unsigned ni = (unsigned)b->nitems, zg = (unsigned)((b->ncirc + 255) / 256);
    std::vector<BState> st(b->ninst);
    const bool policy = o->policy_every > 0 && b->policy;
    BCPol pp{};
    if (policy) {
        pp.params = mlp_device_params(b->policy);
        pp.raise_log = b->policy_mode == 3 ? 0.5f : __builtin_inff();
        pp.average = b->policy_average ? 1 : 0;
        pp.base = b->cbase;
        pp.yc = b->cyc;
        pp.rows = b->crows;
        pp.wacc = b->cwacc;
        pp.ysum = b->cysum;
        pp.hist = b->chist;
        pp.gl = b->cgl;
        CK(hipMemsetAsync(b->chist, 0, sizeof(double) * 6 * (size_t)b->ninst, b->stream));
        k_batch_cpol_base<<<zg, 256, 0, b->stream>>>(b->inst_rho, b->circ_inst_d, b->ncirc, b->cbase);
        CK(hipGetLastError());
    }
    bool re = false;  // rho per edge: after the first policy application
    int par = 0, it = 1;
    while (it < o->max_iters) {
        const int end = std::min(o->max_iters, (it / o->check_every + 1) * o->check_every);
        for (int j = 0; it + j < end; ++j) {
            const bool res = it + j + 1 == end;
            if (re) {
                if (res) k_batch_contact_step<CS, true, true><<<ni, 64, 0, b->stream>>>(a, b->citems_d, par, j);
                else k_batch_contact_step<CS, false, true><<<ni, 64, 0, b->stream>>>(a, b->citems_d, par, j);
            } else {
                if (res) k_batch_contact_step<CS, true, false><<<ni, 64, 0, b->stream>>>(a, b->citems_d, par, j);
                else k_batch_contact_step<CS, false, false><<<ni, 64, 0, b->stream>>>(a, b->citems_d, par, j);
            }
            par ^= 1;
        }
        if (re) k_batch_contact_zreduce<CS, true, true><<<zg, 256, 0, b->stream>>>(a, b->circ_inst_d, b->ncirc, end - it - 1);
        else k_batch_contact_zreduce<CS, true, false><<<zg, 256, 0, b->stream>>>(a, b->circ_inst_d, b->ncirc, end - it - 1);
        k_batch_contact_quality<CS><<<(unsigned)b->cts, 64, 0, b->stream>>>(a, b->ctile_d);
        k_batch_contact_check<<<(unsigned)b->ninst, 64, 0, b->stream>>>(a, end, end >= o->max_iters ? 1 : 0, par);
        CK(hipGetLastError());
        it = end;
        if (it >= o->max_iters) break;
        if (policy && it % o->policy_every == 0) {
            pp.frac = (float)it / (float)o->max_iters;
            pp.exceptions = re ? 1 : 0;
            CK(hipMemsetAsync(b->crows, 0, sizeof(int) * (size_t)b->ncirc, b->stream));
            CK(hipMemsetAsync(b->cwacc, 0, sizeof(long long) * 2 * (size_t)b->ncirc, b->stream));
            CK(hipMemsetAsync(b->cysum, 0, sizeof(long long) * (size_t)b->ninst, b->stream));
            const unsigned nt = (unsigned)b->cts;
            k_batch_cpol_tile<CS, false><<<nt, 256, 0, b->stream>>>(a, pp, b->ctile_d, par);
            k_batch_cpol_final<CS><<<(unsigned)b->ninst, 256, 0, b->stream>>>(a, pp, par);
            k_batch_cpol_tile<CS, true><<<nt, 256, 0, b->stream>>>(a, pp, b->ctile_d, par);
            k_batch_cpol_zprev<CS><<<zg, 256, 0, b->stream>>>(a, pp, b->circ_inst_d, b->ncirc, par);
            CK(hipGetLastError());
            re = true;
        }
        CK(hipMemcpyAsync(st.data(), b->st, sizeof(BState) * st.size(), hipMemcpyDeviceToHost, b->stream));
        CK(hipStreamSynchronize(b->stream));
        bool any = false;
        for (const BState& x : st) any = any || x.active;
        if (!any) break;
    }
    return PARADMM_OK;
}

// A contact solve from the uploaded state: the after-upload pass, the
// entry, one cooperative launch over iterations 1 .. max_iters, the exit
// and the pending U update.
// *ms: the four together (each timed with events; b->last_ms has them;
// the exit includes the pending U update of every edge).
static int batch_contact_solve(paradmm_batch* b, const paradmm_solve_opts* o, float* ms)
{
    if (b->cbpc < 0) {  // not configured: one launch per iteration (measured faster than the persistent grid)
        b->cbpc = 0;
        b->cgrid = 0;
    }
    RC(batch_contact_alloc(b, b->ccs));
    CK(hipMemsetAsync(b->cAc, 0, sizeof(unsigned long long) * 3 * CQ * (size_t)b->ncirc, b->stream));
    CK(hipMemsetAsync(b->cflags, 0, sizeof(unsigned) * (size_t)b->ninst, b->stream));
    float m0 = 0.0f;
    RC(batch_launch(b, o, &m0, 0, 1, FIRST_NBUF));
    BCArgs a = batch_contact_args(b, o, 1, o->max_iters, 0);
    const unsigned nt = (unsigned)b->cts;
    float me = 0.0f, mi = 0.0f, ml = 0.0f;
    CK(hipEventRecord(b->ev0, b->stream));
    if (b->ccs == 64) k_batch_contact_enter<64><<<nt, 256, 0, b->stream>>>(a, b->ctile_d);
    else k_batch_contact_enter<32><<<nt, 256, 0, b->stream>>>(a, b->ctile_d);
    CK(hipGetLastError());
    CK(hipEventRecord(b->ev1, b->stream));
    CK(hipEventSynchronize(b->ev1));
    CK(hipEventElapsedTime(&me, b->ev0, b->ev1));
    if (b->cbpc == 0 || o->policy_every > 0) {  // one launch per iteration (always with a policy)
        CK(hipEventRecord(b->ev0, b->stream));
        RC(b->ccs == 64 ? batch_contact_steps<64>(b, o, a) : batch_contact_steps<32>(b, o, a));
        CK(hipEventRecord(b->ev1, b->stream));
        CK(hipEventSynchronize(b->ev1));
        CK(hipEventElapsedTime(&mi, b->ev0, b->ev1));
    } else {  // one cooperative launch
        CK(hipMemsetAsync(b->bar, 0, sizeof(unsigned) * b->bar_words, b->stream));
        void* args[] = {&a};
        const void* fn = contact_kernel(b->ccs, b->ninst);
        CK(hipEventRecord(b->ev0, b->stream));
        CK(hipLaunchCooperativeKernel(fn, dim3(b->cgrid), dim3(256), args, 0, b->stream));
        CK(hipEventRecord(b->ev1, b->stream));
        CK(hipEventSynchronize(b->ev1));
        CK(hipEventElapsedTime(&mi, b->ev0, b->ev1));
        unsigned aborted = 0;
        CK(hipMemcpy(&aborted, b->bar + (GRIDBAR_GROUPS + 2) * GRIDBAR_LINE, sizeof aborted, hipMemcpyDeviceToHost));
        if (aborted) {
            pdm_set_error(PARADMM_ERR_HIP, "batch: a grid barrier wait timed out (workgroups not co-resident?)");
            return PARADMM_ERR_HIP;
        }
    }
    CK(hipEventRecord(b->ev0, b->stream));
    if (b->ccs == 64) k_batch_contact_leave<64><<<nt, 256, 0, b->stream>>>(a, b->ctile_d);
    else k_batch_contact_leave<32><<<nt, 256, 0, b->stream>>>(a, b->ctile_d);
    CK(hipGetLastError());
    if (b->T == 64) k_batch_finish_u<64><<<(unsigned)b->ntiles, 256, 0, b->stream>>>(b->tiles_d, b->inst_d, b->slots, b->X, b->U, b->Z);
    else k_batch_finish_u<32><<<(unsigned)b->ntiles, 256, 0, b->stream>>>(b->tiles_d, b->inst_d, b->slots, b->X, b->U, b->Z);
    CK(hipGetLastError());
    CK(hipEventRecord(b->ev1, b->stream));
    CK(hipEventSynchronize(b->ev1));
    CK(hipEventElapsedTime(&ml, b->ev0, b->ev1));
    b->last_ms[0] = m0;
    b->last_ms[1] = me;
    b->last_ms[2] = mi;
    b->last_ms[3] = ml;
    *ms = m0 + me + mi + ml;
    return PARADMM_OK;
}

// The learned-rho batch: launches of policy_every iterations; after each,
// the pending U update and the check of every active instance (a
// finish-only launch), then the policy on the instances still active.
static int batch_policy_solve(paradmm_batch* b, const paradmm_solve_opts* o, std::vector<BState>& st, float* total_ms)
{
    const int K = o->policy_every;
    float ms = 0.0f;
    *total_ms = 0.0f;
    int it = 0, first = FIRST_NBUF;
    if (b->policy_average) {
        if (!b->phist) RC(bmalloc(b, (void**)&b->phist, sizeof(double) * 2 * (size_t)b->ninst));
        CK(hipMemsetAsync(b->phist, 0, sizeof(double) * 2 * (size_t)b->ninst, b->stream));
    }
    for (;;) {
        const int end = std::min(it + K, o->max_iters);
        RC(batch_launch(b, o, &ms, it, end, first));
        *total_ms += ms;
        it = end;
        RC(batch_launch(b, o, &ms, it, it, FIRST_FUSED, 1));
        *total_ms += ms;
        CK(hipMemcpy(st.data(), b->st, sizeof(BState) * st.size(), hipMemcpyDeviceToHost));
        if (it >= o->max_iters) break;
        // active tiles in instance order, and each active instance's range
        std::vector<int> pt;
        std::vector<int2> ir;
        for (int i = 0; i < b->ninst; ++i) {
            if (!st[i].active) continue;
            ir.push_back(make_int2(i, (int)pt.size()));
            for (int t = 0; t < b->inst[i].nt; ++t) pt.push_back(b->inst[i].t0 + t);
        }
        if (ir.empty()) break;
        ir.push_back(make_int2(-1, (int)pt.size()));
        const auto t0 = std::chrono::steady_clock::now();
        CK(hipMemcpyAsync(b->ptiles, pt.data(), sizeof(int) * pt.size(), hipMemcpyHostToDevice, b->stream));
        CK(hipMemcpyAsync(b->iranges, ir.data(), sizeof(int2) * ir.size(), hipMemcpyHostToDevice, b->stream));
        BPolicyArgs pa;
        pa.ptiles = b->ptiles;
        pa.inst = b->inst_d;
        pa.tiles = b->tiles_d;
        pa.stride = b->slots;
        pa.X = b->X;
        pa.U = b->U;
        pa.rho = b->rho;
        pa.Z = b->Z;
        pa.Zprev = b->Zprev;
        pa.inst_rho = b->inst_rho;
        pa.frac = (float)it / (float)o->max_iters;
        pa.part = b->ppart;
        pa.cnt = b->pcnt;
        pa.iranges = b->iranges;
        pa.nact = (int)ir.size() - 1;
        pa.scale = b->pscale;
        pa.yout = nullptr;
        pa.raised = nullptr;
        pa.hist = b->policy_average ? b->phist : nullptr;
        if (b->policy_mode == 3) {
            if (!b->yout) RC(bmalloc(b, (void**)&b->yout, sizeof(float) * (size_t)b->slots));
            if (!b->raised) RC(bmalloc(b, (void**)&b->raised, sizeof(int) * (size_t)b->ninst));
            CK(hipMemsetAsync(b->raised, 0, sizeof(int) * (size_t)b->ninst, b->stream));
            pa.yout = b->yout;
            pa.raised = b->raised;
        }
        const unsigned npt = (unsigned)pt.size();
        const float* params = mlp_device_params(b->policy);
        if (b->T == 32) k_batch_policy_eval<32><<<npt, 256, 0, b->stream>>>(pa, params);
        else k_batch_policy_eval<64><<<npt, 256, 0, b->stream>>>(pa, params);
        CK(hipGetLastError());
        k_batch_policy_reduce<<<(unsigned)((pa.nact + 255) / 256), 256, 0, b->stream>>>(pa);
        CK(hipGetLastError());
        if (b->T == 32) k_batch_policy_apply<32><<<npt, 256, 0, b->stream>>>(pa);
        else k_batch_policy_apply<64><<<npt, 256, 0, b->stream>>>(pa);
        CK(hipGetLastError());
        CK(hipStreamSynchronize(b->stream));
        if (b->policy_mode == 3) {
            // an instance with raised edges leaves the uniform tile variant
            std::vector<int> nr(b->ninst);
            CK(hipMemcpy(nr.data(), b->raised, sizeof(int) * nr.size(), hipMemcpyDeviceToHost));
            bool changed = false;
            for (size_t k = 0; k + 1 < ir.size(); ++k) {
                const int i = ir[k].x;
                const int u = b->uni0[i] && nr[i] == 0;
                changed |= u != b->inst[i].uni;
                b->inst[i].uni = u;
            }
            if (changed)
                CK(hipMemcpy(b->inst_d, b->inst.data(), sizeof(BInst) * (size_t)b->ninst, hipMemcpyHostToDevice));
        }
        *total_ms += std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        first = FIRST_CURRENT;
    }
    return PARADMM_OK;
}

static int check_opts(const paradmm_solve_opts* o)
{
    if (!o || o->max_iters < 1 || o->check_every < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "batch: bad solve options");
        return PARADMM_ERR_ARG;
    }
    return PARADMM_OK;
}

extern "C" int paradmm_batch_create(paradmm_batch** out, graph* const* graphs, int count)
{
    if (!out || !graphs || count < 1 || count > BATCH_MAX) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_create: need 1..%d graphs", BATCH_MAX);
        return PARADMM_ERR_ARG;
    }
    for (int i = 0; i < count; ++i)
        if (!graphs[i]) {
            pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_create: graph %d is NULL", i);
            return PARADMM_ERR_ARG;
        }
    paradmm_batch* b = new paradmm_batch();
    b->graphs.assign(graphs, graphs + count);
    b->ninst = count;
    int dev = 0, khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        const int rc = hip_fail(e, "paradmm_batch_create", __FILE__, __LINE__);
        paradmm_batch_destroy(b);
        return rc;
    }
    b->clock_khz = khz;
    *out = b;
    return PARADMM_OK;
}

extern "C" void paradmm_batch_destroy(paradmm_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free_dev(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int paradmm_batch_configure(paradmm_batch* b, int tile, int blocks_per_cu, int barrier)
{
    if (!b || (tile != 32 && tile != 64) || (barrier != 0 && barrier != 1)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_configure: tile must be 32 or 64, barrier 0 or 1");
        return PARADMM_ERR_ARG;
    }
    int grid = 0;
    RC(grid_for(b, tile, blocks_per_cu, &grid));
    if (b->T != tile) RC(batch_layout(b, tile));
    b->blocks_per_cu = blocks_per_cu;
    b->grid = grid;
    b->barrier = barrier;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_configure_contact(paradmm_batch* b, int contact_tile, int blocks_per_cu, int barrier)
{
    if (!b || (contact_tile != 0 && contact_tile != 32 && contact_tile != 64) || (barrier != 0 && barrier != 1)) {
        pdm_set_error(PARADMM_ERR_ARG,
                      "paradmm_batch_configure_contact: contact tile 0 (off), 32 or 64; barrier 0 or 1");
        return PARADMM_ERR_ARG;
    }
    if (contact_tile == 0) {
        b->contact_off = true;
        return PARADMM_OK;
    }
    int grid = 0;
    RC(contact_grid(b, contact_tile, blocks_per_cu, &grid));
    b->contact_off = false;
    b->ccs = contact_tile;
    b->cbpc = blocks_per_cu;
    b->cgrid = grid;
    b->cbarrier = barrier;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_contact_info(paradmm_batch* b, int* contact_tile, int* blocks_per_cu, int* grid,
                                          long* contact_tiles, long* listed_pairs, double* phase_ms)
{
    if (!b) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_contact_info: NULL batch");
        return PARADMM_ERR_ARG;
    }
    if (contact_tile) *contact_tile = b->contact_off ? 0 : b->ccs;
    if (blocks_per_cu) *blocks_per_cu = b->cbpc;
    if (grid) *grid = b->cgrid;
    if (contact_tiles) *contact_tiles = b->ccs_alloc ? b->cts : 0;
    if (phase_ms)
        for (int k = 0; k < 4; ++k) phase_ms[k] = b->last_ms[k];
    if (listed_pairs) {  // the last iteration's list of every instance
        *listed_pairs = 0;
        if (b->ccs_alloc && b->last_ms[2] > 0.0f) {
            std::vector<int> cnt(2 * (size_t)b->cts), par(b->ninst);
            CK(hipMemcpy(cnt.data(), b->ccnt, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost));
            CK(hipMemcpy(par.data(), b->cpar, sizeof(int) * par.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < b->ninst; ++i) {
                const BCInst& c = b->cinst[i];
                for (int t = 0; t < c.nts; ++t) *listed_pairs += cnt[2 * (size_t)c.l0 + (size_t)par[i] * c.nts + t];
            }
        }
    }
    return PARADMM_OK;
}

extern "C" int paradmm_batch_autotune(paradmm_batch* b, int iters, paradmm_batch_config* chosen,
                                      paradmm_batch_config* tried, int max_tried, int* ntried)
{
    if (!b || iters < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_autotune: bad arguments");
        return PARADMM_ERR_ARG;
    }
    // every candidate runs `iters` iterations without convergence checks
    // (after one warm launch); the state keeps iterating from the graphs'
    // state uploaded once per tile side, which does not change the work
    paradmm_solve_opts o;
    memset(&o, 0, sizeof o);
    o.max_iters = iters;
    o.check_every = iters;
    o.tol_primal = o.tol_dual = -1.0;
    o.tol_feas = -1.0;
    paradmm_batch_config best{0, 0, 0, 0.0, 0};
    int n = 0;
    // the plain fixed-rho batch runs in the contact state when it can: then
    // its candidates are the contact kernel's (contact tile side x blocks
    // per CU x barrier), each timed from a fresh upload over iters - 1
    // contact iterations, with the dense layout at tile side 32 (which only
    // the after-upload pass, the entry and the exit use)
    if (!b->T) RC(paradmm_batch_configure(b, 32, 1, 1));
    RC(batch_upload(b));
    if (iters >= 3 && batch_contact_applies(b, &o)) {
        if (b->T != 32) {
            RC(paradmm_batch_configure(b, 32, 1, b->barrier));
        }
        // every candidate starts from the uploaded state: a device copy of
        // it (X, U, N, Z; rho and alpha do not change without a policy)
        // restored before each run instead of a host upload
        const size_t S = (size_t)b->slots, nz = 3 * (size_t)b->ncirc;
        float* snap = nullptr;
        CK(hipMalloc((void**)&snap, sizeof(float) * (9 * S + nz)));
        auto restore = [&](bool save) -> int {
            float* parts[4] = {b->X, b->U, b->Nb, b->Z};
            const size_t len[4] = {3 * S, 3 * S, 3 * S, nz};
            size_t off = 0;
            for (int k = 0; k < 4; ++k) {
                float* dst = save ? snap + off : parts[k];
                const float* src = save ? parts[k] : snap + off;
                CK(hipMemcpyAsync(dst, src, sizeof(float) * len[k], hipMemcpyDeviceToDevice, b->stream));
                off += len[k];
            }
            if (!save) {
                CK(hipMemcpyAsync(b->Zprev, b->Z, sizeof(float) * nz, hipMemcpyDeviceToDevice, b->stream));
                RC(batch_reset_flags(b));
            }
            CK(hipStreamSynchronize(b->stream));
            return PARADMM_OK;
        };
        int rc = restore(true);
        const int sides[2] = {32, 64};
        for (int si = 0; si < 2 && !rc; ++si) {
            const int CS = sides[si];
            int per_cu = 0;
            rc = contact_occupancy(b, CS, &per_cu);
            for (int bpc = 0; !rc && bpc <= per_cu && bpc <= 4; ++bpc)
                for (int bar = 0; !rc && bar < 2; ++bar) {
                    if (bpc == 0 && bar == 1) continue;  // no grid barrier in the launch-per-iteration form
                    float ms = 0.0f;
                    rc = paradmm_batch_configure_contact(b, CS, bpc, bar);
                    if (!rc) rc = restore(false);
                    if (!rc) rc = batch_contact_solve(b, &o, &ms);  // warm
                    if (!rc) rc = restore(false);
                    if (!rc) rc = batch_contact_solve(b, &o, &ms);
                    if (rc) break;
                    const double per = b->last_ms[2] / (iters - 1);
                    const paradmm_batch_config c{b->T, bpc, bar, per, CS};
                    if (tried && n < max_tried) tried[n] = c;
                    ++n;
                    if (best.contact_tile == 0 || per < best.ms_per_iter) best = c;
                }
        }
        (void)hipFree(snap);
        RC(rc);
        if (ntried) *ntried = n;
        RC(paradmm_batch_configure_contact(b, best.contact_tile, best.blocks_per_cu, best.barrier));
        if (chosen) *chosen = best;
        return PARADMM_OK;
    }
    const int tiles[2] = {32, 64};
    for (int ti = 0; ti < 2; ++ti) {
        const int T = tiles[ti];
        int per_cu = 0;
        hipError_t e = T == 32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_solve<32>, 256, 0)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_solve<64>, 256, 0);
        if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor", __FILE__, __LINE__);
        bool uploaded = false;
        for (int bpc = 1; bpc <= per_cu && bpc <= 4; ++bpc)
            for (int bar = 0; bar < 2; ++bar) {
                RC(paradmm_batch_configure(b, T, bpc, bar));
                if (!uploaded) RC(batch_upload(b));
                uploaded = true;
                float ms = 0.0f;
                RC(batch_reset_flags(b));
                RC(batch_launch(b, &o, &ms));  // warm
                RC(batch_reset_flags(b));      // the last check retired every instance
                RC(batch_launch(b, &o, &ms));
                const double per = ms / iters;
                if (tried && n < max_tried) tried[n] = paradmm_batch_config{T, bpc, bar, per, 0};
                ++n;
                if (best.tile == 0 || per < best.ms_per_iter) best = paradmm_batch_config{T, bpc, bar, per, 0};
            }
    }
    if (ntried) *ntried = n;
    RC(paradmm_batch_configure(b, best.tile, best.blocks_per_cu, best.barrier));
    if (chosen) *chosen = best;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_solve(paradmm_batch* b, const paradmm_solve_opts* o, paradmm_solve_result* res,
                                   double* kernel_ms)
{
    if (!b || !res) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_solve: bad arguments");
        return PARADMM_ERR_ARG;
    }
    RC(check_opts(o));
    if (o->policy_every > 0 && (!b->policy || o->check_every != o->policy_every)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_solve: policy_every > 0 needs a policy "
                                       "(paradmm_batch_set_rho_policy) and check_every == policy_every");
        return PARADMM_ERR_ARG;
    }
    if (!b->T) RC(paradmm_batch_configure(b, 64, 1, 1));
    RC(batch_upload(b));
    float ms = 0.0f;
    std::vector<BState> st(b->ninst);
    if (batch_contact_applies(b, o)) {
        RC(batch_contact_solve(b, o, &ms));
    } else if (o->policy_every <= 0) {
        RC(batch_launch(b, o, &ms));
    } else {
        RC(batch_policy_solve(b, o, st, &ms));
    }
    long long t0 = 0;
    CK(hipMemcpy(st.data(), b->st, sizeof(BState) * st.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(&t0, b->t_start, sizeof t0, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->ninst; ++i) {
        paradmm_solve_result& r = res[i];
        r.iterations = st[i].iterations;
        r.converged = st[i].converged;
        r.primal_rms = st[i].primal_rms;
        r.dual_rms = st[i].dual_rms;
        r.max_overlap = st[i].max_overlap;
        r.max_box = st[i].max_box;
        r.objective = st[i].objective;
        r.elapsed_ms = (double)(st[i].t_end - t0) / b->clock_khz;
    }
    if (kernel_ms) *kernel_ms = ms;
    b->solved = true;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_set_rho_policy(paradmm_batch* b, paradmm_rho_mlp* m)
{
    if (!b) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_set_rho_policy: NULL batch");
        return PARADMM_ERR_ARG;
    }
    b->policy = m;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_set_rho_policy_mode(paradmm_batch* b, int mode, int average)
{
    if (!b || (mode != 2 && mode != 3)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_set_rho_policy_mode: mode 2 or 3");
        return PARADMM_ERR_ARG;
    }
    b->policy_mode = mode;
    b->policy_average = average != 0;
    return PARADMM_OK;
}

extern "C" int paradmm_batch_sync_to_host(paradmm_batch* b)
{
    if (!b || !b->solved) {
        pdm_set_error(PARADMM_ERR_STATE, "paradmm_batch_sync_to_host: nothing solved");
        return PARADMM_ERR_STATE;
    }
    const size_t S = (size_t)b->slots;
    std::vector<float> x(3 * S), u(3 * S), r(S), z(3 * (size_t)b->ncirc);
    CK(hipMemcpy(x.data(), b->X, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
    CK(hipMemcpy(u.data(), b->U, sizeof(float) * 3 * S, hipMemcpyDeviceToHost));
    CK(hipMemcpy(r.data(), b->rho, sizeof(float) * S, hipMemcpyDeviceToHost));
    CK(hipMemcpy(z.data(), b->Z, sizeof(float) * z.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->ninst; ++i) {
        graph* G = b->graphs[i];
        const BInst& in = b->inst[i];
        const std::vector<int>& pos = b->pos[i];
        for (int v = 0; v < in.N; ++v)
            for (int k = 0; k < 3; ++k) G->Z[3L * v + k] = z[3L * (in.c0 + v) + k];
        for (int e = 0; e < G->num_edges; ++e) {
            const size_t p = (size_t)in.s0 + (size_t)pos[e];
            const int v = G->edgeix_to_varlistix[e];
            for (int k = 0; k < 3; ++k) {
                const double xe = x[k * S + p], ue = u[k * S + p];
                G->X[3L * e + k] = xe;
                G->U[3L * e + k] = ue;
                G->M[3L * e + k] = xe + ue;
                G->N[3L * e + k] = G->Z[3L * v + k] - ue;
            }
            G->RHOS[e] = r[p];
        }
    }
    return PARADMM_OK;
}

extern "C" int paradmm_batch_info(paradmm_batch* b, int* tile, int* blocks_per_cu, int* grid, long* slots, int* tiles)
{
    if (!b) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_batch_info: NULL batch");
        return PARADMM_ERR_ARG;
    }
    if (tile) *tile = b->T;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    if (grid) *grid = b->grid;
    if (slots) *slots = b->slots;
    if (tiles) *tiles = b->ntiles;
    return PARADMM_OK;
}

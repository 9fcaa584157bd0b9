// dense.hip -- dense pair-tile layout for the all-pairs unit-square packing.
//
// A graph whose factors are exactly: one box and one radius factor per
// circle and one pair factor for every unordered pair of circles (the
// unit-square packing, dim 3) is stored as tiles of T x T circle pairs
// (T = 32 by default, 64 on request).
// Tile (I, J), I <= J, holds two sub-blocks of T*T edge slots:
//   A[il][jl] = the edge of circle i = I*T+il in pair (i, j),
//   B[il][jl] = the edge of circle j = J*T+jl in the same pair,
// so the two edges of every pair factor sit at the same offset in A and B
// and a thread reads both with coalesced vector loads and no transpose.
// Diagonal tiles (I == J) use il < jl for pairs, put the box edge of circle
// c at A[cl][cl] and its radius edge at B[cl][cl]; the rest is padding.
// Every component is its own array of `stride` floats (SoA).
//
// One iteration is two kernels:
//   k_pairtile<DM_FUSED>: per edge U += alpha (X - Z) (the U update of the
//     previous iteration, with its FP64 primal residual), N = Z - U, the
//     prox, M = X + U, and per-tile partial sums of rho*M, rho, rho^2 for
//     every circle of the tile's two blocks; 56 algorithmic bytes per edge
//     (read X, U, rho, alpha; write X, U).
//   k_dense_zreduce: Z_c = sum rho M / sum rho over the circle's T-block
//     partials in fixed order (atomic-free, deterministic), and the FP64
//     dual residual.
// With alpha = 1 the layout runs the contact state instead (below): only
// the pairs whose prox is active keep per-edge values, in per-tile lists,
// and an iteration is one k_contact_tile launch (k_contact_zreduce ends a
// run of them).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "paradmm_gpu.h"
#include "mlp_device.h"
#include "pairtile.h"
#include "contact_device.h"
#include "contact_policy.h"
#include "contact_sparse.h"

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

namespace pdm {


// minimum waves per SIMD requested for the tile kernel (register budget)
#ifndef PDM_PAIRTILE_WAVES
#define PDM_PAIRTILE_WAVES 2
#endif

struct DenseState {
    int N = 0, nb = 0, nt = 0;
    int T = 64;               // circles per tile side (32 or 64)
    int2* d_tiles = nullptr;  // [nt] (I, J)
    float* d_P = nullptr;     // [5][nb][N] partial sums
    double* d_Rp = nullptr;   // [nt] primal residual partials
    double* d_Rd = nullptr;   // [N] dual residual partials
    float* d_aexp = nullptr;  // [N] radius exponent per circle
    float* d_uni = nullptr;   // [2] common rho, alpha (uniform runs)
    int* d_tile_pos = nullptr; // [nt] layout tile index -> launch position in d_tiles
    bool u_pending = false;   // U still lacks the update of the last iteration
    bool m_state = false;     // the X array holds M = X + U and U is stale (inside a captured run)

    // contact state (one rho, alpha = 1; see the contact kernels below):
    // the pair slots of X and U are stale, the contact lists hold the M of
    // the pairs whose prox was active, every other pair edge has M = Z
    bool c_state = false;
    float4* d_L = nullptr;     // [2][nts][cs*cs][2] contact lists
    int* d_cnt = nullptr;      // [2][nts] list lengths
    int cs = 64;               // contact tile side (32 or 64)
    int nbs = 0, nts = 0;      // contact tiling: blocks of cs circles, tiles
    int c_par = 0;             // the list the last iteration wrote (kept on the host, passed to the kernels)
    bool c_rhoe = false;       // rho differs between edges (fixed while in the state)
    float* d_W = nullptr;      // [2][N] sum rho, sum rho^2 per circle (per-edge rho)
    unsigned* d_Lmask = nullptr; // [2][nts][cs][cs/32] the pairs of each list, per tile row
    unsigned long long* d_Ac = nullptr;  // [3][CQ][N] circle sums of the last three launches of a run (the
                                         // spatial form uses the first two)
    float* d_Zb = nullptr;               // [2][N][3] Z inside a run (its canonical arrays hold the run's start)
    int c_run = 0;                       // contact launches since the last run end
    double* d_Rpc = nullptr;   // [N] primal residual per circle
    float* d_Xu = nullptr;     // [CUNARY][N] box and radius edges inside the contact state
    float* d_Zpp = nullptr;    // [N][3] Z two iterations back
    unsigned* d_flag = nullptr; // a circle sum lost a value (non-finite or out of range) since the state was entered

    // spatial form of the contact state (contact_sparse.h): circles in the
    // Morton order of their centres at the entry, tiles only where two
    // blocks' boxes meet; d_L, d_cnt, d_Lmask, cs, nbs, nts above are the
    // current form's lists, the other form's wait in c_alt
    struct ContactLists {
        float4* L = nullptr;
        int* cnt = nullptr;
        unsigned* Lmask = nullptr;
        int cs = 0, nbs = 0, nts = 0;
    };
    ContactLists c_alt;
    bool c_sparse = false;
    int* d_perm = nullptr;     // [N] contact index -> circle
    int* d_inv = nullptr;      // [N] circle -> contact index
    unsigned long long* d_key = nullptr;  // [N] Morton keys (entry)
    float4* d_zcv = nullptr;   // [2][N] (Z_k, c) per contact circle
    float4* d_bb = nullptr;    // [nbs] block boxes
    int row_waves = 4;         // wavefronts per row workgroup (2, 4, 8 or 16)
    int row_split = 1;         // workgroups per block row

    // learned rho applied inside the contact state (dense_contact_policy)
    bool c_base_ok = false;    // d_base holds every circle's unlisted pair edges' rho
    float* d_base = nullptr;   // [N] that rho
    void* d_cpol = nullptr;    // [N] rows, [3][N] outputs, [N] new rho, [N] factor, sums
};

// one workgroup per tile of side T; UNI: one rho and one alpha for every edge
template <int T, int MODE, bool UNI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PDM_PAIRTILE_WAVES))) void k_pairtile(DArgs a)
{
    __shared__ TileShared<T> sh;
    // launch order (a.tiles) differs from layout order: t is the layout index
    const int2 ij0 = a.tiles[blockIdx.x];
    const int2 ij = make_int2(ij0.x, ij0.y & ~TILE_RAISED);
    const int t = ij.x * a.nb - ij.x * (ij.x - 1) / 2 + (ij.y - ij.x);
    if constexpr (UNI) {
        if (ij0.y & TILE_RAISED) {
            // a tile holding edges of their own rho (a rho policy raised
            // them) reads the per-edge rho and alpha streams
            tile_pass<T, MODE, false, false>(a, t, ij.x, ij.y, sh, a.Rp + t, nullptr);
            return;
        }
        // the common values live in device memory, so captured iterations
        // follow later changes of them
        DArgs b = a;
        b.rho_u = a.uni[0];
        b.alpha_u = a.uni[1];
        tile_pass<T, MODE, false, UNI>(b, t, ij.x, ij.y, sh, a.Rp + t, nullptr);
    } else {
        tile_pass<T, MODE, false, UNI>(a, t, ij.x, ij.y, sh, a.Rp + t, nullptr);
    }
}

// Z_c = sum rho M / sum rho over the circle's nb block partials.  A
// workgroup takes 32 circles; its 32 slices of 32 threads sum the blocks
// K = slice, slice + 32, ... (coalesced 128-byte rows of P[q][K][.], four
// blocks in flight per thread), then the slices are added by a fixed-shape
// tree in LDS.
constexpr int ZR_SLICES = 32;

__global__ __launch_bounds__(32 * ZR_SLICES) void k_dense_zreduce(int N, int nb, const float* __restrict__ P,
                                                                  float* __restrict__ Z, float* __restrict__ Zprev,
                                                                  double* __restrict__ Rd)
{
    __shared__ float part[ZR_SLICES][5][32];
    const int ks = threadIdx.x >> 5, cl = threadIdx.x & 31;
    const int c = blockIdx.x * 32 + cl;
    float s[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (c < N)
#pragma unroll 4
        for (int K = ks; K < nb; K += ZR_SLICES)
#pragma unroll
            for (int q = 0; q < 5; ++q) s[q] += P[((long)q * nb + K) * N + c];
#pragma unroll
    for (int q = 0; q < 5; ++q) part[ks][q][cl] = s[q];
    __syncthreads();
    // fixed-shape tree over the slices
    for (int h = ZR_SLICES / 2; h > 0; h >>= 1) {
        if (ks < h)
#pragma unroll
            for (int q = 0; q < 5; ++q) part[ks][q][cl] += part[ks + h][q][cl];
        __syncthreads();
    }
    if (threadIdx.x >= 32 || c >= N) return;
#pragma unroll
    for (int q = 0; q < 5; ++q) s[q] = part[0][q][cl];
    double dz2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float z = s[k] / s[3];
        const float zo = Z[3 * c + k];
        const double d = (double)(z - zo);
        dz2 += d * d;
        Z[3 * c + k] = z;
        Zprev[3 * c + k] = zo;
    }
    Rd[c] = (double)s[4] * dz2;
}

// primal = sum Rp[0..nt), dual = sum Rd[0..N): one workgroup, fixed order.
// flag (or null): set when the contact state's sums lost a value; the
// residuals are then NaN
__global__ __launch_bounds__(256) void k_dense_res_final(int nt, const double* __restrict__ Rp, int N,
                                                         const double* __restrict__ Rd, double* __restrict__ out,
                                                         const unsigned* __restrict__ flag)
{
    __shared__ double sp[256], sq[256];
    double p = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < nt; i += 256) p += Rp[i];
    for (int i = threadIdx.x; i < N; i += 256) q += Rd[i];
    sp[threadIdx.x] = p;
    sq[threadIdx.x] = q;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sp[threadIdx.x] += sp[threadIdx.x + s];
            sq[threadIdx.x] += sq[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool lost = flag && *flag;
        out[0] = lost ? __builtin_nan("") : sp[0];
        out[1] = lost ? __builtin_nan("") : sq[0];
    }
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Launch: nbs + nts wavefronts, the first nbs the blocks' box and radius
// edges (contact_unary), then the contact tiles.
template <int CS, bool RES, bool RE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_contact_tile(CArgs a)
{
    __shared__ ContactWave<CS, RES ? CQ : 3> sh;
    if ((int)blockIdx.x < a.nbs) {
        contact_unary<CS, RES, RE>(a, blockIdx.x);
        return;
    }
    contact_tile_body<CS, RES, RE>(a, (int)blockIdx.x - a.nbs, sh);
}


// Per-edge rho: W[c] = sum rho and W[N + c] = sum rho^2 over circle c's N + 1
// edges, FP64 in a fixed order (lane j's pairs j, j + 64, ..., then the
// wavefront's butterfly).  One wavefront per circle, when the state is
// entered (rho does not change inside it).
__global__ __launch_bounds__(64) void k_contact_weights(CArgs a)
{
    const int c = blockIdx.x, lane = threadIdx.x;
    const int N = a.N;
    const long TT = (long)a.T * a.T;
    double s = 0.0, s2 = 0.0;
    for (int j = lane; j < N; j += 64) {
        if (j == c) continue;
        const long q = j > c ? dense_pair_slot(c, j, a.nb, a.T) : dense_pair_slot(j, c, a.nb, a.T) + TT;
        const double r = a.rho[q];
        s += r;
        s2 += r * r;
    }
    if (lane < 2) {  // the box and radius edges
        const double r = a.rho[dense_unary_slot(c, a.nb, a.T) + lane * TT];
        s += r;
        s2 += r * r;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        a.W[c] = (float)s;
        a.W[N + c] = (float)s2;
    }
}

template <bool RES, bool RE>
__global__ __launch_bounds__(CZ_THREADS) void k_contact_zreduce(CArgs a)
{
    const int c = blockIdx.x * CZ_THREADS + threadIdx.x;
    if (c < a.N) contact_zreduce_circle<RES, RE>(a, c);
}

// Enters the contact state (contact_enter_tile), one workgroup per contact tile.
template <int CS>
__global__ __launch_bounds__(256) void k_contact_enter(CArgs a)
{
    __shared__ ContactEnterShared<CS> sh;
    contact_enter_tile<CS>(a, (int)blockIdx.x, sh);
}

// Leaves the contact state (contact_leave_tile), one workgroup per contact tile.
template <int CS>
__global__ __launch_bounds__(256) void k_contact_leave(CArgs a)
{
    __shared__ ContactLeaveShared<CS> sh;
    contact_leave_tile<CS>(a, (int)blockIdx.x, sh);
}

// ------------------------------------------------------------------ spatial form of the contact state
// (contact_sparse.h): two launches per iteration, the circles, then the
// block rows' tile selection and tiles.

template <int CS, bool RES>
__global__ __launch_bounds__(64) void k_scs_circles(CArgs a)
{
    const int c = blockIdx.x * 64 + threadIdx.x;
    float n[3];
    const bool valid = c < a.N;
    scs_circle<RES>(a, valid ? c : a.N - 1, valid, n);
    scs_block_box<CS>(a, c, valid, n);
}

// split workgroups of SW wavefronts per block row P: part h lists the
// row's running tiles (P, Q), Q >= P, Q = P + h mod split, in Q order in
// LDS (a workgroup prefix over the candidates, SW 64 at a time); part 0
// also runs block P's box and radius edges; wavefront w takes the listed
// tiles w, w + SW, ...
constexpr int SCS_MAXB = 4096;  // blocks of a row (N <= 65,536 circles at tiles of 16)

// minimum waves per SIMD requested for the row kernel (register budget)
#ifndef PDM_SCS_WAVES_PER_EU
#define PDM_SCS_WAVES_PER_EU 4
#endif

template <int CS, bool RES, int SCS_WAVES>
__global__ __launch_bounds__(64 * SCS_WAVES) __attribute__((amdgpu_waves_per_eu(PDM_SCS_WAVES_PER_EU))) void k_scs_rows(
    CArgs a, int split)
{
    __shared__ ContactWave<CS, RES ? CQ : 3> sh[SCS_WAVES];
    __shared__ unsigned short run[SCS_MAXB];
    __shared__ int wsum[SCS_WAVES];
    // split workgroups share a row: part h takes the candidates Q = P + h mod split
    const int P = (int)blockIdx.x / split, h = (int)blockIdx.x % split, nbs = a.nbs;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const float4 bp = a.bb[P];
    int total = 0;
    for (int q0 = P + h; q0 < nbs; q0 += 64 * SCS_WAVES * split) {
        const int Qb = q0 + tid * split;
        const bool r = Qb < nbs && scs_tile_runs(a, bp, P, Qb);
        const unsigned long long b = __ballot(r);
        if (lane == 0) wsum[w] = (int)__popcll(b);
        __syncthreads();
        int base = total, all = 0;
#pragma unroll
        for (int k = 0; k < SCS_WAVES; ++k) {
            base += k < w ? wsum[k] : 0;
            all += wsum[k];
        }
        if (r) run[base + (int)__popcll(b & ((1ull << lane) - 1ull))] = (unsigned short)Qb;
        total += all;
        __syncthreads();
    }
    // the box and radius edges on a wavefront with one tile fewer than the rest, if any
    if (h == 0 && w == (total % SCS_WAVES)) scs_unary<CS, RES>(a, P);
    // the tile index is the wavefront's own (one LDS word for all lanes):
    // read as uniform, the tile's addresses and list resource stay scalar
    for (int k = w; k < total; k += SCS_WAVES) {
        const int Qb = __builtin_amdgcn_readfirstlane((int)run[k]);
        contact_tile_body<CS, RES, false, true>(a, P * nbs - P * (P - 1) / 2 + (Qb - P), P, Qb, sh[w]);
        wave_lds_fence();  // the next tile reuses the scratch
    }
}

template <bool RES>
__global__ __launch_bounds__(CZ_THREADS) void k_scs_zreduce(CArgs a)
{
    const int c = blockIdx.x * CZ_THREADS + threadIdx.x;
    if (c < a.N) scs_zreduce_circle<RES>(a, c);
}

// Morton key of each circle's centre, ties broken by the circle index
__global__ __launch_bounds__(256) void k_scs_keys(CArgs a, unsigned long long* key)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < a.N) key[i] = ((unsigned long long)scs_morton(a.Z[3 * i], a.Z[3 * i + 1]) << 32) | (unsigned)i;
}

// contact order: rank of each key among all N (keys are distinct), by
// counting the smaller ones (O(N^2) compares, once per entry)
__global__ __launch_bounds__(256) void k_scs_rank(int N, const unsigned long long* __restrict__ key, int* perm,
                                                  int* inv)
{
    __shared__ unsigned long long ks[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const unsigned long long my = i < N ? key[i] : ~0ull;
    int r = 0;
    for (int b = 0; b < N; b += 256) {
        __syncthreads();
        ks[threadIdx.x] = b + (int)threadIdx.x < N ? key[b + threadIdx.x] : ~0ull;
        __syncthreads();
        const int m = min(256, N - b);
        for (int q = 0; q < m; ++q) r += ks[q] < my ? 1 : 0;
    }
    if (i < N) {
        perm[r] = i;
        inv[i] = r;
    }
}

// Entry from the steady X, U state: every pair with M = X + U != Z_prev on
// either edge (the test of the id-order entry) goes to list 0 of its
// contact tile, appended by atomics (the order within a list changes no
// value); one workgroup per dense tile, a wavefront along a dense row.
template <int CS>
__global__ __launch_bounds__(256) void k_scs_enter(CArgs a)
{
    constexpr int W = cwords<CS>();
    const int T = a.T, N = a.N;
    int I, J;
    tri_tile((int)blockIdx.x, a.nb, I, J);
    const long S = a.stride, TT = (long)T * T, t0 = (long)blockIdx.x * 2 * TT;
    const int jl = (int)threadIdx.x % T, j = J * T + jl;
    if (j >= N) return;
    const int jc = a.inv[j];
    float zpj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) zpj[k] = a.Zp[3 * j + k];
    for (int il = (int)threadIdx.x / T; il < T; il += 256 / T) {
        const int i = I * T + il;
        if (i >= N || (I == J && il >= jl)) continue;
        const long pa = t0 + (long)il * T + jl, pb = pa + TT;
        float mA[3], mB[3];
        bool ex = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mA[k] = a.X[k * S + pa] + a.U[k * S + pa];
            mB[k] = a.X[k * S + pb] + a.U[k * S + pb];
            ex = ex || mA[k] != a.Zp[3 * i + k] || mB[k] != zpj[k];
        }
        if (!ex) continue;
        const int ic = a.inv[i];
        const bool f = ic < jc;  // circle i is the tile's row circle
        int ct, idx;
        scs_tile_of(f ? ic : jc, f ? jc : ic, CS, a.nbs, ct, idx);
        const int slot = atomicAdd(&a.cnt[ct], 1);
        float4* L0 = a.L + (long)ct * (2 * CS * CS);
        L0[2 * slot] = f ? make_float4(mA[0], mA[1], mA[2], __int_as_float(idx))
                         : make_float4(mB[0], mB[1], mB[2], __int_as_float(idx));
        L0[2 * slot + 1] = f ? make_float4(mB[0], mB[1], mB[2], 0.0f) : make_float4(mA[0], mA[1], mA[2], 0.0f);
        atomicOr(&a.Lmask[(long)ct * CS * W + (idx / CS) * W + ((idx % CS) >> 5)], 1u << ((idx % CS) & 31));
    }
}

// Entry: the box and radius edges into the compact array, in contact order
__global__ __launch_bounds__(256) void k_scs_enter_unary(CArgs a)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.N) return;
    const long N = a.N, S = a.stride, TT = (long)a.T * a.T;
    const long q0 = dense_unary_slot(a.perm[c], a.nb, a.T);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const long q = q0 + side * TT;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.Xu[(6 * side + k) * N + c] = a.X[k * S + q];
            a.Xu[(6 * side + 3 + k) * N + c] = a.U[k * S + q];
        }
        a.Xu[(12 + side) * N + c] = a.rho[q];
    }
}

// Leave, first pass: every pair edge as if in no list (U_k = Z_{k-1} - Z_k,
// X_k = Z_k - U_k, the id-order leave's values for such a pair) and the box
// and radius edges back to their slots; one workgroup per dense tile.
__global__ __launch_bounds__(256) void k_scs_leave_dense(CArgs a)
{
    const int T = a.T, N = a.N;
    int I, J;
    tri_tile((int)blockIdx.x, a.nb, I, J);
    const long S = a.stride, TT = (long)T * T, t0 = (long)blockIdx.x * 2 * TT;
    const int jl = (int)threadIdx.x % T, j = J * T + jl;
    if (I == J && (int)threadIdx.x < T) {
        const int c = I * T + (int)threadIdx.x;
        if (c < N) {
            const int cc = a.inv[c];
            const long q0 = dense_unary_slot(c, a.nb, T);
#pragma unroll
            for (int side = 0; side < 2; ++side)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    a.X[k * S + q0 + side * TT] = a.Xu[(6 * side + k) * (long)N + cc];
                    a.U[k * S + q0 + side * TT] = a.Xu[(6 * side + 3 + k) * (long)N + cc];
                }
        }
    }
    if (j >= N) return;
    float uj[3], xj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float zk = a.Zp[3 * j + k];
        uj[k] = a.Zpp[3 * j + k] - zk;
        xj[k] = zk - uj[k];
    }
    for (int il = (int)threadIdx.x / T; il < T; il += 256 / T) {
        const int i = I * T + il;
        if (i >= N || (I == J && il >= jl)) continue;
        const long pa = t0 + (long)il * T + jl, pb = pa + TT;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float zk = a.Zp[3 * i + k];
            const float u = a.Zpp[3 * i + k] - zk;
            a.U[k * S + pa] = u;
            a.X[k * S + pa] = zk - u;
            a.U[k * S + pb] = uj[k];
            a.X[k * S + pb] = xj[k];
        }
    }
}

// Leave, second pass: the pairs of the last two lists (U_k = M_{k-1} - Z_k
// from the older, X_k = M_k - U_k from the newer) over the first pass's
// values; one workgroup per contact tile.
template <int CS>
__global__ __launch_bounds__(256) void k_scs_leave_lists(CArgs a)
{
    constexpr int CSS = CS * CS, W = cwords<CS>();
    __shared__ unsigned short map[2][CSS];
    __shared__ unsigned mask[2][CS][W];
    const int t = blockIdx.x, tid = threadIdx.x;
    const long lt[2] = {(long)a.par * a.nts + t, (long)(a.par ^ 1) * a.nts + t};
    const int cn[2] = {min(a.cnt[lt[0]], CSS), min(a.cnt[lt[1]], CSS)};
    if (cn[0] == 0 && cn[1] == 0) return;  // the same for the whole workgroup
    const float4* Ls[2] = {a.L + lt[0] * (2 * CSS), a.L + lt[1] * (2 * CSS)};
    for (int e = tid; e < 2 * CS * W; e += 256) (&mask[0][0][0])[e] = 0u;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 2; ++b)
        for (int e = tid; e < cn[b]; e += 256) {
            const int idx = __float_as_int(Ls[b][2 * e].w) & (CSS - 1);
            map[b][idx] = (unsigned short)e;
            atomicOr(&mask[b][idx / CS][(idx % CS) >> 5], 1u << ((idx % CS) & 31));
        }
    __syncthreads();
    int P, Qb;
    tri_tile(t, a.nbs, P, Qb);
    const long S = a.stride;
    for (int idx = tid; idx < CSS; idx += 256) {
        const int il = idx / CS, jl = idx % CS;
        const bool in_k = (mask[0][il][jl >> 5] >> (jl & 31)) & 1u;
        const bool in_k1 = (mask[1][il][jl >> 5] >> (jl & 31)) & 1u;
        if (!in_k && !in_k1) continue;
        const int gi = a.perm[P * CS + il], gj = a.perm[Qb * CS + jl];
        long qi, qj;
        scs_pair_slots(gi, gj, a.nb, a.T, qi, qj);
        float4 ek[2] = {}, ek1[2] = {};
        if (in_k) {
            ek[0] = Ls[0][2 * map[0][idx]];
            ek[1] = Ls[0][2 * map[0][idx] + 1];
        }
        if (in_k1) {
            ek1[0] = Ls[1][2 * map[1][idx]];
            ek1[1] = Ls[1][2 * map[1][idx] + 1];
        }
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int v = side == 0 ? gi : gj;
            const long q = side == 0 ? qi : qj;
            const float mk[3] = {ek[side].x, ek[side].y, ek[side].z};
            const float mk1[3] = {ek1[side].x, ek1[side].y, ek1[side].z};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float zk = a.Zp[3 * v + k];
                const float u = (in_k1 ? mk1[k] : a.Zpp[3 * v + k]) - zk;
                a.U[k * S + q] = u;
                a.X[k * S + q] = (in_k ? mk[k] : zk) - u;
            }
        }
    }
}

// ------------------------------------------------------------------ learned rho in the contact state
// (contact_policy.h): one workgroup per contact tile, one for the graph rho,
// one thread per circle

template <int CS, bool APPLY>
__global__ __launch_bounds__(256) void k_cpol_tile(CPolArgs p)
{
    __shared__ CPolShared<CS> sh;
    cpol_tile_body<CS, APPLY>(p, (int)blockIdx.x, sh);
}

__global__ __launch_bounds__(256) void k_cpol_final(CPolArgs p, long E)
{
    __shared__ long long sh[256];
    cpol_final_body(p, E, sh);
}

__global__ __launch_bounds__(256) void k_cpol_zprev(CPolArgs p)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < p.c.N) cpol_zprev_circle(p, c);
}


__global__ void k_fill_f32(float* p, long n, const float* v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = *v;
}

// ------------------------------------------------------------------ host side

static inline long tile_index(int I, int J, int nb) { return (long)I * nb - (long)I * (I - 1) / 2 + (J - I); }

int dense_structure(graph* G, const std::vector<int>& fkind, int T, std::vector<int>& pos, std::vector<float>& aexp,
                    int& nb_out, int& nt_out)
{
    if (G->num_dims != 3) return -1;
    const int N = G->num_vars, E = G->num_edges, F = G->num_opts;
    if (N < 2) return -1;
    const long npairs = (long)N * (N - 1) / 2;
    if ((long)F != npairs + 2L * N || (long)E != 2 * npairs + 2L * N) return -1;
    std::vector<int> box(N, -1), rad(N, -1);
    aexp.assign(N, 1.0f);
    std::vector<uint64_t> seen(((size_t)N * N + 63) / 64, 0);
    std::vector<int> pair_f;
    pair_f.reserve(npairs);
    const int* ev = G->edgeix_to_varlistix;
    for (int f = 0; f < F; ++f) {
        const int K = fkind[f];
        const int deg = G->operatorix_to_numedges[f];
        const int* el = G->operatorix_to_edgelistix[f];
        if (K == PARADMM_KIND_BOX_3D || K == PARADMM_KIND_RADIUS_3D) {
            if (deg != 1) return -1;
            const int c = ev[el[0]];
            std::vector<int>& slot = K == PARADMM_KIND_BOX_3D ? box : rad;
            if (slot[c] != -1) return -1;
            slot[c] = el[0];
            if (K == PARADMM_KIND_RADIUS_3D) aexp[c] = (float)((const double*)G->list_of_operators_param_pointers[f])[0];
        } else if (K == PARADMM_KIND_PAIR_3D) {
            const int i = ev[el[0]], j = ev[el[1]];
            if (i >= j) return -1;
            const size_t b = (size_t)i * N + j;
            if (seen[b >> 6] >> (b & 63) & 1) return -1;
            seen[b >> 6] |= 1ull << (b & 63);
            pair_f.push_back(f);
        } else {
            return -1;
        }
    }
    if ((long)pair_f.size() != npairs) return -1;
    for (int c = 0; c < N; ++c)
        if (box[c] < 0 || rad[c] < 0) return -1;

    const long TT = (long)T * T;
    const int nb = (N + T - 1) / T;
    const int nt = (int)((long)nb * (nb + 1) / 2);
    const long stride = (long)nt * 2 * TT;
    if (stride > 0x7fffffffL) {
        pdm_set_error(PARADMM_ERR_SHAPE, "dense layout too large (%ld slots)", stride);
        return PARADMM_ERR_SHAPE;
    }
    pos.assign(E, -1);
    for (int f : pair_f) {
        const int* el = G->operatorix_to_edgelistix[f];
        const int i = ev[el[0]], j = ev[el[1]];
        const long b = tile_index(i / T, j / T, nb) * 2 * TT + (long)(i % T) * T + (j % T);
        pos[el[0]] = (int)b;
        pos[el[1]] = (int)(b + TT);
    }
    for (int c = 0; c < N; ++c) {
        const long b = tile_index(c / T, c / T, nb) * 2 * TT + (long)(c % T) * T + (c % T);
        pos[box[c]] = (int)b;
        pos[rad[c]] = (int)(b + TT);
    }
    nb_out = nb;
    nt_out = nt;
    return 0;
}

// Tile side.  A workgroup's 256 threads each take T*T/1024 rows of a tile,
// one after the other: side 32 has a quarter of the serial row loop of side
// 64 and four times the workgroups, so more loads are in flight; measured on
// MI355X it is the faster side from N = 500 to N = 8000 (and pads less).
// PARADMM_DENSE_TILE=64 selects side 64.
static int dense_tile_side()
{
    if (const char* e = getenv("PARADMM_DENSE_TILE")) {
        const int t = atoi(e);
        if (t == 32 || t == 64) return t;
    }
    return 32;
}

int dense_try_finalize(DevGraph* s, graph* G, const std::vector<int>& fkind)
{
    std::vector<int> pos;
    std::vector<float> aexp;
    int nb = 0, nt = 0;
    const int T = dense_tile_side();
    const int rc = dense_structure(G, fkind, T, pos, aexp, nb, nt);
    if (rc) return rc;
    const int N = G->num_vars, E = G->num_edges;
    const int* ev = G->edgeix_to_varlistix;
    const long stride = (long)nt * 2 * T * T;
    // launch order: the diagonal tiles (the longest: pairs, then the unary
    // factors) first, the tiles of the partial last block (the shortest)
    // last, so the final round of workgroups is the cheapest
    std::vector<int2> tiles;
    tiles.reserve(nt);
    const bool ragged = N % T != 0;
    for (int I = 0; I < nb; ++I) tiles.push_back(make_int2(I, I));
    for (int I = 0; I < nb; ++I)
        for (int J = I + 1; J < nb; ++J)
            if (!(ragged && J == nb - 1)) tiles.push_back(make_int2(I, J));
    if (ragged)
        for (int I = 0; I < nb - 1; ++I) tiles.push_back(make_int2(I, nb - 1));

    DenseState* d = new DenseState();
    s->dense = d;
    d->N = N;
    d->T = T;
    d->nb = nb;
    d->nt = nt;
    s->stride = stride;
    s->npos = stride;
    RC(dmalloc(s, (void**)&s->d_pos, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&s->d_edge_var_api, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&d->d_tiles, sizeof(int2) * (size_t)nt));
    RC(dmalloc(s, (void**)&d->d_tile_pos, sizeof(int) * (size_t)nt));
    RC(dmalloc(s, (void**)&d->d_aexp, sizeof(float) * (size_t)N));
    CK(hipMemcpy(s->d_pos, pos.data(), sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_edge_var_api, ev, sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(d->d_tiles, tiles.data(), sizeof(int2) * (size_t)nt, hipMemcpyHostToDevice));
    std::vector<int> tpos(nt);
    for (int k = 0; k < nt; ++k) tpos[tile_index(tiles[k].x, tiles[k].y, nb)] = k;
    CK(hipMemcpy(d->d_tile_pos, tpos.data(), sizeof(int) * (size_t)nt, hipMemcpyHostToDevice));
    CK(hipMemcpy(d->d_aexp, aexp.data(), sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
    return 0;
}

int dense_alloc_state(DevGraph* s)
{
    DenseState* d = s->dense;
    RC(dmalloc(s, (void**)&d->d_P, sizeof(float) * 5 * (size_t)d->N * d->nb));
    RC(dmalloc(s, (void**)&d->d_Rp, sizeof(double) * (size_t)d->nt));
    RC(dmalloc(s, (void**)&d->d_Rd, sizeof(double) * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_uni, sizeof(float) * 2));

    CK(hipMemsetAsync(d->d_Rp, 0, sizeof(double) * (size_t)d->nt, s->stream));
    CK(hipMemsetAsync(d->d_Rd, 0, sizeof(double) * (size_t)d->N, s->stream));
    return PARADMM_OK;
}

void dense_free(DevGraph* s)
{
    delete s->dense;
    s->dense = nullptr;
}

static DArgs dargs(DevGraph* s)
{
    DenseState* d = s->dense;
    DArgs a;
    a.N = d->N;
    a.nb = d->nb;
    a.stride = s->stride;
    a.X = s->d_X;
    a.U = s->d_U;
    a.rho = s->d_rho;
    a.alpha = s->d_alpha;
    a.Nb = s->d_N;
    a.Mb = s->d_M;
    a.Z = s->d_Z;
    a.P = d->d_P;
    a.Rp = d->d_Rp;
    a.aexp = d->d_aexp;
    a.tiles = d->d_tiles;
    a.rho_u = s->rho_u;
    a.alpha_u = s->alpha_u;
    a.uni = d->d_uni;
    return a;
}

static const char* mode_name(int mode)
{
    switch (mode) {
    case DM_FUSED: return "pairtile_fused";
    case DM_FUSED_CURRENT: return "pairtile_fused_current";
    case DM_FUSED_NBUF: return "pairtile_fused_nbuf";
    case DM_LEGACY_X: return "pairtile_legacy_x";
    case DM_LEGACY_Z: return "pairtile_legacy_z";
    case DM_M_ENTER: return "pairtile_m_enter";
    case DM_M: return "pairtile_m";
    case DM_M_EXIT: return "pairtile_m_exit";
    default: return "pairtile_finish";
    }
}

template <int T>
static void launch_pairtile_t(hipStream_t st, const DArgs& a, unsigned nt, int mode, bool uni)
{
    switch (mode) {
    case DM_FUSED:
        if (uni) k_pairtile<T, DM_FUSED, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_FUSED, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_FUSED_CURRENT:
        if (uni) k_pairtile<T, DM_FUSED_CURRENT, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_FUSED_CURRENT, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_FUSED_NBUF:
        if (uni) k_pairtile<T, DM_FUSED_NBUF, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_FUSED_NBUF, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_M_ENTER:
        if (uni) k_pairtile<T, DM_M_ENTER, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_M_ENTER, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_M:
        if (uni) k_pairtile<T, DM_M, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_M, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_M_EXIT:
        if (uni) k_pairtile<T, DM_M_EXIT, true><<<nt, 256, 0, st>>>(a);
        else k_pairtile<T, DM_M_EXIT, false><<<nt, 256, 0, st>>>(a);
        break;
    case DM_LEGACY_X: k_pairtile<T, DM_LEGACY_X, false><<<nt, 256, 0, st>>>(a); break;
    case DM_LEGACY_Z: k_pairtile<T, DM_LEGACY_Z, false><<<nt, 256, 0, st>>>(a); break;
    default: k_pairtile<T, DM_FINISH, false><<<nt, 256, 0, st>>>(a); break;
    }
}

static int launch_pairtile(DevGraph* s, int mode)
{
    DArgs a = dargs(s);
    const unsigned nt = (unsigned)s->dense->nt;
    // the steady-state modes have a variant without the per-edge rho/alpha
    // streams when every edge shares one rho and one alpha
    const bool uni = s->rho_uni && s->alpha_uni && (mode <= DM_FUSED_NBUF || mode >= DM_M_ENTER);
    prof_begin(s, mode_name(mode));
    if (s->dense->T == 32) launch_pairtile_t<32>(s->stream, a, nt, mode, uni);
    else launch_pairtile_t<64>(s->stream, a, nt, mode, uni);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

static int launch_zreduce(DevGraph* s)
{
    DenseState* d = s->dense;
    prof_begin(s, "zreduce");
    k_dense_zreduce<<<(unsigned)((d->N + 31) / 32), 32 * ZR_SLICES, 0, s->stream>>>(d->N, d->nb, d->d_P, s->d_Z,
                                                                               s->d_Zprev, d->d_Rd);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

// ---- contact state, host side

static CArgs cargs(DevGraph* s)
{
    DenseState* d = s->dense;
    CArgs a;
    a.N = d->N;
    a.nbs = d->nbs;
    a.nts = d->nts;
    a.nb = d->nb;
    a.T = d->T;
    a.stride = s->stride;
    a.X = s->d_X;
    a.U = s->d_U;
    a.Z = s->d_Z;
    a.Zp = s->d_Zprev;
    a.Zpp = d->d_Zpp;
    a.aexp = d->d_aexp;
    a.uni = d->d_uni;
    a.rho = s->d_rho;
    a.W = d->d_W;
    a.par = d->c_par;
    a.L = d->d_L;
    a.Lmask = d->d_Lmask;
    a.cnt = d->d_cnt;
    a.Ac = d->d_Ac;
    a.Ap = nullptr;
    a.Acl = nullptr;
    a.Zin = nullptr;
    a.Zpin = nullptr;
    a.Zo = nullptr;
    a.Rpc = d->d_Rpc;
    a.Rd = d->d_Rd;
    a.flag = d->d_flag;
    a.Xu = d->d_Xu;
    if (d->c_sparse) {
        a.perm = d->d_perm;
        a.inv = d->d_inv;
        a.zcv = d->d_zcv;
        a.bb = d->d_bb;
    }
    return a;
}

bool dense_contact_ok(DevGraph* s)
{
    const DenseState* d = s->dense;
    if (!d) return false;
    if (!(s->alpha_uni && s->alpha_u == 1.0f)) return false;
    const char* e = getenv("PARADMM_CONTACT");
    return !(e && e[0] == '0');
}

bool dense_in_contact(DevGraph* s) { return s->dense && s->dense->c_state; }

bool dense_contact_rho_per_edge(DevGraph* s) { return s->dense && s->dense->c_state && s->dense->c_rhoe; }

// Contact tile side: 64 (a quarter of the per-pair overhead of 32) once
// the tiles are enough to fill the GPU, 32 below N = 2500 (the smaller
// tiles spread a smaller packing over more wavefronts; measured crossover
// between N = 2000 and 3000).  The two give the same bits (the circle sums
// are integers).  PARADMM_CONTACT_TILE = 32 or 64 overrides.
static int contact_tile_side(int N)
{
    if (const char* e = getenv("PARADMM_CONTACT_TILE")) {
        const int v = atoi(e);
        if (v == 32 || v == 64) return v;
    }
    return N >= 2500 ? 64 : 32;
}

// The spatial form: from N = 3200 circles (below, the id-order form's one
// launch per iteration is as fast or faster: us per iteration, id-order and
// spatial, N = 2500 9.6, 10.1; N = 3000 10.3, 10.3; N = 3250 11.0, 10.5;
// N = 3500 11.9, 10.7; N = 4000 13.1, 10.8), with one rho, alpha = 1 and
// no rho policy (its application runs in the id-order form).
// PARADMM_CONTACT_SPARSE = 0 or 1 overrides the size rule.
static bool contact_sparse_wanted(DevGraph* s, bool rho_edge)
{
    const DenseState* d = s->dense;
    if (rho_edge || s->rho_policy) return false;
    if ((d->N + 15) / 16 > SCS_MAXB) return false;  // a row's blocks in LDS
    if (const char* e = getenv("PARADMM_CONTACT_SPARSE")) return e[0] == '1';
    return d->N >= 3200;
}

// Spatial tile side: 16 (PARADMM_SPARSE_TILE = 32 overrides): a block of
// 16 circles covers a patch small enough that a circle meets few blocks.
static int sparse_tile_side()
{
    if (const char* e = getenv("PARADMM_SPARSE_TILE"))
        if (atoi(e) == 32) return 32;
    return 16;
}

// the lists of the form wanted from here (the other form's are kept)
static void contact_select_form(DenseState* d, bool sparse)
{
    if (sparse == d->c_sparse) return;
    DenseState::ContactLists cur;
    cur.L = d->d_L;
    cur.cnt = d->d_cnt;
    cur.Lmask = d->d_Lmask;
    cur.cs = d->cs;
    cur.nbs = d->nbs;
    cur.nts = d->nts;
    d->d_L = d->c_alt.L;
    d->d_cnt = d->c_alt.cnt;
    d->d_Lmask = d->c_alt.Lmask;
    d->cs = d->c_alt.cs;
    d->nbs = d->c_alt.nbs;
    d->nts = d->c_alt.nts;
    d->c_alt = cur;
    d->c_sparse = sparse;
}

static int contact_alloc(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d->c_sparse && !d->d_perm) {
        const size_t N = (size_t)d->N;
        RC(dmalloc(s, (void**)&d->d_perm, sizeof(int) * N));
        RC(dmalloc(s, (void**)&d->d_inv, sizeof(int) * N));
        RC(dmalloc(s, (void**)&d->d_key, sizeof(unsigned long long) * N));
        RC(dmalloc(s, (void**)&d->d_zcv, sizeof(float4) * 2 * N));
    }
    if (d->d_L) return PARADMM_OK;
    d->cs = d->c_sparse ? sparse_tile_side() : contact_tile_side(d->N);
    d->nbs = (d->N + d->cs - 1) / d->cs;
    d->nts = d->nbs * (d->nbs + 1) / 2;
    if (d->c_sparse) {
        RC(dmalloc(s, (void**)&d->d_bb, sizeof(float4) * (size_t)d->nbs));
        // Workgroups of 4 wavefronts, each block row split over several of
        // them: the uneven rows are shared out better than by one
        // workgroup per row, and with about 11 running tiles per row most
        // wavefronts take one tile.  max(5, 1024 / nbs) parts (at most 8)
        // while the nbs x 5 workgroups fit on the chip at once (6 per CU at
        // the row kernel's 77 VGPRs: 1536 on 256 CUs), 4 up to 320 blocks
        // (N = 5120), above 3 (us per iteration, 3, 4 and 5 parts: N = 4000
        // 11.7, 11.6, 10.7; N = 4500 12.0, 12.1, 11.6; N = 4750 12.5, 12.6,
        // 12.1; N = 5000 12.3, 12.3, 12.8; N = 5500 13.2, 14.5, 13.7;
        // N = 6000 13.7, 14.7, 14.3; N = 8000 16.3, 17.3, 16.6; N = 10000
        // 18.8, 20.0, 19.2; 8 wavefronts per workgroup or 2 slower at each
        // N).  PARADMM_SPARSE_WAVES (2, 4, 8, 16) and PARADMM_SPARSE_SPLIT
        // (1-8) override.
        const char* e = getenv("PARADMM_SPARSE_WAVES");
        const int wv = e ? atoi(e) : 4;
        d->row_waves = wv == 2 || wv == 8 || wv == 16 ? wv : 4;
        e = getenv("PARADMM_SPARSE_SPLIT");
        const int sp = e ? atoi(e)
                         : d->nbs * 5 <= 1536 ? (1024 / d->nbs > 5 ? 1024 / d->nbs : 5)
                         : d->nbs <= 320      ? 4
                                              : 3;
        d->row_split = sp < 1 ? 1 : sp > 8 ? 8 : sp;
    }
    const size_t nts = (size_t)d->nts, css = (size_t)d->cs * d->cs,
                 words = (size_t)d->cs * (d->cs >= 32 ? d->cs / 32 : 1);
    RC(dmalloc(s, (void**)&d->d_L, sizeof(float4) * 2 * 2 * nts * css));
    RC(dmalloc(s, (void**)&d->d_cnt, sizeof(int) * 2 * nts));
    RC(dmalloc(s, (void**)&d->d_Lmask, sizeof(unsigned) * 2 * nts * words));
    CK(hipMemsetAsync(d->d_cnt, 0, sizeof(int) * 2 * nts, s->stream));
    CK(hipMemsetAsync(d->d_Lmask, 0, sizeof(unsigned) * 2 * nts * words, s->stream));
    if (d->d_Ac) return PARADMM_OK;  // the per-circle arrays are shared by both forms
    RC(dmalloc(s, (void**)&d->d_Ac, sizeof(unsigned long long) * 3 * CQ * (size_t)d->N));
    CK(hipMemsetAsync(d->d_Ac, 0, sizeof(unsigned long long) * 3 * CQ * (size_t)d->N, s->stream));
    RC(dmalloc(s, (void**)&d->d_Zb, sizeof(float) * 2 * 3 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Rpc, sizeof(double) * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Zpp, sizeof(float) * 3 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_W, sizeof(float) * 2 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_flag, sizeof(unsigned)));
    RC(dmalloc(s, (void**)&d->d_base, sizeof(float) * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Xu, sizeof(float) * CUNARY * (size_t)d->N));
    return PARADMM_OK;
}

// res: also the residual sums (the last iteration of a run)
template <typename K, typename... Args>
static void contact_launch(DevGraph* s, const char* name, K kernel, unsigned grid, unsigned block, const Args&... a)
{
    hipEvent_t e0, e1;
    prof_dispatch(s, name, e0, e1);
    if (e0)  // timed, eager
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, e0, e1, 0, a...);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, a...);
}

// Launch j of a run (j = c_run launches since the last run end) applies the
// sums of launch j - 1 to Z as it reads it and adds its own to buffer j % 3
// of three; its box-and-radius wavefronts store its Z in ring slot j % 2 and
// clear buffer (j + 1) % 3.  res (the last iteration of every run, and every
// single iteration): also the residual sums, then k_contact_zreduce ends the
// run into the canonical Z arrays, so outside a run the state is as it
// always was.
// The spatial form's launch j of a run: the circles (Z_k from the sums of
// launch j - 1 in buffer (j - 1) % 2, stored to ring slot j % 2; buffer
// j % 2 cleared), then the rows (the box and radius edges and the running
// tiles add to buffer j % 2); res: then k_scs_zreduce ends the run into
// the canonical arrays and clears both buffers.
static int sparse_iteration(DevGraph* s, bool res)
{
    DenseState* d = s->dense;
    const int j = d->c_run;
    const size_t AQ = (size_t)CQ * d->N, ZN = 3 * (size_t)d->N;
    unsigned long long* A = d->d_Ac;
    CArgs a = cargs(s);
    a.Ac = A + (j % 2) * AQ;
    a.Ap = j ? A + ((j - 1) % 2) * AQ : nullptr;
    a.Zin = j ? d->d_Zb + ((j - 1) % 2) * ZN : nullptr;
    a.Zo = d->d_Zb + (j % 2) * ZN;
    const unsigned cg = (unsigned)((d->N + 63) / 64), rg = (unsigned)(d->nbs * d->row_split);
    const char* cn = res ? "scs_circles_res" : "scs_circles";
    const char* rn = res ? "scs_rows_res" : "scs_rows";
#define SCS_LAUNCH(CS_, W_)                                                                           \
    do {                                                                                              \
        res ? contact_launch(s, cn, k_scs_circles<CS_, true>, cg, 64, a)                              \
            : contact_launch(s, cn, k_scs_circles<CS_, false>, cg, 64, a);                            \
        CK(hipGetLastError());                                                                        \
        res ? contact_launch(s, rn, k_scs_rows<CS_, true, W_>, rg, 64 * (W_), a, d->row_split)        \
            : contact_launch(s, rn, k_scs_rows<CS_, false, W_>, rg, 64 * (W_), a, d->row_split);      \
    } while (0)
    if (d->cs == 16) {
        if (d->row_waves == 16) SCS_LAUNCH(16, 16);
        else if (d->row_waves == 8) SCS_LAUNCH(16, 8);
        else if (d->row_waves == 2) SCS_LAUNCH(16, 2);
        else SCS_LAUNCH(16, 4);
    } else {
        if (d->row_waves == 16) SCS_LAUNCH(32, 16);
        else if (d->row_waves == 8) SCS_LAUNCH(32, 8);
        else if (d->row_waves == 2) SCS_LAUNCH(32, 2);
        else SCS_LAUNCH(32, 4);
    }
#undef SCS_LAUNCH
    CK(hipGetLastError());
    d->c_par ^= 1;
    if (!res) {
        d->c_run = j + 1;
        return PARADMM_OK;
    }
    CArgs f = cargs(s);
    f.Ac = A + (j % 2) * AQ;
    f.Acl = A + ((j + 1) % 2) * AQ;
    f.Zin = d->d_Zb + (j % 2) * ZN;
    f.Zpin = j ? d->d_Zb + ((j - 1) % 2) * ZN : nullptr;
    const unsigned zgrid = (unsigned)((d->N + CZ_THREADS - 1) / CZ_THREADS);
    contact_launch(s, "scs_zreduce_res", k_scs_zreduce<true>, zgrid, CZ_THREADS, f);
    CK(hipGetLastError());
    d->c_run = 0;
    return PARADMM_OK;
}

static int contact_iteration(DevGraph* s, bool res)
{
    DenseState* d = s->dense;
    if (d->c_sparse) return sparse_iteration(s, res);
    const int j = d->c_run;
    const size_t AQ = (size_t)CQ * d->N, ZN = 3 * (size_t)d->N;
    unsigned long long* A = d->d_Ac;
    CArgs a = cargs(s);
    a.Ac = A + (j % 3) * AQ;
    a.Ap = j ? A + ((j - 1) % 3) * AQ : nullptr;
    a.Acl = A + ((j + 1) % 3) * AQ;
    a.Zin = j == 0 ? nullptr : j == 1 ? s->d_Z : d->d_Zb + ((j - 1) % 2) * ZN;
    a.Zo = j ? d->d_Zb + (j % 2) * ZN : nullptr;
    const unsigned g = (unsigned)(d->nbs + d->nts);  // unary wavefronts, then the tiles
    const char* tn = res ? "contact_tile_res" : "contact_tile";
    const bool re = d->c_rhoe;
    if (d->cs == 64) {
        if (res) re ? contact_launch(s, tn, k_contact_tile<64, true, true>, g, 64, a)
                    : contact_launch(s, tn, k_contact_tile<64, true, false>, g, 64, a);
        else re ? contact_launch(s, tn, k_contact_tile<64, false, true>, g, 64, a)
                : contact_launch(s, tn, k_contact_tile<64, false, false>, g, 64, a);
    } else {
        if (res) re ? contact_launch(s, tn, k_contact_tile<32, true, true>, g, 64, a)
                    : contact_launch(s, tn, k_contact_tile<32, true, false>, g, 64, a);
        else re ? contact_launch(s, tn, k_contact_tile<32, false, true>, g, 64, a)
                : contact_launch(s, tn, k_contact_tile<32, false, false>, g, 64, a);
    }
    CK(hipGetLastError());
    d->c_par ^= 1;
    if (!res) {
        d->c_run = j + 1;
        return PARADMM_OK;
    }
    CArgs f = cargs(s);
    f.Ac = A + (j % 3) * AQ;
    f.Acl = A + ((j + 2) % 3) * AQ;  // launch j - 1's, read by launch j
    f.Zin = j == 0 ? s->d_Z : d->d_Zb + (j % 2) * ZN;
    f.Zpin = j == 0 ? s->d_Zprev : j == 1 ? s->d_Z : d->d_Zb + ((j - 1) % 2) * ZN;
    const unsigned zgrid = (unsigned)((d->N + CZ_THREADS - 1) / CZ_THREADS);
    re ? contact_launch(s, "contact_zreduce_res", k_contact_zreduce<true, true>, zgrid, CZ_THREADS, f)
       : contact_launch(s, "contact_zreduce_res", k_contact_zreduce<true, false>, zgrid, CZ_THREADS, f);
    CK(hipGetLastError());
    d->c_run = 0;
    return PARADMM_OK;
}

int dense_contact_parity(DevGraph* s) { return s->dense ? s->dense->c_par : 0; }

bool dense_contact_policy_ok(DevGraph* s)
{
    const DenseState* d = s->dense;
    if (!d || !d->c_state || !d->c_base_ok) return false;
    const char* e = getenv("PARADMM_CONTACT_POLICY");
    return !(e && e[0] == '0');
}

int dense_contact_policy(DevGraph* s, const float* params, float frac, int mode, bool average, double* hist)
{
    DenseState* d = s->dense;
    const int N = d->N;
    // [0, 16): output sum, graph rho; then rows [N] (int), the rho sums [2][N]
    // (int64), the outputs [3][N] (float): the first three start at zero
    const size_t zwords = 4 + (size_t)N + 4 * (size_t)N;
    if (!d->d_cpol) RC(dmalloc(s, &d->d_cpol, 4 * zwords + sizeof(float) * 3 * (size_t)N + 8));
    CPolArgs p;
    p.c = cargs(s);
    p.params = params;
    p.frac = frac;
    p.raise_log = mode == 3 ? 0.5f : __builtin_inff();
    p.exceptions = d->c_rhoe ? 1 : 0;
    p.average = average ? 1 : 0;
    p.base = d->d_base;
    p.ysum = (long long*)d->d_cpol;
    p.gl = (float*)((char*)d->d_cpol + 8);
    p.rows = (int*)((char*)d->d_cpol + 16);
    p.wacc = (long long*)((char*)d->d_cpol + 16 + 8 * (((size_t)N + 1) / 2));
    p.yc = (float*)((char*)d->d_cpol + 4 * zwords + 8);
    p.hist = hist;
    // a fill kernel, not a memset: the solve may be captured by a caller
    prof_begin(s, "cpol_fill");
    k_fill_u32<<<(unsigned)((zwords + 4 + 255) / 256), 256, 0, s->stream>>>((unsigned*)d->d_cpol, (int)(zwords + 4), 0u);
    prof_end(s);
    CK(hipGetLastError());
    const unsigned cg = (unsigned)((N + 255) / 256);
    prof_begin(s, "cpol_rows");
    if (d->cs == 64) k_cpol_tile<64, false><<<(unsigned)d->nts, 256, 0, s->stream>>>(p);
    else k_cpol_tile<32, false><<<(unsigned)d->nts, 256, 0, s->stream>>>(p);
    prof_end(s);
    CK(hipGetLastError());
    prof_begin(s, "cpol_final");
    k_cpol_final<<<1, 256, 0, s->stream>>>(p, (long)s->E);
    prof_end(s);
    CK(hipGetLastError());
    prof_begin(s, "cpol_apply");
    if (d->cs == 64) k_cpol_tile<64, true><<<(unsigned)d->nts, 256, 0, s->stream>>>(p);
    else k_cpol_tile<32, true><<<(unsigned)d->nts, 256, 0, s->stream>>>(p);
    prof_end(s);
    CK(hipGetLastError());
    prof_begin(s, "cpol_zprev");
    k_cpol_zprev<<<cg, 256, 0, s->stream>>>(p);
    prof_end(s);
    CK(hipGetLastError());
    d->c_rhoe = true;  // rho per edge from here (the circles' sums are set)
    return PARADMM_OK;
}

int dense_contact_info(DevGraph* s, int* tile_side, long* listed_pairs)
{
    DenseState* d = s->dense;
    *tile_side = 0;
    *listed_pairs = 0;
    if (!d || !d->c_state) return PARADMM_OK;
    std::vector<int> cnt((size_t)d->nts);
    CK(hipMemcpyAsync(cnt.data(), d->d_cnt + (size_t)d->c_par * d->nts, sizeof(int) * (size_t)d->nts,
                      hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    long n = 0;
    for (int c : cnt) n += c;
    *tile_side = d->cs;
    *listed_pairs = n;
    return PARADMM_OK;
}

void dense_contact_set_parity(DevGraph* s, int p)
{
    if (s->dense) s->dense->c_par = p & 1;
}

int dense_contact_enter(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d->c_state) return PARADMM_OK;
    if (!d->u_pending || d->m_state || !dense_contact_ok(s)) {
        pdm_set_error(PARADMM_ERR_STATE, "dense: the contact state is entered from the steady state with one rho "
                                         "and alpha = 1");
        return PARADMM_ERR_STATE;
    }
    // one rho everywhere (a policy's raised edges make it per edge)
    const bool rho_edge = !(s->rho_uni && s->policy_raised == 0);
    contact_select_form(d, contact_sparse_wanted(s, rho_edge));
    RC(contact_alloc(s));
    d->c_par = 0;  // the entry writes list 0
    d->c_run = 0;
    // a fresh state: no sum has lost a value (a fill kernel, not a memset,
    // as the entry may run inside a captured stream)
    k_fill_u32<<<1, 64, 0, s->stream>>>(d->d_flag, 1, 0u);
    CK(hipGetLastError());
    d->c_rhoe = rho_edge;
    // with one rho it is every circle's base rho (learned rho in the
    // id-order state; the spatial form takes no policy application)
    d->c_base_ok = !d->c_rhoe && !d->c_sparse;
    if (d->c_base_ok) {
        k_fill_f32<<<(unsigned)((d->N + 255) / 256), 256, 0, s->stream>>>(d->d_base, d->N, d->d_uni);
        CK(hipGetLastError());
    }
    const CArgs a = cargs(s);
    if (d->c_sparse) {
        const unsigned cg = (unsigned)((d->N + 255) / 256);
        const size_t words = (size_t)d->nts * d->cs * (d->cs >= 32 ? d->cs / 32 : 1);
        prof_begin(s, "scs_enter");
        k_scs_keys<<<cg, 256, 0, s->stream>>>(a, d->d_key);
        k_scs_rank<<<cg, 256, 0, s->stream>>>(d->N, d->d_key, d->d_perm, d->d_inv);
        // both lists empty (list 0 is filled below), list 0's masks clear
        k_fill_u32<<<(unsigned)((2 * (size_t)d->nts + 255) / 256), 256, 0, s->stream>>>(
            (unsigned*)d->d_cnt, 2 * d->nts, 0u);
        k_fill_u32<<<(unsigned)((words + 255) / 256), 256, 0, s->stream>>>(d->d_Lmask, (int)words, 0u);
        CK(hipGetLastError());
        const CArgs b = cargs(s);  // with the contact order
        if (d->cs == 16) k_scs_enter<16><<<(unsigned)d->nt, 256, 0, s->stream>>>(b);
        else k_scs_enter<32><<<(unsigned)d->nt, 256, 0, s->stream>>>(b);
        k_scs_enter_unary<<<cg, 256, 0, s->stream>>>(b);
        prof_end(s);
        CK(hipGetLastError());
        d->c_state = true;
        return contact_iteration(s, true);
    }
    if (d->c_rhoe) {
        prof_begin(s, "contact_weights");
        k_contact_weights<<<(unsigned)d->N, 64, 0, s->stream>>>(a);
        prof_end(s);
        CK(hipGetLastError());
    }
    prof_begin(s, "contact_enter");
    if (d->cs == 64) k_contact_enter<64><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
    else k_contact_enter<32><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    d->c_state = true;
    return contact_iteration(s, true);
}

int dense_contact_leave(DevGraph* s)
{
    DenseState* d = s->dense;
    if (!d || !d->c_state) return PARADMM_OK;
    const CArgs a = cargs(s);
    prof_begin(s, "contact_leave");
    if (d->c_sparse) {
        k_scs_leave_dense<<<(unsigned)d->nt, 256, 0, s->stream>>>(a);
        if (d->cs == 16) k_scs_leave_lists<16><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
        else k_scs_leave_lists<32><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
        prof_end(s);
        CK(hipGetLastError());
        d->c_state = false;
        return PARADMM_OK;
    }
    if (d->cs == 64) k_contact_leave<64><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
    else k_contact_leave<32><<<(unsigned)d->nts, 256, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    d->c_state = false;
    return PARADMM_OK;
}

int dense_iteration(DevGraph* s, int from_n, int mphase)
{
    DenseState* d = s->dense;
    if (d->c_state || mphase == DENSE_CONTACT) {
        // in the contact state every iteration is a contact iteration; only
        // the middle ones of a captured run skip the residual sums
        if (!d->c_state || from_n) {
            pdm_set_error(PARADMM_ERR_STATE, "dense: a contact iteration outside the contact state");
            return PARADMM_ERR_STATE;
        }
        return contact_iteration(s, mphase != DENSE_CONTACT);
    }
    int mode = from_n ? DM_FUSED_NBUF : (d->u_pending ? DM_FUSED : DM_FUSED_CURRENT);
    // a run of captured iterations with alpha = 1: enter the M state, stay,
    // and leave it on the run's last iteration (the state between runs is
    // always X and U)
    if (mphase == DENSE_M_ENTER) {
        if (mode != DM_FUSED || d->m_state) {
            pdm_set_error(PARADMM_ERR_STATE, "dense: the M state is entered from a pending U update only");
            return PARADMM_ERR_STATE;
        }
        mode = DM_M_ENTER;
    } else if (mphase == DENSE_M_MID || mphase == DENSE_M_EXIT) {
        if (!d->m_state) {
            pdm_set_error(PARADMM_ERR_STATE, "dense: not in the M state");
            return PARADMM_ERR_STATE;
        }
        mode = mphase == DENSE_M_MID ? DM_M : DM_M_EXIT;
    } else if (d->m_state) {
        pdm_set_error(PARADMM_ERR_STATE, "dense: a plain iteration in the M state");
        return PARADMM_ERR_STATE;
    }
    RC(launch_pairtile(s, mode));
    RC(launch_zreduce(s));
    d->u_pending = true;
    d->m_state = mode == DM_M_ENTER || mode == DM_M;
    return PARADMM_OK;
}

bool dense_steady(DevGraph* s) { return s->dense && s->dense->u_pending; }

float* dense_uni_ptr(DevGraph* s) { return s->dense ? s->dense->d_uni : nullptr; }

int2* dense_tile_flags(DevGraph* s, const int** tile_pos, int* nt, int* tile_slots)
{
    if (!s->dense) return nullptr;
    *tile_pos = s->dense->d_tile_pos;
    *nt = s->dense->nt;
    *tile_slots = 2 * s->dense->T * s->dense->T;
    return s->dense->d_tiles;
}

void dense_reset(DevGraph* s)
{
    if (s->dense) {
        s->dense->u_pending = false;
        s->dense->m_state = false;
        s->dense->c_state = false;
    }
}

int dense_finish(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d && d->c_state) RC(dense_contact_leave(s));
    if (!d || !d->u_pending) return PARADMM_OK;
    if (d->m_state) {
        pdm_set_error(PARADMM_ERR_STATE, "dense: finish inside the M state");
        return PARADMM_ERR_STATE;
    }
    RC(launch_pairtile(s, DM_FINISH));
    d->u_pending = false;
    return PARADMM_OK;
}

int dense_legacy_xpass(DevGraph* s)
{
    RC(dense_contact_leave(s));
    return launch_pairtile(s, DM_LEGACY_X);
}

int dense_legacy_zpass(DevGraph* s)
{
    RC(dense_contact_leave(s));
    RC(launch_pairtile(s, DM_LEGACY_Z));
    return launch_zreduce(s);
}

int dense_residual_reduce(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d->c_state) {  // per-circle residuals of the last contact iteration
        k_dense_res_final<<<1, 256, 0, s->stream>>>(d->N, d->d_Rpc, d->N, d->d_Rd, s->d_res, d->d_flag);
        CK(hipGetLastError());
        return PARADMM_OK;
    }
    RC(dense_finish(s));
    k_dense_res_final<<<1, 256, 0, s->stream>>>(d->nt, d->d_Rp, d->N, d->d_Rd, s->d_res, nullptr);
    CK(hipGetLastError());
    return PARADMM_OK;
}

}  // namespace pdm

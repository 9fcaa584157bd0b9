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
    unsigned long long* d_Ac = nullptr;  // [3][CQ][N] circle sums of the last three launches of a run
    float* d_Zb = nullptr;               // [2][N][3] Z inside a run (its canonical arrays hold the run's start)
    int c_run = 0;                       // contact launches since the last run end
    double* d_Rpc = nullptr;   // [N] primal residual per circle
    float* d_Zpp = nullptr;    // [N][3] Z two iterations back
    unsigned* d_flag = nullptr; // a circle sum lost a value (non-finite or out of range) since the state was entered

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

// ------------------------------------------------------------------ contact state
//
// With alpha = 1 the U update is U_{k+1} = M_k - Z_{k+1}.  A pair factor
// whose prox is inactive returns X = N = Z_k - U_k, so M_k = X + U = Z_k:
// its edges carry no state of their own, and at the next iteration
// U = Z_k - Z_{k+1}, N = Z_{k+1} - U, the same for every inactive pair of a
// circle.  So only the pairs whose prox was active (the contacts, about 1%
// of all pairs for N = 5000 circles) need per-edge values.  The contact
// state covers the pairs with contact tiles of CS x CS circles (its own
// tiling, independent of the dense layout's); a contact tile keeps a list
// of its active pairs, each entry the M of both edges and the pair's slot,
// and an iteration reads that list, tests every pair of the tile against
// the circles' common N, runs the prox on the colliding ones and writes the
// new list.  The box and radius edges stay in their dense X and U slots.
//
// Per circle a tile sums, over its pair edges that are in either list,
//   a = M - Z_k                      (consensus: Z_{k+1} = Z_k + sum a / (N + 1)),
//   g = (X - Z_k) - (Z_k - Z_{k-1})  and |g|^2,
// where a and g are 0 on every other edge.  The primal residual
// sum |X - Z_{k+1}|^2 is then, per circle, |g - e|^2 summed over all N + 1
// edges with e = Z_{k+1} - 2 Z_k + Z_{k-1}: S2 - 2 e.S1 + (N + 1)|e|^2.
// The sums run in fixed point, each value rounded to the nearest multiple
// of 2^-31 (64-bit LDS atomics inside a tile, then 64-bit global atomics of
// the tile's nonzero sums into the circle's accumulator), so they do not
// depend on the order the lanes or the tiles add them: runs are bitwise
// reproducible.  |g|^2 is held exactly to 2^-62 as two sums, its multiple
// of 2^-31 and the remainder in units of 2^-62, so the primal residual does
// not lose the small terms of a nearly converged iterate.  A value the
// fixed point cannot hold (non-finite, or |v| >= 2^30) is not added: it
// sets the state's flag, and the residuals read NaN from then on.
//
// The per-pair test replaces 24 bytes of M per edge read and written by
// ~10 VALU operations; the state is Z, Z_prev, Z_prev2 and the lists.

constexpr int CQ = 8;         // per-circle sums: a (3), g (3), |g|^2 (multiple of 2^-31, remainder in 2^-62)
constexpr double CFIX_INV = 1.0 / 2147483648.0;  // 2^-31
constexpr double CFIX2_INV = CFIX_INV * CFIX_INV; // 2^-62

struct CArgs {
    int N;
    int nbs, nts;          // contact tiling: blocks of CS circles, tiles
    int nb, T;             // dense layout: blocks, tile side (box and radius slots)
    long stride;
    float* X;
    float* U;
    float* Z;
    float* Zp;
    float* Zpp;
    const float* aexp;
    const float* uni;      // {rho, alpha}
    const float* rho;      // per-edge rho (dense slots), read when rho differs between edges
    float* W;              // [2][N] sum rho, sum rho^2 over each circle's edges (per-edge rho)
    int par;               // the list the last iteration wrote
    float4* L;             // [2][nts][CS*CS][2]
    unsigned* Lmask;       // [2][nts][CS][CS/32] pairs of each list, per tile row
    int* cnt;              // [2][nts]
    unsigned long long* Ac; // [CQ][N] this launch's circle sums, 2^-31 fixed point (starts at zero)
    const unsigned long long* Ap; // [CQ][N] the previous launch's sums, not yet applied to Z (or null)
    unsigned long long* Acl; // [CQ][N] the sums two launches back, cleared here (or null)
    const float* Zin;      // the Z the pending sums Ap apply to
    const float* Zpin;     // (run end) the Z before Zin
    float* Zo;             // Zin + Ap / (N + 1): this launch's Z, stored by the box-and-radius wavefronts
    double* Rpc;
    double* Rd;
    unsigned* flag;        // set when a circle sum could not take a value
};

// v rounded to the nearest multiple of 2^-31, in units of 2^-31:
// v = k + f with k = rint(v) and f = v - k exact (|f| <= 1/2), so
// v 2^31 = k 2^31 + rint(f 2^31) needs no FP64.  A value outside
// |v| < 2^30 (NaN included) gives 0 and sets bad.
__device__ __forceinline__ long long cfix(float v, unsigned& bad)
{
    if (!(fabsf(v) < 1073741824.0f)) {
        bad = 1u;
        return 0;
    }
    const float k = rintf(v);
    const float f = v - k;
    return (long long)(int)k * 2147483648LL + (long long)(int)rintf(f * 2147483648.0f);
}

// v >= 0 exactly to 2^-62 as hi 2^-31 + lo 2^-62: hi = the multiple of
// 2^-31 below v, lo = rint((v - hi 2^-31) 2^62) with the remainder exact
__device__ __forceinline__ void cfix2(float v, long long& hi, long long& lo, unsigned& bad)
{
    if (!(fabsf(v) < 1073741824.0f)) {
        bad = 1u;
        hi = lo = 0;
        return;
    }
    const float t = v * 2147483648.0f;  // exact
    const float k = floorf(t);
    hi = (long long)k;
    lo = (long long)rintf((t - k) * 2147483648.0f);  // t - k exact, in [0, 1)
}

// a circle sum (units of 2^-31) as fp32: one rounding, no FP64
__device__ __forceinline__ float csum_f32(unsigned long long v) { return (float)(long long)v * 4.656612873077393e-10f; }

template <typename SH>
__device__ __forceinline__ void cadd(SH& sh, unsigned long long* acc, float v)
{
    atomicAdd(acc, (unsigned long long)cfix(v, sh.bad));
}

template <typename SH>
__device__ __forceinline__ void cadd2(SH& sh, unsigned long long* acc, float v)
{
    long long hi, lo;
    cfix2(v, hi, lo, sh.bad);
    atomicAdd(acc, (unsigned long long)hi);
    atomicAdd(acc + 1, (unsigned long long)lo);
}

// (I, J) of tile t of a triangular tiling with nb blocks: row I starts at
// I nb - I (I - 1) / 2 (the float estimate is corrected by the loops)
__device__ __forceinline__ void tri_tile(int t, int nb, int& I, int& J)
{
    const float b = 2.0f * nb + 1.0f;
    I = (int)((b - sqrtf(fmaxf(b * b - 8.0f * (float)t, 0.0f))) * 0.5f);
    I = max(0, min(I, nb - 1));
    while (I > 0 && I * nb - I * (I - 1) / 2 > t) --I;
    while (I + 1 < nb && (I + 1) * nb - (I + 1) * I / 2 <= t) ++I;
    J = I + (t - (I * nb - I * (I - 1) / 2));
}

// dense-layout slot of the edge of circle i in pair (i, j), i < j (the
// edge of j is T*T further)
__device__ __forceinline__ long dense_pair_slot(int i, int j, int nb, int T)
{
    const int I = i / T, J = j / T;
    const long t = (long)I * nb - (long)I * (I - 1) / 2 + (J - I);
    return t * (2L * T * T) + (long)(i % T) * T + (j % T);
}

// box slot of circle c (its radius slot is T*T further)
__device__ __forceinline__ long dense_unary_slot(int c, int nb, int T)
{
    const int I = c / T;
    const long t = (long)I * nb - (long)I * (I - 1) / 2;
    return t * (2L * T * T) + (long)(c % T) * T + (c % T);
}

// Exclusive position of this thread's n items among the workgroup's (in
// thread order) and the workgroup total.  Contains one __syncthreads.
__device__ __forceinline__ int block_offset(int n, int (&wsum)[4], int& total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int s = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(s, o, 64);
        s += lane >= o ? y : 0;
    }
    if (lane == 63) wsum[w] = s;
    __syncthreads();
    int base = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        base += q < w ? wsum[q] : 0;
        total += wsum[q];
    }
    return base + s - n;
}

// Orders one wavefront's LDS accesses across its lanes.  A wavefront's LDS
// operations execute in issue order, so a wavefront-scope fence (a compiler
// ordering point, no instruction) suffices; s_waitcnt 0 would also wait for
// the wavefront's list stores to reach memory.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Shared scratch of one contact tile (one wavefront).  The pairs to project
// form one stream per tile: the previous list's pairs, then the colliding
// pairs found by the tests, which pass through a ring in LDS.
constexpr int CRING = 2048;  // > 63 pending + the 1024 pairs one test step can add

template <int CS, int Q>
struct ContactWave {
    unsigned long long acc[2 * CS][Q];  // fixed-point circle sums: rows (block P), then columns (block Q)
    float zc[2 * CS][6];                // Z_k and c = Z_k - Z_{k-1} of the tile's 2 CS circles
    unsigned short ring[CRING];         // colliding pairs not in the previous list (slot)
    unsigned mask[CS][CS / 32];         // pairs in the previous list, per row
    unsigned nmask[CS][CS / 32];        // pairs in the new list
    unsigned bad;                       // a circle sum could not take a value
};

// One pair of the contact tile through the prox: N from its list entry (ex)
// or the circles' common N, the circle sums; returns whether the prox was
// active, with the new M of both edges.
template <int CS, int Q, bool RE>
__device__ __forceinline__ bool contact_pair(ContactWave<CS, Q>& sh, int idx, bool ex, float4 pa, float4 pb,
                                             bool diag, float rA, float rB, float (&mA)[3], float (&mB)[3])
{
    const int rl = idx / CS, cl = idx % CS;
    const int cs = diag ? cl : CS + cl;
    float zA[3], zB[3], cA[3], cB[3], uA[3], uB[3], nA[3], nB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zA[k] = sh.zc[rl][k];
        cA[k] = sh.zc[rl][3 + k];
        zB[k] = sh.zc[CS + cl][k];
        cB[k] = sh.zc[CS + cl][3 + k];
    }
    if (ex) {
        const float ma[3] = {pa.x, pa.y, pa.z}, mb[3] = {pb.x, pb.y, pb.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uA[k] = ma[k] - zA[k];
            uB[k] = mb[k] - zB[k];
            nA[k] = zA[k] - uA[k];
            nB[k] = zB[k] - uB[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uA[k] = -cA[k];  // Z_{k-1} - Z_k, exactly
            uB[k] = -cB[k];
            nA[k] = zA[k] + cA[k];  // Z_k - (Z_{k-1} - Z_k), exactly
            nB[k] = zB[k] + cB[k];
        }
    }
    float xA[3] = {nA[0], nA[1], nA[2]}, xB[3] = {nB[0], nB[1], nB[2]};
    const bool act = pair_project(nA[0], nA[1], nA[2], nB[0], nB[1], nB[2], rA, rA, rB, rB, xA[0], xA[1], xA[2],
                                  xB[0], xB[1], xB[2]);
    if constexpr (Q == CQ) if (act || ex) {  // g = 0 on the other edges
        float gA2 = 0.0f, gB2 = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float gA = (xA[k] - zA[k]) - cA[k];
            const float gB = (xB[k] - zB[k]) - cB[k];
            cadd(sh, &sh.acc[rl][3 + k], gA);
            cadd(sh, &sh.acc[cs][3 + k], gB);
            gA2 += gA * gA;
            gB2 += gB * gB;
        }
        cadd2(sh, &sh.acc[rl][6], gA2);
        cadd2(sh, &sh.acc[cs][6], gB2);
    }
    if (act) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mA[k] = xA[k] + uA[k];
            mB[k] = xB[k] + uB[k];
            // per-edge rho: the consensus weighs each deviation with its edge's rho
            cadd(sh, &sh.acc[rl][k], RE ? rA * (mA[k] - zA[k]) : mA[k] - zA[k]);
            cadd(sh, &sh.acc[cs][k], RE ? rB * (mB[k] - zB[k]) : mB[k] - zB[k]);
        }
    }
    return act;
}

// One round of the tile's pair stream: items d .. d + 63 (item < cp: entry
// of the previous list, else ring slot item - cp); the active pairs are
// appended to the new list in lane order.
template <int CS, int Q, bool RE>
__device__ __forceinline__ void contact_round(ContactWave<CS, Q>& sh, const CArgs& a, int P, int Qb, int d, int avail,
                                              int cp, const float4* Lp, float4 pre0, float4 pre1, float4* Ln,
                                              int& base, bool diag, float rho)
{
    constexpr int CSS = CS * CS;
    const int lane = threadIdx.x;
    const int item = d + lane;
    bool act = false;
    int idx = 0;
    float mA[3], mB[3];
    if (item < avail) {
        const bool ex = item < cp;
        float4 pa = pre0, pb = pre1;
        if (ex) {
            if (d > 0) {
                pa = Lp[2 * item];
                pb = Lp[2 * item + 1];
            }
            idx = __float_as_int(pa.w) & (CSS - 1);
        } else {
            idx = sh.ring[(item - cp) & (CRING - 1)];
        }
        float rA = rho, rB = rho;
        if constexpr (RE) {  // the pair's own rho from its dense slots
            const long q = dense_pair_slot(P * CS + idx / CS, Qb * CS + idx % CS, a.nb, a.T);
            rA = a.rho[q];
            rB = a.rho[q + (long)a.T * a.T];
        }
        act = contact_pair<CS, Q, RE>(sh, idx, ex, pa, pb, diag, rA, rB, mA, mB);
    }
    const unsigned long long bal = __ballot(act);
    if (act) {
        const int slot = base + (int)__popcll(bal & ((1ull << lane) - 1ull));
        Ln[2 * slot] = make_float4(mA[0], mA[1], mA[2], __int_as_float(idx));
        Ln[2 * slot + 1] = make_float4(mB[0], mB[1], mB[2], 0.0f);
        atomicOr(&sh.nmask[idx / CS][(idx % CS) >> 5], 1u << (idx & 31));
    }
    base += (int)__popcll(bal);
}

// One contact iteration of one contact tile (P, Q) of side CS: a wavefront
// of 64 lanes, lane l owning rows RL (l / 8) .. +RL-1 and columns
// RL (l % 8) .. +RL-1 (RL = CS / 8).  The tile's pair stream is the
// previous list's pairs, then, two rows of every lane at a time, the pairs
// that collide with the circles' common N and are not in that list (lane
// order, then pair order); whenever 64 items are pending they go through
// the prox as one round.  The new list is written in stream order.  RES:
// also the residual sums (the last iteration before residuals are read).
// RE: rho differs between edges; the pairs taken through the prox read
// theirs from the dense slots (inactive pairs need none).
// Circle cc's consensus of this launch, z = Z_k, and c = Z_k - Z_{k-1}.
// Inside a run the previous launch's sums are applied here, by every
// wavefront that needs the circle (the same operations as the run's last
// update in k_contact_zreduce, so the same bits): the Z update needs no
// launch of its own between iterations.
template <bool RE>
__device__ __forceinline__ void contact_zc(const CArgs& a, int cc, float (&z)[3], float (&c)[3])
{
    if (a.Ap) {
        const float den = RE ? a.W[cc] : (float)(a.N + 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float zk = a.Zin[3 * cc + k];
            z[k] = zk + csum_f32(a.Ap[(long)k * a.N + cc]) / den;
            c[k] = z[k] - zk;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            z[k] = a.Z[3 * cc + k];
            c[k] = z[k] - a.Zp[3 * cc + k];
        }
    }
}

// The box and radius edges of the CS circles of block P, one circle per
// lane: the plain U update, prox and X, U stores of the fused iteration,
// their sums added to the circles' accumulators.  These are wavefronts of
// their own in the tile launch, so the diagonal tiles are no longer than
// the rest.
template <int CS, bool RES, bool RE>
__device__ __forceinline__ void contact_unary(const CArgs& a, int P)
{
    constexpr int Q = RES ? CQ : 3;
    const long S = a.stride, TT = (long)a.T * a.T;
    for (int e = threadIdx.x; e < CS; e += 64) {
        const int cc = P * CS + e;
        if (cc >= a.N) break;
        float z[3], cz[3];
        contact_zc<RE>(a, cc, z, cz);
        if (a.Zo)
#pragma unroll
            for (int k = 0; k < 3; ++k) a.Zo[3 * cc + k] = z[k];
        if (a.Acl)
#pragma unroll
            for (int q = 0; q < CQ; ++q) a.Acl[(long)q * a.N + cc] = 0ull;
        const long p0 = dense_unary_slot(cc, a.nb, a.T);
        long long s[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) s[q] = 0;
        float g2 = 0.0f;
        unsigned bad = 0u;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const long q = p0 + side * TT;
            const float r = RE ? a.rho[q] : a.uni[0];
            float x[3], u[3], nn[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = a.X[k * S + q];
                u[k] = (x[k] + a.U[k * S + q]) - z[k];
                nn[k] = z[k] - u[k];
            }
            if (side == 0) {
                box3(x[0], x[1], x[2], nn[0], nn[1], nn[2]);
            } else {
                x[0] = nn[0];
                x[1] = nn[1];
                x[2] = max_alpha_radius(nn[2], r, a.aexp[cc]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a.X[k * S + q] = x[k];
                a.U[k * S + q] = u[k];
                s[k] += cfix(RE ? r * ((x[k] + u[k]) - z[k]) : (x[k] + u[k]) - z[k], bad);
                if constexpr (RES) {
                    const float g = (x[k] - z[k]) - cz[k];
                    s[3 + k] += cfix(g, bad);
                    g2 += g * g;
                }
            }
            if constexpr (RES) {
                long long hi, lo;
                cfix2(g2, hi, lo, bad);
                s[6] += hi;
                s[7] += lo;
                g2 = 0.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (s[q]) atomicAdd(&a.Ac[(long)q * a.N + cc], (unsigned long long)s[q]);
        if (bad) atomicOr(a.flag, 1u);
    }
}

// Launch: nbs + nts wavefronts, the first nbs the blocks' box and radius
// edges (contact_unary), then the contact tiles.
template <int CS, bool RES, bool RE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_contact_tile(CArgs a)
{
    constexpr int Q = RES ? CQ : 3;
    constexpr int CSS = CS * CS, RL = CS / 8, W = CS / 32;
    __shared__ ContactWave<CS, Q> sh;
    const int lane = threadIdx.x;
    const int N = a.N, nbs = a.nbs;
    if ((int)blockIdx.x < nbs) {
        contact_unary<CS, RES, RE>(a, blockIdx.x);
        return;
    }
    const int t = blockIdx.x - nbs;
    int P, Qb;
    tri_tile(t, nbs, P, Qb);
    const bool diag = P == Qb;
    const long tp = (long)a.par * a.nts + t, tn = (long)(a.par ^ 1) * a.nts + t;
    const float4* Lp = a.L + tp * (2 * CSS);
    float4* Ln = a.L + tn * (2 * CSS);
    const int cp = min(a.cnt[tp], CSS);
    const float rho = a.uni[0];

    // circle table: entries 0..CS-1 the rows' circles, CS..2CS-1 the columns'
    for (int e = lane; e < 2 * CS; e += 64) {
        const int cc = min((e < CS ? P * CS + e : Qb * CS + (e - CS)), N - 1);
        float z[3], c[3];
        contact_zc<RE>(a, cc, z, c);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sh.zc[e][k] = z[k];
            sh.zc[e][3 + k] = c[k];
        }
    }
    for (int e = lane; e < CS * W; e += 64) {
        (&sh.mask[0][0])[e] = a.Lmask[tp * CS * W + e];
        (&sh.nmask[0][0])[e] = 0u;
    }
    for (int k = lane; k < 2 * CS * Q; k += 64) (&sh.acc[0][0])[k] = 0ull;
    if (lane == 0) sh.bad = 0u;
    wave_lds_fence();
    // the first 64 entries of the previous list, in flight during the tests
    float4 pre0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pre1 = pre0;
    if (lane < cp) {
        pre0 = Lp[2 * lane];
        pre1 = Lp[2 * lane + 1];
    }

    const int r0 = (lane >> 3) * RL, c0 = (lane & 7) * RL;
    const bool interior = !diag && (Qb + 1) * CS <= N;
    float nc[RL][3];
#pragma unroll
    for (int c = 0; c < RL; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) nc[c][k] = sh.zc[CS + c0 + c][k] + sh.zc[CS + c0 + c][3 + k];
    int base = 0;           // entries of the new list so far
    int done = 0;           // stream items taken through the prox
    int avail = cp;         // stream items so far (previous list + ring)
    while (avail - done >= 64) {
        contact_round<CS, Q, RE>(sh, a, P, Qb, done, avail, cp, Lp, pre0, pre1, Ln, base, diag, rho);
        done += 64;
    }
#pragma unroll 1
    for (int h = 0; h < RL / 2; ++h) {
        // the tests of rows r0 + 2h, r0 + 2h + 1
        unsigned qm = 0u;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int rl = r0 + 2 * h + r;
            float nr[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) nr[k] = sh.zc[rl][k] + sh.zc[rl][3 + k];
#pragma unroll
            for (int c = 0; c < RL; ++c) {
                const bool hit = pair_hits(nr[0], nr[1], nr[2], nc[c][0], nc[c][1], nc[c][2]);
                qm |= hit ? 1u << (RL * r + c) : 0u;
            }
            const unsigned ex = (sh.mask[rl][c0 >> 5] >> (c0 & 31)) & ((1u << RL) - 1u);
            qm &= ~(ex << (RL * r));
        }
        if (!interior) {  // pairs of the tile: not below the diagonal, not past N
            unsigned vmask = 0u;
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int c = 0; c < RL; ++c) {
                    const int rl = r0 + 2 * h + r, cl = c0 + c;
                    const bool vm = (P * CS + rl < N) && (Qb * CS + cl < N) && (!diag || rl < cl);
                    vmask |= vm ? 1u << (RL * r + c) : 0u;
                }
            qm &= vmask;
        }
        const int n = __popc(qm);
        int s = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(s, o, 64);
            s += lane >= o ? y : 0;
        }
        int off = avail - cp + s - n;
        avail += __shfl(s, 63, 64);
        unsigned m = qm;
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1u;
            sh.ring[(off++) & (CRING - 1)] = (unsigned short)((r0 + 2 * h + b / RL) * CS + c0 + b % RL);
        }
        wave_lds_fence();
        while (avail - done >= 64) {
            contact_round<CS, Q, RE>(sh, a, P, Qb, done, avail, cp, Lp, pre0, pre1, Ln, base, diag, rho);
            done += 64;
        }
    }
    if (avail > done) contact_round<CS, Q, RE>(sh, a, P, Qb, done, avail, cp, Lp, pre0, pre1, Ln, base, diag, rho);
    if (lane == 0) a.cnt[tn] = base;

    wave_lds_fence();
    for (int e = lane; e < CS * W; e += 64) a.Lmask[tn * CS * W + e] = (&sh.nmask[0][0])[e];
    // the tile's nonzero circle sums into the circles' accumulators (most
    // circles of a tile have no active pair in it)
    for (int e = lane; e < (diag ? CS : 2 * CS); e += 64) {
        const int circ = e < CS ? P * CS + e : Qb * CS + (e - CS);
        if (circ < N)
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const unsigned long long v = sh.acc[e][q];
                if (v) atomicAdd(&a.Ac[(long)q * N + circ], v);
            }
    }
    if (lane == 0 && sh.bad) atomicOr(a.flag, 1u);
}

// The end of a run: Z_{k+1} = Z_k + sum a / (N + 1) from the last launch's
// sums (RE, rho per edge: Z_k + sum rho a / sum rho) into the canonical Z,
// Z_prev, Z_prev2 (Zin = Z_k, Zpin = Z_{k-1}: canonical or inside the run's
// ring), the clearing of the sum buffers and, RES, the per-circle primal
// and dual residuals.  One thread per circle.
constexpr int CZ_THREADS = 256;

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
    constexpr int Q = RES ? CQ : 3;
    const int c = blockIdx.x * CZ_THREADS + threadIdx.x;
    const int N = a.N;
    if (c >= N) return;
    unsigned long long sa[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) sa[k] = a.Ac[(long)k * N + c];
    double s[Q];
#pragma unroll
    for (int q = 3; q < Q; ++q) s[q] = (double)(long long)a.Ac[(long)q * N + c] * (q == 7 ? CFIX2_INV : CFIX_INV);
    float zk[3], zp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zk[k] = a.Zin[3 * c + k];
        zp[k] = a.Zpin[3 * c + k];
    }
    // every sum buffer of the run is zero for the next one
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
        a.Ac[(long)q * N + c] = 0ull;
        a.Acl[(long)q * N + c] = 0ull;
    }
    const float deg = (float)(N + 1);
    double p = 0.0, dz2 = 0.0;
    if constexpr (RES) p = s[6] + s[7];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // the same operations as contact_zc's, so the same bits
        const float zn = zk[k] + csum_f32(sa[k]) / (RE ? a.W[c] : deg);
        if constexpr (RES) {
            const float dz = zn - zk[k];
            const double e = (double)(dz - (zk[k] - zp[k]));
            p += -2.0 * e * s[3 + k] + (double)deg * e * e;
            dz2 += (double)dz * dz;
        }
        a.Zpp[3 * c + k] = zp[k];
        a.Zp[3 * c + k] = zk[k];
        a.Z[3 * c + k] = zn;
    }
    if constexpr (RES) {
        const float rho = a.uni[0];
        // S2 - 2 e.S1 + (N + 1)|e|^2 >= 0; the roundings of the two sides can
        // leave a tiny negative
        a.Rpc[c] = p > 0.0 ? p : 0.0;
        a.Rd[c] = RE ? (double)a.W[a.N + c] * dz2 : (double)rho * rho * (double)deg * dz2;
    }
}

// Enters the contact state from the steady X, U state (U one update behind,
// Z = Z_k, Z_prev = Z_{k-1}): list 0 of contact tile blockIdx.x receives
// every pair with M = X + U != Z_{k-1} on either edge, exactly, so the first
// contact iteration computes the same U = M - Z_k as the fused one.  256
// threads; thread: column tid % CS of rows tid / CS, + 256 / CS, ... (a
// wavefront reads consecutive slots of one row of the dense tiles).
template <int CS>
__global__ __launch_bounds__(256) void k_contact_enter(CArgs a)
{
    constexpr int CSS = CS * CS, W = CS / 32, RS = 256 / CS, CPT = CSS / 256;
    __shared__ int wsum[4];
    __shared__ unsigned emask[CS][W];
    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    const int N = a.N;
    const long S = a.stride, TT = (long)a.T * a.T;
    int P, Qb;
    tri_tile(t, a.nbs, P, Qb);
    if (tid < CS * W) (&emask[0][0])[tid] = 0u;
    __syncthreads();
    float4* L0 = a.L + (long)t * (2 * CSS);
    const int jl = tid % CS, r0 = tid / CS;
    const int j = Qb * CS + jl;
    const int jc = min(j, N - 1);
    float zpj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) zpj[k] = a.Zp[3 * jc + k];
    unsigned exm = 0u;
#pragma unroll 4
    for (int c = 0; c < CPT; ++c) {
        const int il = r0 + c * RS, i = P * CS + il;
        const bool vm = (i < N) && (j < N) && (P != Qb || il < jl);
        if (!vm) continue;
        const long pa = dense_pair_slot(i, j, a.nb, a.T), pb = pa + TT;
        bool ex = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float mA = a.X[k * S + pa] + a.U[k * S + pa];
            const float mB = a.X[k * S + pb] + a.U[k * S + pb];
            ex = ex || mA != a.Zp[3 * i + k] || mB != zpj[k];
        }
        if (ex) {
            exm |= 1u << c;
            atomicOr(&emask[il][jl >> 5], 1u << (jl & 31));
        }
    }
    int total = 0;
    int slot = block_offset(__popc(exm), wsum, total);
    while (exm) {
        const int c = __ffs(exm) - 1;
        exm &= exm - 1u;
        const int il = r0 + c * RS, i = P * CS + il;
        const long pa = dense_pair_slot(i, j, a.nb, a.T), pb = pa + TT;
        float mA[3], mB[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mA[k] = a.X[k * S + pa] + a.U[k * S + pa];
            mB[k] = a.X[k * S + pb] + a.U[k * S + pb];
        }
        L0[2 * slot] = make_float4(mA[0], mA[1], mA[2], __int_as_float(il * CS + jl));
        L0[2 * slot + 1] = make_float4(mB[0], mB[1], mB[2], 0.0f);
        ++slot;
    }
    if (tid == 0) a.cnt[t] = total;
    if (tid < CS * W) a.Lmask[(long)t * CS * W + tid] = (&emask[0][0])[tid];
}

// Leaves the contact state after iteration k (Z_prev = Z_k, Z_prev2 =
// Z_{k-1}, list par = L_k, the other = L_{k-1}): every pair edge gets
//   U_k = M_{k-1} - Z_k,  X_k = M_k - U_k
// with M = Z of its circle off the lists: the steady X, U state.  256
// threads per contact tile, as k_contact_enter.
template <int CS>
__global__ __launch_bounds__(256) void k_contact_leave(CArgs a)
{
    constexpr int CSS = CS * CS, W = CS / 32, RS = 256 / CS, CPT = CSS / 256;
    __shared__ unsigned short map[2][CSS];
    __shared__ unsigned mask[2][CS][W];
    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    const int N = a.N;
    const long S = a.stride, TT = (long)a.T * a.T;
    int P, Qb;
    tri_tile(t, a.nbs, P, Qb);
    const long lt[2] = {(long)a.par * a.nts + t, (long)(a.par ^ 1) * a.nts + t};
    const float4* Ls[2] = {a.L + lt[0] * (2 * CSS), a.L + lt[1] * (2 * CSS)};
    const int cn[2] = {min(a.cnt[lt[0]], CSS), min(a.cnt[lt[1]], CSS)};
    if (tid < 2 * CS * W) (&mask[0][0][0])[tid] = 0u;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 2; ++b)
        for (int e = tid; e < cn[b]; e += 256) {
            const int idx = __float_as_int(Ls[b][2 * e].w) & (CSS - 1);
            map[b][idx] = (unsigned short)e;
            atomicOr(&mask[b][idx / CS][(idx % CS) >> 5], 1u << (idx & 31));
        }
    __syncthreads();
    const int jl = tid % CS, r0 = tid / CS;
    const int j = Qb * CS + jl;
#pragma unroll 2
    for (int c = 0; c < CPT; ++c) {
        const int il = r0 + c * RS, i = P * CS + il;
        const bool vm = (i < N) && (j < N) && (P != Qb || il < jl);
        if (!vm) continue;
        const int ic = i;
        const int idx = il * CS + jl;
        const bool in_k = (mask[0][il][jl >> 5] >> (jl & 31)) & 1u;
        const bool in_k1 = (mask[1][il][jl >> 5] >> (jl & 31)) & 1u;
        float4 ek[2] = {}, ek1[2] = {};
        if (in_k) {
            ek[0] = Ls[0][2 * map[0][idx]];
            ek[1] = Ls[0][2 * map[0][idx] + 1];
        }
        if (in_k1) {
            ek1[0] = Ls[1][2 * map[1][idx]];
            ek1[1] = Ls[1][2 * map[1][idx] + 1];
        }
        const long pa = dense_pair_slot(i, j, a.nb, a.T), pb = pa + TT;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int v = side == 0 ? ic : j;
            const long q = side == 0 ? pa : pb;
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
//
// paradmm_gpu_solve applies the rho policy (modes 2 and 3: one rho per graph
// from the mean of the network's log10 rho over all edges, in mode 3 plus the
// edges whose own output exceeds that mean by more than raise_log) without
// leaving the contact state.  At a policy point (the end of a run: lists par
// = M_j of the last iteration and other = M_{j-1}, Z = Z_j, Z_prev = Z_{j-1},
// Z_prev2 = Z_{j-2}) an edge's policy inputs are those of the steady state
// k_contact_leave would materialise: X_j = M_j - U_{j-1} with
// U_{j-1} = M_{j-1} - Z_{j-1}, where M = Z of the circle off the lists.
// A pair edge in neither list has X_j = Z_{j-1} - (Z_{j-2} - Z_{j-1}) and, if
// its rho is its circle's base rho (the rho every unlisted pair edge of the
// circle took at the last application), exactly the inputs of every other
// such edge of the circle: the network runs once per circle for them, and
// once per edge for the listed pairs, the pairs of either list, the pairs
// off their circles' base rho, and the box and radius edges.
//
// The scaled dual U_j = M_j - Z_j is rescaled by rho_old / rho_new: a listed
// pair's M becomes Z + s (M - Z); the box and radius edges' stored U (one
// update behind) becomes s (X + U - Z) - X + Z; an unlisted pair edge's
// U_j = Z_{j-1} - Z_j is implicit in its circle's Z_{j-1}, so the circle's
// Z_prev becomes Z - s (Z - Z_prev), a virtual Z_{j-1} that the next contact
// iteration reads for every unlisted pair (and consistently in its residual
// terms).  An unlisted pair whose edges' factors are not their circles'
// joins the list with its rescaled M.  The virtual Z_prev is valid only for
// the next contact iteration: the solve applies the policy only where more
// iterations follow, and no state is left in between.

constexpr float CPOL_FIX = 268435456.0f;  // network outputs summed in units of 2^-28
constexpr float CPOL_WFIX = 16777216.0f;  // sums of rho in units of 2^-24
constexpr float CPOL_W2FIX = 4096.0f;     // sums of rho^2 in units of 2^-12

struct CPolArgs {
    CArgs c;
    const float* params;   // the network's parameters
    float frac;            // iteration fraction (network input)
    float raise_log;       // mode 3: an edge keeps its own rho above graph rho * 10^raise_log; +inf in mode 2
    int exceptions;        // rho may differ from a circle's base rho on unlisted pairs (per-edge rho)
    int average;           // running mean of the estimates
    float* base;           // [N] base rho of each circle's unlisted pair edges
    float* yc;             // [3][N] outputs: unlisted pair edges, box edge, radius edge
    int* rows;             // [N] pair edges of each circle evaluated on their own
    long long* wacc;       // [2][N] per circle, over its pair edges evaluated on their own: sum of
                           // (rho_new - circle's new unlisted rho) (2^-24) and of the same for rho^2 (2^-12)
    long long* ysum;       // [1] sum of the outputs of the pair edges evaluated on their own (units of 2^-28)
    double* hist;          // [6] running sums of the per-type means (shared with the per-edge policy)
    float* gl;             // [2] the new graph rho and its log10
};

__device__ __forceinline__ long long cpol_fix(float y) { return (long long)rintf(y * CPOL_FIX); }

// the new rho of an edge whose network output is y: its own where it exceeds
// the graph rho by more than the raise factor, else the graph rho
__device__ __forceinline__ float cpol_rho(const CPolArgs& p, float y, float g, float lg)
{
    return y - lg > p.raise_log ? exp10f(y) : g;
}

// policy inputs of an edge from its X, its circle's Z and Z_prev and its rho
__device__ __forceinline__ void cpol_features(const float (&x)[3], const float* z, const float* zp, float rho, int type,
                                              float frac, float (&f)[MLP_NIN])
{
    // no contraction: the per-edge application (mlp.hip) computes the same
    // inputs bit for bit
#pragma clang fp contract(off)
    float r2 = 0.0f, dz2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = x[k] - z[k];
        const float q = z[k] - zp[k];
        r2 += d * d;
        dz2 += q * q;
    }
    edge_features(sqrtf(r2), rho * sqrtf(dz2), rho, type, frac, f);
}

// Appends the items a workgroup's threads take (in thread order) to q at n;
// returns the new count (every thread).  Two __syncthreads.
template <int CSS>
__device__ __forceinline__ int cpol_append(unsigned short (&q)[CSS], int n, bool take, unsigned short code,
                                           int (&wsum)[4])
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long b = __ballot(take);
    if (lane == 0) wsum[wv] = (int)__popcll(b);
    __syncthreads();
    int base = n, total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        base += k < wv ? wsum[k] : 0;
        total += wsum[k];
    }
    const int at = base + (int)__popcll(b & ((1ull << lane) - 1ull));
    if (take && at < CSS) q[at] = code;
    __syncthreads();
    return min(n + total, CSS);
}

// One contact tile, four wavefronts: the pairs evaluated on their own, in a
// fixed order (the par list, then the pairs of the other list only, then
// with per-edge rho the unlisted pairs off their circles' base rho), 32
// pairs (64 edges) per wavefront and network step.  !APPLY: their outputs'
// sum and each circle's count of them; a diagonal tile also evaluates its
// circles' unlisted pair edges, box and radius edges.  APPLY: each edge's
// new rho and the rescaled M (in place for the par list, appended for an
// unlisted pair whose factors are not its circles'), the circles' rho sums
// corrected for them; then every other pair of the tile takes its circles'
// new rho (after the evaluated pairs have read their old one), and a
// diagonal tile also rescales its circles' box and radius edges.
template <int CS, bool APPLY>
__global__ __launch_bounds__(256) void k_cpol_tile(CPolArgs p)
{
    constexpr int CSS = CS * CS, W = CS / 32;
    __shared__ MlpShared w;
    __shared__ TileScratch ts[4];
    __shared__ float yv[4][64];
    __shared__ float zc[2 * CS][9];       // Z, Z_prev, Z_prev2 of the tile's circles
    __shared__ unsigned short omap[CSS];  // other list: pair -> entry
    __shared__ unsigned omask[CS][W], pmask[CS][W], rmask[CS][W];
    __shared__ unsigned short q[CSS];     // the tile's pairs to evaluate: kind << 14 | entry or pair
    __shared__ int wsum[4];
    const CArgs& a = p.c;
    float* const rhow = const_cast<float*>(a.rho);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = blockIdx.x;
    const int N = a.N;
    const long TT = (long)a.T * a.T;
    int P, Qb;
    tri_tile(t, a.nbs, P, Qb);
    const bool diag = P == Qb;
    const long tp = (long)a.par * a.nts + t, to = (long)(a.par ^ 1) * a.nts + t;
    float4* Lp = a.L + tp * (2 * CSS);
    const float4* Lo = a.L + to * (2 * CSS);
    const int cp = min(a.cnt[tp], CSS), co = min(a.cnt[to], CSS);
    stage_weights(p.params, w);
    for (int e = tid; e < 2 * CS; e += 256) {
        const int cc = min(e < CS ? P * CS + e : Qb * CS + (e - CS), N - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zc[e][k] = a.Z[3 * cc + k];
            zc[e][3 + k] = a.Zp[3 * cc + k];
            zc[e][6 + k] = a.Zpp[3 * cc + k];
        }
    }
    for (int e = tid; e < CS * W; e += 256) {
        (&pmask[0][0])[e] = a.Lmask[tp * CS * W + e];
        (&omask[0][0])[e] = a.Lmask[to * CS * W + e];
        if (APPLY) (&rmask[0][0])[e] = 0u;
    }
    for (int e = tid; e < co; e += 256) omap[__float_as_int(Lo[2 * e].w) & (CSS - 1)] = (unsigned short)e;
    for (int e = tid; e < cp; e += 256) q[e] = (unsigned short)e;
    __syncthreads();
    auto inmask = [&](const unsigned (&m)[CS][W], int idx) { return (m[idx / CS][(idx % CS) >> 5] >> (idx & 31)) & 1u; };
    // the queue: par entries, other-only entries, off-base unlisted pairs
    int n = cp;
    for (int e0 = 0; e0 < co; e0 += 256) {
        const int e = e0 + tid;
        const bool take = e < co && !inmask(pmask, __float_as_int(Lo[2 * e].w) & (CSS - 1));
        n = cpol_append(q, n, take, (unsigned short)((1u << 14) | (unsigned)e), wsum);
    }
    if (p.exceptions) {
        for (int i0 = 0; i0 < CSS; i0 += 256) {
            const int idx = i0 + tid;
            const int il = idx / CS, jl = idx % CS;
            const int ci = P * CS + il, cj = Qb * CS + jl;
            bool take = false;
            if (ci < N && cj < N && (!diag || il < jl) && !inmask(pmask, idx) && !inmask(omask, idx)) {
                const long qa = dense_pair_slot(ci, cj, a.nb, a.T);
                take = a.rho[qa] != p.base[ci] || a.rho[qa + TT] != p.base[cj];
            }
            n = cpol_append(q, n, take, (unsigned short)((2u << 14) | (unsigned)idx), wsum);
        }
    }
    if (APPLY) {
        for (int e = tid; e < n; e += 256) {
            const unsigned code = q[e];
            const int k2 = (int)(code >> 14), en = (int)(code & 0x3fffu);
            const int idx = k2 == 2 ? en : __float_as_int((k2 == 0 ? Lp : Lo)[2 * en].w) & (CSS - 1);
            atomicOr(&rmask[idx / CS][(idx % CS) >> 5], 1u << (idx & 31));
        }
        __syncthreads();
    }
    const float g = APPLY ? p.gl[0] : 0.0f, lg = APPLY ? p.gl[1] : 0.0f;
    long long acc = 0;
    for (int q0 = 0; q0 < n; q0 += 128) {
        const int qi = q0 + 32 * wv + (lane & 31);
        const int side = lane >> 5;  // 0: the pair's first circle (row), 1: its second (column)
        const bool valid = qi < n;
        float f[MLP_NIN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        int kind = 0, ent = 0, idx = 0, zr = 0, circ = 0;
        long slot = 0;
        float m[3] = {0.0f, 0.0f, 0.0f}, rho = 1.0f;
        if (valid) {
            const unsigned code = q[qi];
            kind = (int)(code >> 14);
            ent = (int)(code & 0x3fffu);
            idx = kind == 2 ? ent : __float_as_int((kind == 0 ? Lp : Lo)[2 * ent].w) & (CSS - 1);
            const int il = idx / CS, jl = idx % CS;
            zr = side ? CS + jl : il;
            circ = side ? Qb * CS + jl : P * CS + il;
            slot = dense_pair_slot(P * CS + il, Qb * CS + jl, a.nb, a.T) + (side ? TT : 0);
            rho = a.rho[slot];
            // M_j: the par entry, else Z_{j-1}; M_{j-1}: the other entry, else Z_{j-2}
            float mo[3];
            const bool in_o = kind == 1 || (kind == 0 && inmask(omask, idx));
            const float4 eo = in_o ? Lo[2 * (kind == 1 ? ent : omap[idx]) + side] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 ep = kind == 0 ? Lp[2 * ent + side] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float eov[3] = {eo.x, eo.y, eo.z}, epv[3] = {ep.x, ep.y, ep.z};
            float x[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mo[k] = in_o ? eov[k] : zc[zr][6 + k];
                m[k] = kind == 0 ? epv[k] : zc[zr][3 + k];
                const float u = mo[k] - zc[zr][3 + k];
                x[k] = m[k] - u;
            }
            cpol_features(x, &zc[zr][0], &zc[zr][3], rho, PDM_EDGE_PAIR, p.frac, f);
        }
        const float y = policy_wave_eval(w, ts[wv], yv[wv], f, lane);
        if constexpr (!APPLY) {
            if (valid) {
                acc += cpol_fix(y);
                atomicAdd(&p.rows[circ], 1);
            }
        } else {
            float rn = 0.0f, sf = 0.0f, rv = 0.0f;
            if (valid) {
                rn = cpol_rho(p, y, g, lg);
                rv = cpol_rho(p, p.yc[circ], g, lg);  // the circle's new unlisted rho
                sf = rho / rn;
            }
            // an unlisted pair stays implicit if both edges take their circles' factors
            const bool own = valid && (rho != p.base[circ] || rn != rv);
            const bool both_own = own || __shfl_xor(own ? 1 : 0, 32, 64);
            float mn[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) mn[k] = zc[zr][k] + sf * (m[k] - zc[zr][k]);
            if (valid) {
                rhow[slot] = rn;  // every evaluated edge, implicit or not
                const long long dw = (long long)rintf((rn - rv) * CPOL_WFIX);
                const long long dw2 = (long long)rintf((rn * rn - rv * rv) * CPOL_W2FIX);
                if (dw) atomicAdd((unsigned long long*)&p.wacc[circ], (unsigned long long)dw);
                if (dw2) atomicAdd((unsigned long long*)&p.wacc[N + circ], (unsigned long long)dw2);
            }
            if (valid && kind == 0) Lp[2 * ent + side] = make_float4(mn[0], mn[1], mn[2], side ? 0.0f : __int_as_float(idx));
            // the appended entries: lane of side 0 reserves the slot
            const bool app = valid && kind != 0 && both_own;
            int at = 0;
            if (app && side == 0) at = atomicAdd(&a.cnt[tp], 1);
            at = __shfl(at, lane & 31, 64);
            if (app && at >= CSS) {  // cannot happen for distinct pairs: the state is lost
                if (side == 0) atomicOr(a.flag, 1u);
            } else if (app) {
                Lp[2 * at + side] = make_float4(mn[0], mn[1], mn[2], side ? 0.0f : __int_as_float(idx));
                if (side == 0) atomicOr(&a.Lmask[tp * CS * W + (idx / CS) * W + ((idx % CS) >> 5)], 1u << (idx & 31));
            }
        }
    }
    if constexpr (!APPLY) {
        // a diagonal tile: its circles' unlisted pair edges (wave 0), box
        // edges (wave 1) and radius edges (wave 2)
        if (diag && wv < 3) {
            const int kind = wv;
            const int c = P * CS + lane;
            const bool valid = lane < CS && c < N;
            float f[MLP_NIN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (valid) {
                float x[3], rho;
                if (kind == 0) {  // an unlisted pair edge: X_j = Z_{j-1} - (Z_{j-2} - Z_{j-1})
#pragma unroll
                    for (int k = 0; k < 3; ++k) x[k] = zc[lane][3 + k] - (zc[lane][6 + k] - zc[lane][3 + k]);
                    rho = p.base[c];
                } else {
                    const long qq = dense_unary_slot(c, a.nb, a.T) + (kind == 2 ? TT : 0);
#pragma unroll
                    for (int k = 0; k < 3; ++k) x[k] = a.X[k * a.stride + qq];
                    rho = a.rho[qq];
                }
                cpol_features(x, &zc[lane][0], &zc[lane][3], rho,
                              kind == 0 ? PDM_EDGE_PAIR : kind == 1 ? PDM_EDGE_CONTAINER : PDM_EDGE_RADIUS, p.frac, f);
            }
            const float y = policy_wave_eval(w, ts[wv], yv[wv], f, lane);
            if (valid) p.yc[(long)kind * N + c] = y;
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0 && acc) atomicAdd((unsigned long long*)p.ysum, (unsigned long long)acc);
    } else {
        // every pair not evaluated on its own: its circles' new rho
        for (int idx = tid; idx < CSS; idx += 256) {
            const int il = idx / CS, jl = idx % CS;
            const int ci = P * CS + il, cj = Qb * CS + jl;
            if (ci >= N || cj >= N || (diag && il >= jl) || inmask(rmask, idx)) continue;
            const long qa = dense_pair_slot(ci, cj, a.nb, a.T);
            rhow[qa] = cpol_rho(p, p.yc[ci], g, lg);
            rhow[qa + TT] = cpol_rho(p, p.yc[cj], g, lg);
        }
        // a diagonal tile: its circles' box and radius edges, their stored U
        // (one update behind) rescaled so the pending update is s U_j
        if (diag)
            for (int e = tid; e < 2 * CS; e += 256) {
                const int c = P * CS + (e % CS), kind = 1 + e / CS;
                if (c >= N) continue;
                const long qq = dense_unary_slot(c, a.nb, a.T) + (kind == 2 ? TT : 0);
                const float r = cpol_rho(p, p.yc[(long)kind * N + c], g, lg);
                const float sf = a.rho[qq] / r;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float x = a.X[k * a.stride + qq], z = a.Z[3 * c + k];
                    a.U[k * a.stride + qq] = sf * ((x + a.U[k * a.stride + qq]) - z) - x + z;
                }
                rhow[qq] = r;
            }
    }
}

// The graph rho: 10^(mean output over all edges) -- the pair edges evaluated
// on their own (ysum), (N - 1 - rows) times each circle's unlisted pair
// output, its box and radius outputs -- or with average the mean of this and
// the earlier estimates (the per-type policy's running sums, every type the
// same).  One workgroup; integer sums, so the order does not matter.
__global__ __launch_bounds__(256) void k_cpol_final(CPolArgs p, long E)
{
    __shared__ long long sh[256];
    const int N = p.c.N;
    long long s = 0;
    for (int c = threadIdx.x; c < N; c += 256)
        s += (long long)(N - 1 - p.rows[c]) * cpol_fix(p.yc[c]) + cpol_fix(p.yc[(long)N + c]) +
             cpol_fix(p.yc[2L * N + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double m = (double)(sh[0] + *p.ysum) / (double)CPOL_FIX / (double)E;
    if (p.average) {
        for (int k = 0; k < 3; ++k) {
            p.hist[2 * k] += m;
            p.hist[2 * k + 1] += 1.0;
        }
        m = p.hist[0] / p.hist[1];
    }
    const float g = exp10f((float)m);
    p.gl[0] = g;
    p.gl[1] = log10f(g);
}

// Per circle, last: the virtual Z_{j-1} that carries the unlisted pair
// edges' rescaled U, the base rho for the next application, and the
// circle's sums of rho and rho^2 over its N + 1 edges for the per-edge-rho
// consensus: (N - 1) times the unlisted rho, corrected for the pair edges
// evaluated on their own, plus the box and radius edges.
__global__ __launch_bounds__(256) void k_cpol_zprev(CPolArgs p)
{
    const CArgs& a = p.c;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int N = a.N;
    if (c >= N) return;
    const float g = p.gl[0], lg = p.gl[1];
    const float rv = cpol_rho(p, p.yc[c], g, lg);
    const float sf = p.base[c] / rv;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float z = a.Z[3 * c + k];
        a.Zp[3 * c + k] = z - sf * (z - a.Zp[3 * c + k]);
    }
    p.base[c] = rv;
    const long TT = (long)a.T * a.T, q0 = dense_unary_slot(c, a.nb, a.T);
    const float rb = a.rho[q0], rr = a.rho[q0 + TT];
    const double w = (double)(N - 1) * rv + (double)p.wacc[c] / (double)CPOL_WFIX + rb + rr;
    const double w2 = (double)(N - 1) * ((double)rv * rv) + (double)p.wacc[N + c] / (double)CPOL_W2FIX +
                      (double)rb * rb + (double)rr * rr;
    a.W[c] = (float)w;
    a.W[N + c] = (float)w2;
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

static int contact_alloc(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d->d_L) return PARADMM_OK;
    d->cs = contact_tile_side(d->N);
    d->nbs = (d->N + d->cs - 1) / d->cs;
    d->nts = d->nbs * (d->nbs + 1) / 2;
    const size_t nts = (size_t)d->nts, css = (size_t)d->cs * d->cs, words = (size_t)d->cs * (d->cs / 32);
    RC(dmalloc(s, (void**)&d->d_L, sizeof(float4) * 2 * 2 * nts * css));
    RC(dmalloc(s, (void**)&d->d_cnt, sizeof(int) * 2 * nts));
    RC(dmalloc(s, (void**)&d->d_Lmask, sizeof(unsigned) * 2 * nts * words));
    CK(hipMemsetAsync(d->d_cnt, 0, sizeof(int) * 2 * nts, s->stream));
    CK(hipMemsetAsync(d->d_Lmask, 0, sizeof(unsigned) * 2 * nts * words, s->stream));
    RC(dmalloc(s, (void**)&d->d_Ac, sizeof(unsigned long long) * 3 * CQ * (size_t)d->N));
    CK(hipMemsetAsync(d->d_Ac, 0, sizeof(unsigned long long) * 3 * CQ * (size_t)d->N, s->stream));
    RC(dmalloc(s, (void**)&d->d_Zb, sizeof(float) * 2 * 3 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Rpc, sizeof(double) * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Zpp, sizeof(float) * 3 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_W, sizeof(float) * 2 * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_flag, sizeof(unsigned)));
    RC(dmalloc(s, (void**)&d->d_base, sizeof(float) * (size_t)d->N));
    return PARADMM_OK;
}

// res: also the residual sums (the last iteration of a run)
template <typename K>
static void contact_launch(DevGraph* s, const char* name, K kernel, unsigned grid, unsigned block, const CArgs& a)
{
    hipEvent_t e0, e1;
    prof_dispatch(s, name, e0, e1);
    if (e0)  // timed, eager
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, e0, e1, 0, a);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, a);
}

// Launch j of a run (j = c_run launches since the last run end) applies the
// sums of launch j - 1 to Z as it reads it and adds its own to buffer j % 3
// of three; its box-and-radius wavefronts store its Z in ring slot j % 2 and
// clear buffer (j + 1) % 3.  res (the last iteration of every run, and every
// single iteration): also the residual sums, then k_contact_zreduce ends the
// run into the canonical Z arrays, so outside a run the state is as it
// always was.
static int contact_iteration(DevGraph* s, bool res)
{
    DenseState* d = s->dense;
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
    RC(contact_alloc(s));
    d->c_par = 0;  // the entry writes list 0
    d->c_run = 0;
    // a fresh state: no sum has lost a value (a fill kernel, not a memset,
    // as the entry may run inside a captured stream)
    k_fill_u32<<<1, 64, 0, s->stream>>>(d->d_flag, 1, 0u);
    CK(hipGetLastError());
    // one rho everywhere (a policy's raised edges make it per edge)
    d->c_rhoe = !(s->rho_uni && s->policy_raised == 0);
    // with one rho it is every circle's base rho (learned rho in the state)
    d->c_base_ok = !d->c_rhoe;
    if (d->c_base_ok) {
        k_fill_f32<<<(unsigned)((d->N + 255) / 256), 256, 0, s->stream>>>(d->d_base, d->N, d->d_uni);
        CK(hipGetLastError());
    }
    const CArgs a = cargs(s);
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

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
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "paradmm_gpu.h"
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
    float4* d_L = nullptr;     // [2][nt][T*T][2] contact lists
    int* d_cnt = nullptr;      // [2][nt] list lengths
    int* d_cur = nullptr;      // [1] the list the last iteration wrote
    float* d_Pc = nullptr;     // [CQ][nb][N] circle partials of the contact iteration
    double* d_Rpc = nullptr;   // [N] primal residual per circle
    float* d_Zpp = nullptr;    // [N][3] Z two iterations back
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
__global__ __launch_bounds__(256) void k_dense_res_final(int nt, const double* __restrict__ Rp, int N,
                                                         const double* __restrict__ Rd, double* __restrict__ out)
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
        out[0] = sp[0];
        out[1] = sq[0];
    }
}

// ------------------------------------------------------------------ contact state
//
// With alpha = 1 the U update is U_{k+1} = M_k - Z_{k+1}.  A pair factor
// whose prox is inactive returns X = N = Z_k - U_k, so M_k = X + U = Z_k:
// its edges carry no state of their own, and at the next iteration
// U = Z_k - Z_{k+1}, N = Z_{k+1} - U, the same for every inactive pair of a
// circle.  So only the pairs whose prox was active (the contacts, about 1%
// of all pairs for N = 5000 circles) need per-edge values: a tile keeps a
// list of them, each entry the M of both edges and the pair's slot, and an
// iteration reads that list, tests every pair of the tile against the
// circles' common N, runs the prox on the colliding ones and writes the new
// list.  The box and radius edges stay in their X and U slots.
//
// Per circle the tile sums, over its pair edges that are in either list,
//   a = M - Z_k                    (consensus: Z_{k+1} = Z_k + sum a / (N + 1)),
//   g = (X - Z_k) - (Z_k - Z_{k-1})  and |g|^2,
// where a and g are 0 on every other edge.  The primal residual
// sum |X - Z_{k+1}|^2 is then, per circle, |g - e|^2 summed over all N + 1
// edges with e = Z_{k+1} - 2 Z_k + Z_{k-1}: S2 - 2 e.S1 + (N + 1)|e|^2.
// The sums run in 2^-31 fixed point (64-bit LDS atomics), so they do not
// depend on the order the lanes add them: runs are bitwise reproducible.
//
// The per-pair test replaces 24 bytes of M per edge read and written by
// ~10 VALU operations; the state is Z, Z_prev, Z_prev2 and the lists.

constexpr int CT = 32;        // tile side of the contact kernels
constexpr int CTT = CT * CT;
constexpr int CQ = 7;         // per-circle sums: a (3), g (3), |g|^2
constexpr double CFIX_INV = 1.0 / 2147483648.0;  // 2^-31

struct CArgs {
    int N, nb, nt;
    long stride;
    const int2* tiles;     // launch order
    float* X;
    float* U;              // box and radius slots of the diagonal tiles
    float* Z;
    float* Zp;
    float* Zpp;
    const float* aexp;
    const float* uni;      // {rho, alpha}
    int* cur;
    float4* L;             // [2][nt][CTT][2]
    int* cnt;              // [2][nt]
    float* Pc;             // [CQ][nb][N]
    double* Rpc;
    double* Rd;
};

// acc += v in 2^-31 fixed point: v = k + f with k = rint(v) and f = v - k
// exact (|f| <= 1/2), so v 2^31 = k 2^31 + f 2^31 needs no FP64 (|v| < 2^31)
__device__ __forceinline__ void cadd(unsigned long long* acc, float v)
{
    const float k = rintf(v);
    const float f = v - k;
    const long long q = (long long)(int)k * 2147483648LL + (long long)(int)(f * 2147483648.0f);
    atomicAdd(acc, (unsigned long long)q);
}

// Exclusive position of this lane's n items among the workgroup's (lanes in
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

// Shared scratch of one wavefront's tile.  Q circle sums: a (3), and with
// the residual terms g (3), |g|^2.
template <int Q>
struct ContactWave {
    unsigned long long acc[2 * CT][Q];  // fixed-point circle sums: rows (block I), then columns (block J)
    float zc[2 * CT][6];                // Z_k and c = Z_k - Z_{k-1} of the tile's 64 circles
    unsigned short queue[CTT];          // colliding pairs not in the previous list (slot)
    unsigned mask[CT];                  // pairs in the previous list, one word per row
};

// (I, J) of layout tile t: row I starts at I nb - I (I - 1) / 2 (the
// float estimate is corrected by the loops; nb < 1449 for any dense layout,
// so every product fits in 32 bits)
__device__ __forceinline__ void tile_of(int t, int nb, int& I, int& J)
{
    const float b = 2.0f * nb + 1.0f;
    I = (int)((b - sqrtf(fmaxf(b * b - 8.0f * (float)t, 0.0f))) * 0.5f);
    I = max(0, min(I, nb - 1));
    while (I > 0 && I * nb - I * (I - 1) / 2 > t) --I;
    while (I + 1 < nb && (I + 1) * nb - (I + 1) * I / 2 <= t) ++I;
    J = I + (t - (I * nb - I * (I - 1) / 2));
}

// One contact iteration: one wavefront per layout tile, four tiles per
// workgroup, no workgroup barrier.
//   pass 1: lane l tests the 16 pairs of rows 4 (l / 8) .. +3 and columns
//     4 (l % 8) .. +3 with the circles' common N; the colliding ones not in
//     the previous list go to a queue (lane order, then pair order);
//   pass 2: the previous list (its order), then the queue, one pair per
//     lane: its N (Z - U from its list entry, or the common one), the
//     prox, the new list (in that order) and the circle sums.
// RES: also the residual sums (the last iteration before residuals are read).
template <bool RES>
__global__ __launch_bounds__(256) void k_contact_tile(CArgs a)
{
    constexpr int Q = RES ? CQ : 3;
    __shared__ ContactWave<Q> shw[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + w;
    if (t >= a.nt) return;  // whole wavefronts only
    ContactWave<Q>& sh = shw[w];
    const int N = a.N, nb = a.nb;
    int I, J;
    tile_of(t, nb, I, J);
    const int par = *a.cur;
    const long tp = (long)par * a.nt + t;
    const float4* Lp = a.L + tp * (2 * CTT);
    float4* Ln = a.L + ((long)(par ^ 1) * a.nt + t) * (2 * CTT);
    const int cp = a.cnt[tp];
    const float rho = a.uni[0];

    {   // circle table: lanes 0..31 the rows' circles, 32..63 the columns'
        const int cc = min(lane < CT ? I * CT + lane : J * CT + (lane - CT), N - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float z = a.Z[3 * cc + k];
            sh.zc[lane][k] = z;
            sh.zc[lane][3 + k] = z - a.Zp[3 * cc + k];
        }
    }
    for (int k = lane; k < 2 * CT * Q; k += 64) (&sh.acc[0][0])[k] = 0ull;
    if (lane < CT) sh.mask[lane] = 0u;
    wave_sync();
    for (int e = lane; e < cp; e += 64) {
        const int idx = __float_as_int(Lp[2 * e].w);
        atomicOr(&sh.mask[idx >> 5], 1u << (idx & 31));
    }
    wave_sync();

    // pass 1
    const int r0 = (lane >> 3) * 4, c0 = (lane & 7) * 4;
    // N of a circle's pairs off the lists: Z_k - (Z_{k-1} - Z_k) = Z_k + c
    float nr[4][3], nc[4][3];
    unsigned exb = 0u;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            nr[r][k] = sh.zc[r0 + r][k] + sh.zc[r0 + r][3 + k];
            nc[r][k] = sh.zc[CT + c0 + r][k] + sh.zc[CT + c0 + r][3 + k];
        }
        exb |= ((sh.mask[r0 + r] >> c0) & 0xFu) << (4 * r);
    }
    if (cp == 0) exb = 0u;
    const bool interior = I != J && (J + 1) * CT <= N;
    unsigned qm = 0u;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool hit = pair_hits(nr[r][0], nr[r][1], nr[r][2], nc[c][0], nc[c][1], nc[c][2]);
            qm |= hit ? 1u << (4 * r + c) : 0u;
        }
    qm &= ~exb;
    if (!interior) {  // pairs of the tile: below the diagonal and past N are not
        unsigned vmask = 0u;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int rl = r0 + r, cl = c0 + c;
                const bool vm = (I * CT + rl < N) && (J * CT + cl < N) && (I != J || rl < cl);
                vmask |= vm ? 1u << (4 * r + c) : 0u;
            }
        qm &= vmask;
    }
    int qn;
    {
        const int n = __popc(qm);
        int s = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(s, o, 64);
            s += lane >= o ? y : 0;
        }
        qn = __shfl(s, 63, 64);
        int off = s - n;
        unsigned m = qm;
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1u;
            sh.queue[off++] = (unsigned short)((r0 + (b >> 2)) * CT + c0 + (b & 3));
        }
    }
    wave_sync();

    // pass 2
    int base = 0;
    const int total = cp + qn;
    for (int q0 = 0; q0 < total; q0 += 64) {
        const int q = q0 + lane;
        bool act = false;
        int idx = 0;
        float mA[3], mB[3];
        if (q < total) {
            const bool ex = q < cp;
            float4 pa, pb;
            if (ex) {
                pa = Lp[2 * q];
                pb = Lp[2 * q + 1];
                idx = __float_as_int(pa.w);
            } else {
                idx = sh.queue[q - cp];
            }
            const int rl = idx >> 5, cl = idx & 31;
            const int cs = I == J ? cl : CT + cl;
            float zA[3], zB[3], cA[3], cB[3], uA[3], uB[3], nA[3], nB[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                zA[k] = sh.zc[rl][k];
                cA[k] = sh.zc[rl][3 + k];
                zB[k] = sh.zc[CT + cl][k];
                cB[k] = sh.zc[CT + cl][3 + k];
                uA[k] = -cA[k];  // Z_{k-1} - Z_k, exactly
                uB[k] = -cB[k];
                nA[k] = zA[k] + cA[k];
                nB[k] = zB[k] + cB[k];
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
            }
            float xA[3] = {nA[0], nA[1], nA[2]}, xB[3] = {nB[0], nB[1], nB[2]};
            act = pair_project(nA[0], nA[1], nA[2], nB[0], nB[1], nB[2], rho, rho, rho, rho, xA[0], xA[1], xA[2],
                               xB[0], xB[1], xB[2]);
            if constexpr (RES) {
                float gA2 = 0.0f, gB2 = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float gA = (xA[k] - zA[k]) - cA[k];
                    const float gB = (xB[k] - zB[k]) - cB[k];
                    cadd(&sh.acc[rl][3 + k], gA);
                    cadd(&sh.acc[cs][3 + k], gB);
                    gA2 += gA * gA;
                    gB2 += gB * gB;
                }
                cadd(&sh.acc[rl][6], gA2);
                cadd(&sh.acc[cs][6], gB2);
            }
            if (act) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    mA[k] = xA[k] + uA[k];
                    mB[k] = xB[k] + uB[k];
                    cadd(&sh.acc[rl][k], mA[k] - zA[k]);
                    cadd(&sh.acc[cs][k], mB[k] - zB[k]);
                }
            }
        }
        const unsigned long long bal = __ballot(act);
        if (act) {
            const int slot = base + (int)__popcll(bal & ((1ull << lane) - 1ull));
            Ln[2 * slot] = make_float4(mA[0], mA[1], mA[2], __int_as_float(idx));
            Ln[2 * slot + 1] = make_float4(mB[0], mB[1], mB[2], 0.0f);
        }
        base += (int)__popcll(bal);
    }
    if (lane == 0) a.cnt[(long)(par ^ 1) * a.nt + t] = base;

    // the diagonal tile's box and radius edges: the plain U update, prox
    // and X, U stores of the fused iteration
    if (I == J && lane < CT && I * CT + lane < N) {
        const int cc = I * CT + lane;
        const long S = a.stride;
        float z[3], cz[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            z[k] = sh.zc[lane][k];
            cz[k] = sh.zc[lane][3 + k];
        }
        const long p0 = (long)t * (2 * CTT) + (long)lane * CT + lane;
        float g2 = 0.0f;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const long q = p0 + side * CTT;
            float x[3], u[3], n[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = a.X[k * S + q];
                u[k] = (x[k] + a.U[k * S + q]) - z[k];
                n[k] = z[k] - u[k];
            }
            if (side == 0) {
                box3(x[0], x[1], x[2], n[0], n[1], n[2]);
            } else {
                x[0] = n[0];
                x[1] = n[1];
                x[2] = max_alpha_radius(n[2], rho, a.aexp[cc]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a.X[k * S + q] = x[k];
                a.U[k * S + q] = u[k];
                cadd(&sh.acc[lane][k], (x[k] + u[k]) - z[k]);
                if constexpr (RES) {
                    const float g = (x[k] - z[k]) - cz[k];
                    cadd(&sh.acc[lane][3 + k], g);
                    g2 += g * g;
                }
            }
        }
        if constexpr (RES) cadd(&sh.acc[lane][6], g2);
    }
    wave_sync();
    // circle partials: Pc[q][K][c] for the rows (K = J) and columns (K = I)
    const bool on = !(I == J && lane >= CT);
    const int circ = lane < CT ? I * CT + lane : J * CT + (lane - CT);
    const int K = lane < CT ? J : I;
    if (on && circ < N)
#pragma unroll
        for (int q = 0; q < Q; ++q)
            a.Pc[((long)q * nb + K) * N + circ] = (float)((double)(long long)sh.acc[lane][q] * CFIX_INV);
}

// Z_{k+1} = Z_k + sum a / (N + 1) over the circle's nb block partials in
// fixed order, the Z history shift, the list flip and, RES, the per-circle
// primal and dual residuals.  Workgroup = CZ_CIRC circles (one tile block:
// 128-byte rows of the partials) x CZ_SLICES slices of the blocks, added by
// a fixed-shape tree.
constexpr int CZ_CIRC = 32, CZ_SLICES = 32;

template <bool RES>
__global__ __launch_bounds__(CZ_CIRC * CZ_SLICES) void k_contact_zreduce(CArgs a)
{
    constexpr int Q = RES ? CQ : 3;
    __shared__ float part[CZ_SLICES][Q][CZ_CIRC];
    const int ks = threadIdx.x / CZ_CIRC, cl = threadIdx.x % CZ_CIRC;
    const int c = blockIdx.x * CZ_CIRC + cl;
    const int N = a.N, nb = a.nb;
    float s[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) s[q] = 0.0f;
    if (c < N)
#pragma unroll 4
        for (int K = ks; K < nb; K += CZ_SLICES)
#pragma unroll
            for (int q = 0; q < Q; ++q) s[q] += a.Pc[((long)q * nb + K) * N + c];
#pragma unroll
    for (int q = 0; q < Q; ++q) part[ks][q][cl] = s[q];
    __syncthreads();
    for (int h = CZ_SLICES / 2; h > 0; h >>= 1) {
        if (ks < h)
#pragma unroll
            for (int q = 0; q < Q; ++q) part[ks][q][cl] += part[ks + h][q][cl];
        __syncthreads();
    }
    if (threadIdx.x >= CZ_CIRC || c >= N) return;
    const float deg = (float)(N + 1);
    double p = 0.0, dz2 = 0.0;
    if constexpr (RES) p = (double)part[0][6][cl];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float zk = a.Z[3 * c + k], zp = a.Zp[3 * c + k];
        const float zn = zk + part[0][k][cl] / deg;
        if constexpr (RES) {
            const float dz = zn - zk;
            const double e = (double)(dz - (zk - zp));
            p += -2.0 * e * (double)part[0][3 + k][cl] + (double)deg * e * e;
            dz2 += (double)dz * dz;
        }
        a.Zpp[3 * c + k] = zp;
        a.Zp[3 * c + k] = zk;
        a.Z[3 * c + k] = zn;
    }
    if constexpr (RES) {
        const float rho = a.uni[0];
        a.Rpc[c] = p;
        a.Rd[c] = (double)rho * rho * (double)deg * dz2;
    }
    if (c == 0) *a.cur ^= 1;
}

// Enters the contact state from the steady X, U state (U one update behind,
// Z = Z_k, Z_prev = Z_{k-1}): list 0 receives every pair with
// M = X + U != Z_{k-1} on either edge, exactly, so the first contact
// iteration computes the same U = M - Z_k as the fused one.
__global__ __launch_bounds__(256) void k_contact_enter(CArgs a)
{
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int2 ij0 = a.tiles[blockIdx.x];
    const int I = ij0.x, J = ij0.y & ~TILE_RAISED;
    const long t = (long)I * a.nb - (long)I * (I - 1) / 2 + (J - I);
    const int N = a.N;
    const long S = a.stride;
    float4* L0 = a.L + t * (2 * CTT);
    const int il = tid >> 3, c0 = (tid & 7) * 4;
    const int i = I * CT + il;
    const int ic = min(i, N - 1);
    float zpi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) zpi[k] = a.Zp[3 * ic + k];
    unsigned exm = 0u;
    float mA[4][3], mB[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int jl = c0 + c, j = J * CT + jl;
        bool vm = (i < N) && (j < N);
        if (I == J) vm = vm && (il < jl);
        if (!vm) continue;
        const long pa = t * (2 * CTT) + (long)il * CT + jl, pb = pa + CTT;
        bool ex = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mA[c][k] = a.X[k * S + pa] + a.U[k * S + pa];
            mB[c][k] = a.X[k * S + pb] + a.U[k * S + pb];
            ex = ex || mA[c][k] != zpi[k] || mB[c][k] != a.Zp[3 * j + k];
        }
        exm |= ex ? 1u << c : 0u;
    }
    int total = 0;
    int slot = block_offset(__popc(exm), wsum, total);
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if ((exm >> c) & 1u) {
            L0[2 * slot] = make_float4(mA[c][0], mA[c][1], mA[c][2], __int_as_float(il * CT + c0 + c));
            L0[2 * slot + 1] = make_float4(mB[c][0], mB[c][1], mB[c][2], 0.0f);
            ++slot;
        }
    if (tid == 0) a.cnt[t] = total;
}

// Leaves the contact state after iteration k (Z_prev = Z_k, Z_prev2 =
// Z_{k-1}, list cur = L_k, the other = L_{k-1}): every pair edge gets
//   U_k = M_{k-1} - Z_k,  X_k = M_k - U_k
// with M = Z of its circle off the lists: the steady X, U state.
__global__ __launch_bounds__(256) void k_contact_leave(CArgs a)
{
    __shared__ unsigned short map[2][CTT];
    __shared__ unsigned mask[2][CT];
    const int tid = threadIdx.x;
    const int2 ij0 = a.tiles[blockIdx.x];
    const int I = ij0.x, J = ij0.y & ~TILE_RAISED;
    const long t = (long)I * a.nb - (long)I * (I - 1) / 2 + (J - I);
    const int N = a.N;
    const long S = a.stride;
    const int par = *a.cur;
    const float4* Ls[2] = {a.L + ((long)par * a.nt + t) * (2 * CTT), a.L + ((long)(par ^ 1) * a.nt + t) * (2 * CTT)};
    const int cn[2] = {a.cnt[(long)par * a.nt + t], a.cnt[(long)(par ^ 1) * a.nt + t]};
    if (tid < CT) mask[0][tid] = mask[1][tid] = 0u;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 2; ++b)
        for (int e = tid; e < cn[b]; e += 256) {
            const int idx = __float_as_int(Ls[b][2 * e].w);
            map[b][idx] = (unsigned short)e;
            atomicOr(&mask[b][idx >> 5], 1u << (idx & 31));
        }
    __syncthreads();
    const int il = tid >> 3, c0 = (tid & 7) * 4;
    const int i = I * CT + il;
    const int ic = min(i, N - 1);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int jl = c0 + c, j = J * CT + jl;
        bool vm = (i < N) && (j < N);
        if (I == J) vm = vm && (il < jl);
        if (!vm) continue;
        const int idx = il * CT + jl;
        const bool in_k = (mask[0][il] >> jl) & 1u, in_k1 = (mask[1][il] >> jl) & 1u;
        float4 ek[2] = {}, ek1[2] = {};
        if (in_k) { ek[0] = Ls[0][2 * map[0][idx]]; ek[1] = Ls[0][2 * map[0][idx] + 1]; }
        if (in_k1) { ek1[0] = Ls[1][2 * map[1][idx]]; ek1[1] = Ls[1][2 * map[1][idx] + 1]; }
        const long pa = t * (2 * CTT) + (long)idx, pb = pa + CTT;
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
    a.nb = d->nb;
    a.nt = d->nt;
    a.stride = s->stride;
    a.tiles = d->d_tiles;
    a.X = s->d_X;
    a.U = s->d_U;
    a.Z = s->d_Z;
    a.Zp = s->d_Zprev;
    a.Zpp = d->d_Zpp;
    a.aexp = d->d_aexp;
    a.uni = d->d_uni;
    a.cur = d->d_cur;
    a.L = d->d_L;
    a.cnt = d->d_cnt;
    a.Pc = d->d_Pc;
    a.Rpc = d->d_Rpc;
    a.Rd = d->d_Rd;
    return a;
}

bool dense_contact_ok(DevGraph* s)
{
    const DenseState* d = s->dense;
    if (!d || d->T != CT) return false;
    if (!(s->rho_uni && s->alpha_uni && s->alpha_u == 1.0f && s->policy_raised == 0)) return false;
    const char* e = getenv("PARADMM_CONTACT");
    return !(e && e[0] == '0');
}

bool dense_in_contact(DevGraph* s) { return s->dense && s->dense->c_state; }

static int contact_alloc(DevGraph* s)
{
    DenseState* d = s->dense;
    if (d->d_L) return PARADMM_OK;
    RC(dmalloc(s, (void**)&d->d_L, sizeof(float4) * 2 * 2 * (size_t)d->nt * CTT));
    RC(dmalloc(s, (void**)&d->d_cnt, sizeof(int) * 2 * (size_t)d->nt));
    RC(dmalloc(s, (void**)&d->d_cur, sizeof(int)));
    RC(dmalloc(s, (void**)&d->d_Pc, sizeof(float) * CQ * (size_t)d->nb * d->N));
    RC(dmalloc(s, (void**)&d->d_Rpc, sizeof(double) * (size_t)d->N));
    RC(dmalloc(s, (void**)&d->d_Zpp, sizeof(float) * 3 * (size_t)d->N));
    return PARADMM_OK;
}

// res: also the residual sums (the last iteration of a run)
static int contact_iteration(DevGraph* s, bool res)
{
    DenseState* d = s->dense;
    const CArgs a = cargs(s);
    const unsigned tiles_grid = (unsigned)((d->nt + 3) / 4);
    prof_begin(s, res ? "contact_tile_res" : "contact_tile");
    if (res) k_contact_tile<true><<<tiles_grid, 256, 0, s->stream>>>(a);
    else k_contact_tile<false><<<tiles_grid, 256, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    prof_begin(s, res ? "contact_zreduce_res" : "contact_zreduce");
    const unsigned zgrid = (unsigned)((d->N + CZ_CIRC - 1) / CZ_CIRC);
    if (res) k_contact_zreduce<true><<<zgrid, CZ_CIRC * CZ_SLICES, 0, s->stream>>>(a);
    else k_contact_zreduce<false><<<zgrid, CZ_CIRC * CZ_SLICES, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
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
    CK(hipMemsetAsync(d->d_cur, 0, sizeof(int), s->stream));
    const CArgs a = cargs(s);
    prof_begin(s, "contact_enter");
    k_contact_enter<<<(unsigned)d->nt, 256, 0, s->stream>>>(a);
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
    k_contact_leave<<<(unsigned)d->nt, 256, 0, s->stream>>>(a);
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
        k_dense_res_final<<<1, 256, 0, s->stream>>>(d->N, d->d_Rpc, d->N, d->d_Rd, s->d_res);
        CK(hipGetLastError());
        return PARADMM_OK;
    }
    RC(dense_finish(s));
    k_dense_res_final<<<1, 256, 0, s->stream>>>(d->nt, d->d_Rp, d->N, d->d_Rd, s->d_res);
    CK(hipGetLastError());
    return PARADMM_OK;
}

}  // namespace pdm

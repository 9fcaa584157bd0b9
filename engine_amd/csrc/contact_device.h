// contact_device.h -- device code of the contact state of the dense
// pair-tile layout, shared by the single-instance kernels (dense.hip) and
// the batched solver (batch.hip).  Internal.
#pragma once
#include <hip/hip_runtime.h>

#include "prox_device.h"

namespace pdm {

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

// Inside the contact state the box and radius edges of circle c live in a
// compact per-circle array, component f at Xu[f N + c]: f = 0..2 the box
// edge's X, 3..5 its U (one update behind), 6..8 and 9..11 the radius
// edge's, 12 and 13 their rho.  (In the dense layout they sit on the
// diagonals of the dense tiles, T + 1 floats apart: one cache line per
// component and circle.)  The entry copies them in, the exit back.
constexpr int CUNARY = 14;

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
    float* Xu;             // [CUNARY][N] the box and radius edges' state inside the contact state
    // spatial form (contact_sparse.h): circles in contact order, a tile only
    // where its two blocks' boxes meet; null in the id-order form
    const int* perm = nullptr;   // [N] contact index -> circle
    const int* inv = nullptr;    // [N] circle -> contact index
    float4* zcv = nullptr;       // [2][N] per contact circle: (Z_k, c) of this launch, written before the tiles
    float4* bb = nullptr;        // [nbs] per block: box of its circles' common N (x lo, x hi, y lo, y hi)
};

// words of a contact tile row's pair mask
template <int CS>
constexpr int cwords() { return CS >= 32 ? CS / 32 : 1; }

// v rounded to the nearest multiple of 2^-31, in units of 2^-31:
// v = k + f with k = rint(v) and f = v - k exact (|f| <= 1/2), so
// v 2^31 = k 2^31 + rint(f 2^31) needs no FP64.  A value outside
// |v| < 2^30 (NaN included) gives 0 and sets bad (a select, no branch).
__device__ __forceinline__ long long cfix(float v, unsigned& bad)
{
    const bool ok = fabsf(v) < 1073741824.0f;
    bad |= ok ? 0u : 1u;
    const float w = ok ? v : 0.0f;
    const float k = rintf(w);
    const float f = w - k;
    return (long long)(int)k * 2147483648LL + (long long)(int)rintf(f * 2147483648.0f);
}

// v >= 0 exactly to 2^-62 as hi 2^-31 + lo 2^-62: hi = the multiple of
// 2^-31 below v, lo = rint((v - hi 2^-31) 2^62) with the remainder exact
__device__ __forceinline__ void cfix2(float v, long long& hi, long long& lo, unsigned& bad)
{
    const bool ok = fabsf(v) < 1073741824.0f;
    bad |= ok ? 0u : 1u;
    const float t = (ok ? v : 0.0f) * 2147483648.0f;  // exact
    const float k = floorf(t);
    hi = (long long)k;
    lo = (long long)rintf((t - k) * 2147483648.0f);  // t - k exact, in [0, 1)
}

// a circle sum (units of 2^-31) as fp32: one rounding, no FP64
__device__ __forceinline__ float csum_f32(unsigned long long v) { return (float)(long long)v * 4.656612873077393e-10f; }

__device__ __forceinline__ void cadd(unsigned& bad, unsigned long long* acc, float v)
{
    atomicAdd(acc, (unsigned long long)cfix(v, bad));
}

__device__ __forceinline__ void cadd2(unsigned& bad, unsigned long long* acc, float v)
{
    long long hi, lo;
    cfix2(v, hi, lo, bad);
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

// Exclusive prefix over the wavefront's lanes of n (0 <= n < 2^BITS), and
// the total, by one ballot per bit of n: no LDS round trips (a shuffle
// scan takes six dependent ds_bpermute).
template <int BITS>
__device__ __forceinline__ int wave_prefix_small(int n, int& total)
{
    int pre = 0, all = 0;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const unsigned long long m = __ballot((n >> b) & 1);
        pre += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
        all += (int)__popcll(m) << b;
    }
    total = all;
    return pre;
}

// Shared scratch of one contact tile (one wavefront).  The pairs to project
// form one stream per tile: the previous list's pairs, then the colliding
// pairs found by the tests, which pass through a ring in LDS.
// ring capacity: more than 63 pending items plus the 2 RL 64 pairs one test
// step can add (RL = CS / 8): 1024 for CS = 32, 2048 for CS = 64
template <int CS>
constexpr int contact_ring() { return CS == 16 ? 512 : CS == 32 ? 1024 : 2048; }

template <int CS, int Q>
struct ContactWave {
    unsigned long long acc[2 * CS][Q];  // fixed-point circle sums: rows (block P), then columns (block Q)
    float zc[2 * CS][6];                // Z_k and c = Z_k - Z_{k-1} of the tile's 2 CS circles
    unsigned short ring[contact_ring<CS>()];  // colliding pairs not in the previous list (slot)
    unsigned mask[CS][cwords<CS>()];    // pairs in the previous list, per row
    unsigned nmask[CS][cwords<CS>()];   // pairs in the new list
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
    unsigned bad = 0u;  // in a register: one LDS write per pair, not a branch per sum
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
            cadd(bad, &sh.acc[rl][3 + k], gA);
            cadd(bad, &sh.acc[cs][3 + k], gB);
            gA2 += gA * gA;
            gB2 += gB * gB;
        }
        cadd2(bad, &sh.acc[rl][6], gA2);
        cadd2(bad, &sh.acc[cs][6], gB2);
    }
    if (act) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mA[k] = xA[k] + uA[k];
            mB[k] = xB[k] + uB[k];
            // per-edge rho: the consensus weighs each deviation with its edge's rho
            cadd(bad, &sh.acc[rl][k], RE ? rA * (mA[k] - zA[k]) : mA[k] - zA[k]);
            cadd(bad, &sh.acc[cs][k], RE ? rB * (mB[k] - zB[k]) : mB[k] - zB[k]);
        }
    }
    if (bad) sh.bad = 1u;
    return act;
}

// One round of the tile's pair stream: items d .. d + 63 (item < cp: entry
// of the previous list, (pa, pb) loaded by the caller; else ring slot
// item - cp); the active pairs are appended to the new list in lane order.
template <int CS, int Q, bool RE>
__device__ __forceinline__ void contact_round(ContactWave<CS, Q>& sh, const CArgs& a, int P, int Qb, int d, int avail,
                                              int cp, float4 pa, float4 pb, float4* Ln, int& base, bool diag,
                                              float rho)
{
    constexpr int CSS = CS * CS;
    const int lane = threadIdx.x & 63;
    const int item = d + lane;
    bool act = false;
    int idx = 0;
    float mA[3] = {0.0f, 0.0f, 0.0f}, mB[3] = {0.0f, 0.0f, 0.0f};
    if (item < avail) {
        const bool ex = item < cp;
        if (ex) {
            idx = __float_as_int(pa.w) & (CSS - 1);
        } else {
            idx = sh.ring[(item - cp) & (contact_ring<CS>() - 1)];
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
    // the active pairs' entries in lane order after the list's `base`; the
    // stores are issued by every lane (an inactive lane's offset lies past
    // the list, where a raw buffer store is dropped), so no branch sits
    // around them and waiting for a later load does not wait for them (the
    // raw descriptor: stride 0, the list's CSS * 32 bytes, word 3 the
    // 32-bit data format)
    const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc(Ln, 0, CSS * 32, 0x00020000);
    const int slot = base + (int)__popcll(bal & ((1ull << lane) - 1ull));
    const int off = act ? slot * 32 : CSS * 32;
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 ea = {__float_as_uint(mA[0]), __float_as_uint(mA[1]), __float_as_uint(mA[2]), (unsigned)idx};
    const u4 eb = {__float_as_uint(mB[0]), __float_as_uint(mB[1]), __float_as_uint(mB[2]), 0u};
    __builtin_amdgcn_raw_buffer_store_b128(ea, lr, off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(eb, lr, off + 16, 0, 0);
    if (act) atomicOr(&sh.nmask[idx / CS][(idx % CS) >> 5], 1u << ((idx % CS) & 31));
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
    for (int e = (int)(threadIdx.x & 63); e < CS; e += 64) {
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
        const long N = a.N;
        long long s[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) s[q] = 0;
        float g2 = 0.0f;
        unsigned bad = 0u;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            float* xu = a.Xu + 6 * side * N + cc;
            const float r = RE ? a.Xu[(12 + side) * N + cc] : a.uni[0];
            float x[3], u[3], nn[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = xu[k * N];
                u[k] = (x[k] + xu[(3 + k) * N]) - z[k];
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
                xu[k * N] = x[k];
                xu[(3 + k) * N] = u[k];
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


// Contact tile t = (P, Qb) of the iteration (one wavefront, lane =
// threadIdx.x % 64, sh its own scratch).
template <int CS, bool RES, bool RE, bool SPATIAL = false>
__device__ __forceinline__ void contact_tile_body(const CArgs& a, const int t, const int P, const int Qb,
                                                  ContactWave<CS, RES ? CQ : 3>& sh)
{
    constexpr int Q = RES ? CQ : 3;
    constexpr int CSS = CS * CS, RL = CS / 8, W = cwords<CS>();
    // spatial form: the circles' (Z_k, c) come from the launch before, and a
    // list of length 0 may have a stale mask (its tile was not launched)
    constexpr bool SP = SPATIAL;
    const int lane = threadIdx.x & 63;
    const int N = a.N;
    const bool diag = P == Qb;
    const long tp = (long)a.par * a.nts + t, tn = (long)(a.par ^ 1) * a.nts + t;
    const float4* Lp = a.L + tp * (2 * CSS);
    float4* Ln = a.L + tn * (2 * CSS);
    const int cp = min(a.cnt[tp], CSS);
    const float rho = a.uni[0];

    // The tile's first loads are issued together, at clamped indices with
    // no branch, before any is waited for (each behind a branch of its own
    // was a round trip in turn): the circle table's (Z_k, c), entries
    // 0..CS-1 the rows' circles, CS..2CS-1 the columns' (lane e + 64 i:
    // entry e + 64 i); the previous list's mask (every tile's mask is inside
    // the buffer; a stale one, spatial form, is replaced by 0); spatial form,
    // the previous list's first 64 entries (within the list's capacity,
    // before its length is known; entries past it are not used).
    constexpr int NZ = (2 * CS + 63) / 64, NM = (CS * W + 63) / 64;
    float zt[NZ][3], ct[NZ][3];
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
        const int e = min(lane + 64 * i, 2 * CS - 1);
        const int cc = min((e < CS ? P * CS + e : Qb * CS + (e - CS)), N - 1);
        if constexpr (SP) {
            const float4 u = a.zcv[2 * cc], v = a.zcv[2 * cc + 1];
            zt[i][0] = u.x, zt[i][1] = u.y, zt[i][2] = u.z;
            ct[i][0] = u.w, ct[i][1] = v.x, ct[i][2] = v.y;
        } else {
            contact_zc<RE>(a, cc, zt[i], ct[i]);
        }
    }
    unsigned mt[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) mt[i] = a.Lmask[tp * CS * W + min(lane + 64 * i, CS * W - 1)];
    // the first 64 entries of the previous list, in flight during the tests
    // (id-order form: only those of the list, after the table is written)
    float4 pre0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pre1 = pre0;
    if constexpr (SP) {
        pre0 = Lp[2 * lane];
        pre1 = Lp[2 * lane + 1];
    }
    // the writes at the same clamped indices, with no branch (a branch
    // would take the loads with it): lanes past the last entry write the
    // last entry's values again
#pragma unroll
    for (int i = 0; i < NZ; ++i) {
        const int e = min(lane + 64 * i, 2 * CS - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sh.zc[e][k] = zt[i][k];
            sh.zc[e][3 + k] = ct[i][k];
        }
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const int e = min(lane + 64 * i, CS * W - 1);
        (&sh.mask[0][0])[e] = (!SP || cp) ? mt[i] : 0u;
        (&sh.nmask[0][0])[e] = 0u;
    }
    for (int k = lane; k < 2 * CS * Q; k += 64) (&sh.acc[0][0])[k] = 0ull;
    if (lane == 0) sh.bad = 0u;
    wave_lds_fence();
    if (!SP && lane < cp) {
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
    // the previous list's entries of the round at `done` (cur) and of the
    // one after it (nxt, in flight while the round runs; the round's list
    // stores are issued after those loads and are not waited for: id-order
    // form 0.0180 -> 0.0176 ms per iteration at N = 5000, spatial form
    // 0.0253 -> 0.0244 at N = 10000)
    float4 cur0 = pre0, cur1 = pre1, nxt0 = pre0, nxt1 = pre1;
    auto next_entries = [&]() {
        const int it = done + 64 + lane;
        if (done + 64 < cp && it < cp) {
            nxt0 = Lp[2 * it];
            nxt1 = Lp[2 * it + 1];
        }
    };
    auto take_round = [&]() {
        next_entries();
        contact_round<CS, Q, RE>(sh, a, P, Qb, done, avail, cp, cur0, cur1, Ln, base, diag, rho);
        done += 64;
        cur0 = nxt0;
        cur1 = nxt1;
    };
    while (avail - done >= 64) take_round();
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
        // n <= 2 RL: 3, 4 or 5 bits for tiles of 16, 32, 64
        int tot;
        int off = avail - cp + wave_prefix_small<RL == 2 ? 3 : RL == 4 ? 4 : 5>(__popc(qm), tot);
        avail += tot;
        unsigned m = qm;
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1u;
            sh.ring[(off++) & (contact_ring<CS>() - 1)] = (unsigned short)((r0 + 2 * h + b / RL) * CS + c0 + b % RL);
        }
        wave_lds_fence();
        while (avail - done >= 64) take_round();
    }
    if (avail > done) contact_round<CS, Q, RE>(sh, a, P, Qb, done, avail, cp, cur0, cur1, Ln, base, diag, rho);
    // spatial form, where a wavefront takes tile after tile: entries loaded
    // ahead and not used (no round took them) are waited for here, before
    // the tile's stores and atomics; still in flight at the next tile, the
    // first reuse of their registers would wait for those stores and
    // atomics too (the memory counter is in order).  (A wavefront of the
    // id-order form has no next tile.)
    if constexpr (SP)
        asm volatile("" ::"v"(cur0.x), "v"(cur0.y), "v"(cur0.z), "v"(cur0.w), "v"(cur1.x), "v"(cur1.y), "v"(cur1.z),
                     "v"(cur1.w));
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

// Contact tile t of the iteration, (P, Qb) found from t.
template <int CS, bool RES, bool RE, bool SPATIAL = false>
__device__ __forceinline__ void contact_tile_body(const CArgs& a, const int t, ContactWave<CS, RES ? CQ : 3>& sh)
{
    int P, Qb;
    tri_tile(t, a.nbs, P, Qb);
    contact_tile_body<CS, RES, RE, SPATIAL>(a, t, P, Qb, sh);
}

// The end of a run: Z_{k+1} = Z_k + sum a / (N + 1) from the last launch's
// sums (RE, rho per edge: Z_k + sum rho a / sum rho) into the canonical Z,
// Z_prev, Z_prev2 (Zin = Z_k, Zpin = Z_{k-1}: canonical or inside the run's
// ring), the clearing of the sum buffers and, RES, the per-circle primal
// and dual residuals.  One thread per circle.
constexpr int CZ_THREADS = 256;


template <bool RES, bool RE>
__device__ __forceinline__ void contact_zreduce_circle(const CArgs& a, const int c)
{
    constexpr int Q = RES ? CQ : 3;
    const int N = a.N;
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
struct ContactEnterShared {
    int wsum[4];
    unsigned emask[CS][cwords<CS>()];
};

template <int CS>
__device__ __forceinline__ void contact_enter_tile(const CArgs& a, const int t, ContactEnterShared<CS>& sh)
{
    constexpr int CSS = CS * CS, W = cwords<CS>(), RS = 256 / CS, CPT = CSS / 256;
    int(&wsum)[4] = sh.wsum;
    unsigned(&emask)[CS][W] = sh.emask;
    const int tid = threadIdx.x;
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
    // a diagonal tile: its circles' box and radius edges into the compact array
    if (P == Qb && tid < CS) {
        const int c = P * CS + tid;
        if (c < N) {
            const long q0 = dense_unary_slot(c, a.nb, a.T);
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const long q = q0 + side * TT;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    a.Xu[(6 * side + k) * (long)N + c] = a.X[k * S + q];
                    a.Xu[(6 * side + 3 + k) * (long)N + c] = a.U[k * S + q];
                }
                a.Xu[(12 + side) * (long)N + c] = a.rho[q];
            }
        }
    }
}


// Leaves the contact state after iteration k (Z_prev = Z_k, Z_prev2 =
// Z_{k-1}, list par = L_k, the other = L_{k-1}): every pair edge gets
//   U_k = M_{k-1} - Z_k,  X_k = M_k - U_k
// with M = Z of its circle off the lists: the steady X, U state.  256
// threads per contact tile, as k_contact_enter.
template <int CS>
struct ContactLeaveShared {
    unsigned short map[2][CS * CS];
    unsigned mask[2][CS][cwords<CS>()];
};

template <int CS>
__device__ __forceinline__ void contact_leave_tile(const CArgs& a, const int t, ContactLeaveShared<CS>& sh)
{
    constexpr int CSS = CS * CS, W = cwords<CS>(), RS = 256 / CS, CPT = CSS / 256;
    unsigned short(&map)[2][CSS] = sh.map;
    unsigned(&mask)[2][CS][W] = sh.mask;
    const int tid = threadIdx.x;
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
            atomicOr(&mask[b][idx / CS][(idx % CS) >> 5], 1u << ((idx % CS) & 31));
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
    // a diagonal tile: its circles' box and radius edges back to their slots
    if (P == Qb && tid < CS) {
        const int c = P * CS + tid;
        if (c < N) {
            const long q0 = dense_unary_slot(c, a.nb, a.T);
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const long q = q0 + side * TT;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    a.X[k * S + q] = a.Xu[(6 * side + k) * (long)N + c];
                    a.U[k * S + q] = a.Xu[(6 * side + 3 + k) * (long)N + c];
                }
            }
        }
    }
}

}  // namespace pdm

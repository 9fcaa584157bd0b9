// contact_policy.h -- device code of the learned rho policy applied inside
// the contact state, shared by the single-instance kernels (dense.hip) and
// the batched solver (batch.hip).  Internal.
#pragma once
#include <hip/hip_runtime.h>

#include "contact_device.h"
#include "mlp_device.h"

namespace pdm {

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

// LDS of one policy tile (one workgroup of four wavefronts)
template <int CS>
struct CPolShared {
    MlpShared w;
    TileScratch ts[4];
    float yv[4][64];
    float zc[2 * CS][9];             // Z, Z_prev, Z_prev2 of the tile's circles
    unsigned short omap[CS * CS];    // other list: pair -> entry
    unsigned omask[CS][cwords<CS>()], pmask[CS][cwords<CS>()], rmask[CS][cwords<CS>()];
    unsigned short q[CS * CS];       // the tile's pairs to evaluate: kind << 14 | entry or pair
    int wsum[4];
};

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
__device__ __forceinline__ void cpol_tile_body(const CPolArgs& p, const int t, CPolShared<CS>& sh)
{
    constexpr int CSS = CS * CS, W = cwords<CS>();
    MlpShared& w = sh.w;
    TileScratch(&ts)[4] = sh.ts;
    float(&yv)[4][64] = sh.yv;
    float(&zc)[2 * CS][9] = sh.zc;
    unsigned short(&omap)[CSS] = sh.omap;
    unsigned(&omask)[CS][W] = sh.omask;
    unsigned(&pmask)[CS][W] = sh.pmask;
    unsigned(&rmask)[CS][W] = sh.rmask;
    unsigned short(&q)[CSS] = sh.q;
    int(&wsum)[4] = sh.wsum;
    const CArgs& a = p.c;
    float* const rhow = const_cast<float*>(a.rho);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
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
    auto inmask = [&](const unsigned (&m)[CS][W], int idx) { return (m[idx / CS][(idx % CS) >> 5] >> ((idx % CS) & 31)) & 1u; };
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
            atomicOr(&rmask[idx / CS][(idx % CS) >> 5], 1u << ((idx % CS) & 31));
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
                if (side == 0) atomicOr(&a.Lmask[tp * CS * W + (idx / CS) * W + ((idx % CS) >> 5)], 1u << ((idx % CS) & 31));
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
                } else {  // the compact box (kind 1) or radius (kind 2) edge
                    const int side = kind - 1;
#pragma unroll
                    for (int k = 0; k < 3; ++k) x[k] = a.Xu[(6 * side + k) * (long)N + c];
                    rho = a.Xu[(12 + side) * (long)N + c];
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
                const int side = kind - 1;
                float* xu = a.Xu + 6L * side * N + c;
                const float r = cpol_rho(p, p.yc[(long)kind * N + c], g, lg);
                const float sf = a.Xu[(12 + side) * (long)N + c] / r;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float x = xu[k * (long)N], z = a.Z[3 * c + k];
                    xu[(3 + k) * (long)N] = sf * ((x + xu[(3 + k) * (long)N]) - z) - x + z;
                }
                rhow[qq] = r;  // the dense slot (read after the state is left) and the compact copy
                a.Xu[(12 + side) * (long)N + c] = r;
            }
    }
}


// The graph rho: 10^(mean output over all edges) -- the pair edges evaluated
// on their own (ysum), (N - 1 - rows) times each circle's unlisted pair
// output, its box and radius outputs -- or with average the mean of this and
// the earlier estimates (the per-type policy's running sums, every type the
// same).  One workgroup; integer sums, so the order does not matter.
// sh: 256 int64 of LDS; one workgroup of 256 threads
__device__ __forceinline__ void cpol_final_body(const CPolArgs& p, const long E, long long (&sh)[256])
{
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
__device__ __forceinline__ void cpol_zprev_circle(const CPolArgs& p, const int c)
{
    const CArgs& a = p.c;
    const int N = a.N;
    const float g = p.gl[0], lg = p.gl[1];
    const float rv = cpol_rho(p, p.yc[c], g, lg);
    const float sf = p.base[c] / rv;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float z = a.Z[3 * c + k];
        a.Zp[3 * c + k] = z - sf * (z - a.Zp[3 * c + k]);
    }
    p.base[c] = rv;
    const float rb = a.Xu[12L * N + c], rr = a.Xu[13L * N + c];
    const double w = (double)(N - 1) * rv + (double)p.wacc[c] / (double)CPOL_WFIX + rb + rr;
    const double w2 = (double)(N - 1) * ((double)rv * rv) + (double)p.wacc[N + c] / (double)CPOL_W2FIX +
                      (double)rb * rb + (double)rr * rr;
    a.W[c] = (float)w;
    a.W[N + c] = (float)w2;
}


}  // namespace pdm

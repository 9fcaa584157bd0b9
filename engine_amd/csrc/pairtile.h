// pairtile.h -- the pair-tile pass of the all-pairs unit-square packing,
// shared by the single-instance dense layout (dense.hip) and the batched
// persistent solver (batch.hip).
//
// Layout of one instance (N circles, blocks of T circles, nb = ceil(N/T)):
// tile (I, J), I <= J, local index I*nb - I(I-1)/2 + (J-I), holds two
// sub-blocks of T*T edge slots:
//   A[il][jl] = the edge of circle i = I*T+il in pair (i, j),
//   B[il][jl] = the edge of circle j = J*T+jl in the same pair,
// so both edges of a pair sit at the same offset and a thread reads four
// pairs with float4 loads and no transpose.  Diagonal tiles use il < jl for
// pairs and hold circle c's box edge at A[c][c] and its radius edge at
// B[c][c]; the rest is padding (rho = 0, never read as a pair).
// Each edge component is an array of `stride` floats (SoA).
//
// One tile pass does, per edge, in the steady state (DM_FUSED):
//   U += alpha (X - Z)    the U update of the previous iteration (FP64 primal
//                         residual partial),
//   N = Z - U, X = prox(N), M = X + U,
//   per-tile partial sums of rho*M (3), rho, rho^2 for the tile's 2T circles.
// The Z consensus then sums a circle's nb block partials in fixed order.
#pragma once
#include <hip/hip_runtime.h>

#include "prox_device.h"

namespace pdm {

enum {
    DM_FUSED = 0,          // U update, N = Z - U, prox, partials      (steady state)
    DM_FUSED_CURRENT = 1,  // N = Z - U (U already current), prox, partials
    DM_FUSED_NBUF = 2,     // N from the N buffer, prox, partials       (after upload)
    DM_LEGACY_X = 3,       // N from the N buffer, prox                 (updateX)
    DM_LEGACY_Z = 4,       // partials of the stored M                   (updateZ)
    DM_FINISH = 5,         // U update only (+ primal residual)
    // alpha = 1 on every edge: U_{k+1} = U_k + X_{k+1} - Z_{k+1} = M_k - Z_{k+1}
    // with M_k = X_{k+1} + U_k, so between the first and the last iteration
    // of a run the X array can hold M alone and U is not stored: 24 instead
    // of 48 bytes per edge per iteration
    DM_M_ENTER = 6,        // as DM_FUSED, but the X array receives M = X + U and U is not stored
    DM_M = 7,              // X array holds M: U = M - Z, N = Z - U, prox, partials, X array receives the new M
    DM_M_EXIT = 8          // X array holds M: U = M - Z, N = Z - U, prox, partials, X and U stored (DM_FUSED's state)
};

// U + alpha (X - Z) of one edge component, d = X - Z.  For alpha = 1 it is
// rounded as (X + U) - Z, the value the M state computes from its stored
// M = X + U, so a captured run (which keeps M) and plain iterations give
// the same bits.
__device__ __forceinline__ float u_update(float u, float x, float z, float alpha, float d)
{
    return alpha == 1.0f ? (x + u) - z : u + alpha * d;
}

// One instance as the tile code sees it: pointers already offset to the
// instance's first slot / circle / partial.
struct DArgs {
    int N, nb;
    long stride;      // component stride of the edge arrays (whole buffer)
    float* X;         // [3][stride]
    float* U;         // [3][stride]
    const float* rho;
    const float* alpha;
    const float* Nb;  // [3][stride]
    const float* Mb;  // [3][stride]
    const float* Z;   // [N][3]
    float* P;         // [5][nb][N] block partials
    double* Rp;       // [nt] primal partials (single-instance kernel)
    const float* aexp;
    const int2* tiles;
    float rho_u, alpha_u;  // the common rho and alpha when UNI (per-edge arrays unread)
    const float* uni;      // device copy {rho_u, alpha_u} (kernels replayed from a graph read it here)
};

// Edge streams are read and written once per iteration; PDM_DENSE_NT=1 marks
// those accesses non-temporal.
#ifndef PDM_DENSE_NT
#define PDM_DENSE_NT 1
#endif
typedef float nf4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float* p)
{
#if PDM_DENSE_NT
    const nf4 v = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
#else
    return *reinterpret_cast<const float4*>(p);
#endif
}
__device__ __forceinline__ void st4(float* p, const float4& v)
{
#if PDM_DENSE_NT
    const nf4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<nf4*>(p));
#else
    *reinterpret_cast<float4*>(p) = v;
#endif
}
__device__ __forceinline__ float& cmp(float4& v, int c) { return (&v.x)[c]; }

// Thread geometry of a T x T tile on 256 threads: each thread owns four
// consecutive columns of RPT rows.
template <int T>
struct TG {
    static_assert(T == 32 || T == 64, "tile side 32 or 64");
    static constexpr int CG = T / 4;     // column groups
    static constexpr int RG = 256 / CG;  // row groups
    static constexpr int RPT = T / RG;   // rows per thread
    static constexpr int TT = T * T;     // slots per sub-block
};

// sum over the CG lanes of a row group (contiguous lanes)
template <int CG>
__device__ __forceinline__ float xsum_cg(float v)
{
#pragma unroll
    for (int o = 1; o < CG; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Shared scratch of one tile pass.
template <int T>
struct TileShared {
    float srow[5][T];
    float scol[4][5][T];
    double sred[4][3];
};

// The pairs of one tile.  VAR 0: interior tile (no masks), 1: tile touching
// the padded end (circle masks), 2: diagonal tile (masks, unary slots kept).
// QUAL: also the largest pair overlap r_i + r_j - |c_i - c_j| of Z (FP64).
// UNI: every edge has the same rho and alpha (a.rho_u, a.alpha_u): the
// per-edge arrays are not read (48 instead of 56 bytes per edge).
template <int T, int MODE, int VAR, bool QUAL, bool UNI = false>
__device__ __forceinline__ void pair_tile(const DArgs& a, const int t, const int I, const int J, TileShared<T>& sh,
                                          double& pr, double& qov)
{
    using G = TG<T>;
    constexpr bool MIN = (MODE == DM_M || MODE == DM_M_EXIT);      // the X array holds M
    constexpr bool MOUT = (MODE == DM_M_ENTER || MODE == DM_M);    // the X array receives M
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH || MODE == DM_M_ENTER);
    constexpr bool PROX = (MODE <= DM_LEGACY_X || MODE >= DM_M_ENTER);
    constexpr bool NBUF = (MODE == DM_FUSED_NBUF || MODE == DM_LEGACY_X);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z || MODE >= DM_M_ENTER);
    constexpr bool PART_M = (MODE == DM_LEGACY_Z);
    constexpr bool WRITE_U = (MODE == DM_FUSED || MODE == DM_FINISH || MODE == DM_M_EXIT);
    constexpr bool WRITE_X = PROX;
    // a diagonal tile's vector stores also cover its unary slots (processed
    // afterwards by diag_unary, which reads their old X) and its padding:
    // X is loaded there so masked slots are written back unchanged
    constexpr bool NEED_X = APPLY_U || (VAR == 2 && WRITE_X) || MIN;
    constexpr bool NEED_U = (APPLY_U || (PROX && !NBUF) || (PART && !PART_M)) && !MIN;
    constexpr bool NEED_A = APPLY_U;

    const int N = a.N;
    const long S = a.stride;
    const int tid = threadIdx.x;
    const int rg = tid / G::CG;           // row group: rows RPT*rg .. RPT*rg+RPT-1
    const int c0 = (tid % G::CG) * 4;     // first of this thread's 4 columns
    const int base = t * (2 * G::TT);     // slots fit in 31 bits (checked at layout time)

    float zj[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = min(J * T + c0 + c, N - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) zj[c][k] = a.Z[3 * j + k];
    }
    float cacc[4][5];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 5; ++q) cacc[c][q] = 0.0f;

#pragma unroll 1
    for (int r = 0; r < G::RPT; ++r) {
        const int il = rg * G::RPT + r;
        const int i = I * T + il;
        const int ic = min(i, N - 1);
        float zi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) zi[k] = a.Z[3 * ic + k];
        const int pa = base + il * T + c0;  // 32-bit lane offsets off uniform bases
        const int pb = pa + G::TT;

        float racc[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // squared primal differences of this row's 4 pairs (<= 24 terms) in
        // fp32; the accumulation across rows, threads and tiles is FP64
        float prow = 0.0f;
        // a lane whose 4 slots hold no pair edge (columns past N, the lower
        // half of a diagonal tile) reads and writes nothing: padding costs
        // no HBM traffic
        bool on = true;
        if constexpr (VAR == 1) on = J * T + c0 < N;
        if constexpr (VAR == 2) on = (c0 + 3 > il) && (J * T + max(c0, il + 1) < N);
        if (on) {
        float4 xa[3], xb[3], ua[3], ub[3], na[3], nbv[3], ma[3], mb[3];
        float4 ra, rb, aa, ab;
        if constexpr (MIN)
#pragma unroll
            for (int k = 0; k < 3; ++k) ua[k] = ub[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (UNI) {
            ra = rb = make_float4(a.rho_u, a.rho_u, a.rho_u, a.rho_u);
            if constexpr (NEED_A) aa = ab = make_float4(a.alpha_u, a.alpha_u, a.alpha_u, a.alpha_u);
        } else {
            ra = ld4(a.rho + pa);
            rb = ld4(a.rho + pb);
            if constexpr (NEED_A) { aa = ld4(a.alpha + pa); ab = ld4(a.alpha + pb); }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (NEED_X) { xa[k] = ld4((a.X + k * S) + pa); xb[k] = ld4((a.X + k * S) + pb); }
            if constexpr (NEED_U) { ua[k] = ld4((a.U + k * S) + pa); ub[k] = ld4((a.U + k * S) + pb); }
            if constexpr (NBUF) { na[k] = ld4((a.Nb + k * S) + pa); nbv[k] = ld4((a.Nb + k * S) + pb); }
            if constexpr (PART_M) { ma[k] = ld4((a.Mb + k * S) + pa); mb[k] = ld4((a.Mb + k * S) + pb); }
        }

#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int jl = c0 + c;
            // pair slots only; the diagonal's unary factors run in diag_unary
            bool vm = true;
            if constexpr (VAR >= 1) vm = (i < N) && (J * T + jl < N);
            if constexpr (VAR == 2) vm = vm && (il < jl);
            if constexpr (QUAL) {
                const double dx = (double)zi[0] - (double)zj[c][0], dy = (double)zi[1] - (double)zj[c][1];
                const double ov = (double)zi[2] + (double)zj[c][2] - sqrt(dx * dx + dy * dy);
                qov = (vm && ov > qov) ? ov : qov;
            }
            float xA[3], xB[3], uA[3], uB[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (NEED_U) { uA[k] = cmp(ua[k], c); uB[k] = cmp(ub[k], c); }
                if constexpr (MIN) { uA[k] = cmp(xa[k], c) - zi[k]; uB[k] = cmp(xb[k], c) - zj[c][k]; }
            }
            if constexpr (APPLY_U) {
                const float alA = cmp(aa, c), alB = cmp(ab, c);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float dA = cmp(xa[k], c) - zi[k];
                    const float dB = cmp(xb[k], c) - zj[c][k];
                    // alpha = 1: (X + U) - Z, the rounding of the M state's M - Z
                    const float nuA = u_update(uA[k], cmp(xa[k], c), zi[k], alA, dA);
                    const float nuB = u_update(uB[k], cmp(xb[k], c), zj[c][k], alB, dB);
                    uA[k] = vm ? nuA : uA[k];
                    uB[k] = vm ? nuB : uB[k];
                    prow += vm ? dA * dA + dB * dB : 0.0f;
                }
            }
            if constexpr (PROX) {
                float nA[3], nB[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    nA[k] = NBUF ? cmp(na[k], c) : zi[k] - uA[k];
                    nB[k] = NBUF ? cmp(nbv[k], c) : zj[c][k] - uB[k];
                    xA[k] = nA[k];
                    xB[k] = nB[k];
                }
                const float rA = cmp(ra, c), rB = cmp(rb, c);
                if (vm)
                    pair_project(nA[0], nA[1], nA[2], nB[0], nB[1], nB[2], rA, rA, rB, rB,
                                 xA[0], xA[1], xA[2], xB[0], xB[1], xB[2]);
            }
            if constexpr (PART) {
                const float rA = cmp(ra, c), rB = cmp(rb, c);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float mA = PART_M ? cmp(ma[k], c) : xA[k] + uA[k];
                    const float mB = PART_M ? cmp(mb[k], c) : xB[k] + uB[k];
                    racc[k] += vm ? rA * mA : 0.0f;
                    cacc[c][k] += vm ? rB * mB : 0.0f;
                }
                racc[3] += vm ? rA : 0.0f;
                racc[4] += vm ? rA * rA : 0.0f;
                cacc[c][3] += vm ? rB : 0.0f;
                cacc[c][4] += vm ? rB * rB : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (WRITE_X) {
                    const float oA = MOUT ? xA[k] + uA[k] : xA[k];
                    const float oB = MOUT ? xB[k] + uB[k] : xB[k];
                    if constexpr (VAR == 2) {  // masked slots keep their X (or M)
                        cmp(xa[k], c) = vm ? oA : cmp(xa[k], c);
                        cmp(xb[k], c) = vm ? oB : cmp(xb[k], c);
                    } else {
                        cmp(xa[k], c) = oA;
                        cmp(xb[k], c) = oB;
                    }
                }
                if constexpr (WRITE_U) { cmp(ua[k], c) = uA[k]; cmp(ub[k], c) = uB[k]; }
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (WRITE_X) { st4((a.X + k * S) + pa, xa[k]); st4((a.X + k * S) + pb, xb[k]); }
            if constexpr (WRITE_U) { st4((a.U + k * S) + pa, ua[k]); st4((a.U + k * S) + pb, ub[k]); }
        }
        }  // on
        if constexpr (APPLY_U) pr += (double)prow;
        if constexpr (PART) {
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const float v = xsum_cg<G::CG>(racc[q]);
                if ((tid % G::CG) == 0) sh.srow[q][il] = v;
            }
        }
    }
    if constexpr (PART) {
        // column sums: across the row groups of the wave, then across waves
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                float v = cacc[c][q];
#pragma unroll
                for (int o = G::CG; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
                if ((tid & 63) < G::CG) sh.scol[tid >> 6][q][c0 + c] = v;
            }
    }
}

// The unary factors of a diagonal tile: box edge A[c][c] and radius edge
// B[c][c] of circle c = I*T + tid, one lane per circle, run after the pair
// sweep; their partial sums are added to the circle's row sum.  QUAL: the
// circle's container violation and radius (FP64).
//
// The body is a separate (not inlined) function so the pair sweep keeps its
// registers; it takes scalars and pointers by value and returns its sums by
// value: a reference to the kernel's argument block would make the caller
// spill that block to scratch memory in every workgroup.
struct UnaryAcc {
    double pr, qbox, qsum;
};

template <int T, int MODE, bool QUAL, bool UNI>
__device__ __noinline__ UnaryAcc diag_unary_body(float* X, float* U, const float* rho, const float* alpha,
                                                 const float* Nb, const float* Mb, const float* Zv,
                                                 const float* aexp, const long S, const int N, const float rho_u,
                                                 const float alpha_u, const int t, const int I, float* srow,
                                                 const int cl)
{
    constexpr int TT = T * T;
    constexpr bool MIN = (MODE == DM_M || MODE == DM_M_EXIT);
    constexpr bool MOUT = (MODE == DM_M_ENTER || MODE == DM_M);
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH || MODE == DM_M_ENTER);
    constexpr bool PROX = (MODE <= DM_LEGACY_X || MODE >= DM_M_ENTER);
    constexpr bool NBUF = (MODE == DM_FUSED_NBUF || MODE == DM_LEGACY_X);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z || MODE >= DM_M_ENTER);
    constexpr bool PART_M = (MODE == DM_LEGACY_Z);
    constexpr bool WRITE_U = (MODE == DM_FUSED || MODE == DM_FINISH || MODE == DM_M_EXIT);
    UnaryAcc acc = {0.0, 0.0, 0.0};
    const int c = I * T + cl;
    if (c >= N) return acc;
    const long p[2] = {(long)t * (2 * TT) + (long)cl * T + cl, (long)t * (2 * TT) + TT + (long)cl * T + cl};
    float z[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) z[k] = Zv[3 * c + k];
    if constexpr (QUAL) {
        const double cx = z[0], cy = z[1], r = z[2];
        const double v[5] = {r - cx, cx + r - 1.0, r - cy, cy + r - 1.0, -r};
#pragma unroll
        for (int k = 0; k < 5; ++k) acc.qbox = v[k] > acc.qbox ? v[k] : acc.qbox;
        acc.qsum = r;
    }
    float part[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float prs = 0.0f;
#pragma unroll
    for (int side = 0; side < 2; ++side) {  // 0: box (A), 1: radius (B)
        const long q = p[side];
        const float r = UNI ? rho_u : rho[q];
        float u[3], x[3], n[3], m[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (MIN) {
                x[k] = X[k * S + q];  // M
                u[k] = x[k] - z[k];
            } else {
                u[k] = U[k * S + q];
                if constexpr (APPLY_U) x[k] = X[k * S + q];
            }
        }
        if constexpr (APPLY_U) {
            const float al = UNI ? alpha_u : alpha[q];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float d = x[k] - z[k];
                u[k] = u_update(u[k], x[k], z[k], al, d);
                prs += d * d;
            }
        }
        if constexpr (WRITE_U)
#pragma unroll
            for (int k = 0; k < 3; ++k) U[k * S + q] = u[k];
        if constexpr (PROX) {
#pragma unroll
            for (int k = 0; k < 3; ++k) n[k] = NBUF ? Nb[k * S + q] : z[k] - u[k];
            if (side == 0) {
                box3(x[0], x[1], x[2], n[0], n[1], n[2]);
            } else {
                x[0] = n[0];
                x[1] = n[1];
                x[2] = max_alpha_radius(n[2], r, aexp[c]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) X[k * S + q] = MOUT ? x[k] + u[k] : x[k];
        }
        if constexpr (PART) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[k] = PART_M ? Mb[k * S + q] : x[k] + u[k];
                part[k] += r * m[k];
            }
            part[3] += r;
            part[4] += r * r;
        }
    }
    if constexpr (APPLY_U) acc.pr = (double)prs;
    if constexpr (PART) {
#pragma unroll
        for (int k = 0; k < 5; ++k) srow[k * T + cl] += part[k];
    }
    return acc;
}

template <int T, int MODE, bool QUAL, bool UNI = false>
__device__ __forceinline__ void diag_unary(const DArgs& a, const int t, const int I, TileShared<T>& sh, double& pr,
                                           double& qbox, double& qsum)
{
    const UnaryAcc r = diag_unary_body<T, MODE, QUAL, UNI>(a.X, a.U, a.rho, a.alpha, a.Nb, a.Mb, a.Z, a.aexp, a.stride,
                                                           a.N, a.rho_u, a.alpha_u, t, I, &sh.srow[0][0],
                                                           (int)threadIdx.x);
    pr += r.pr;
    qbox = r.qbox > qbox ? r.qbox : qbox;
    qsum += r.qsum;
}

__device__ __forceinline__ double wmax64d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ double wsum64d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One whole tile: pairs, the diagonal's unary factors, the circles' block
// partials (P[q][K][c]: sum over the tile's edges of circle c in block K),
// the tile's FP64 primal partial (*rp) and, with QUAL, its quality partials
// q3 = {max pair overlap, max container violation, sum of radii}.
// Called by all 256 threads of a workgroup; ends with a barrier-free
// epilogue after one __syncthreads, so consecutive calls need a barrier in
// between (the caller's loop provides it).
template <int T, int MODE, bool QUAL, bool UNI = false>
__device__ __forceinline__ void tile_pass(const DArgs& a, const int t, const int I, const int J, TileShared<T>& sh,
                                          double* rp, double* q3)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH || MODE == DM_M_ENTER);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z || MODE >= DM_M_ENTER);
    const int tid = threadIdx.x;
    double pr = 0.0, qov = 0.0, qbox = 0.0, qsum = 0.0;
    if (I == J) {
        pair_tile<T, MODE, 2, QUAL, UNI>(a, t, I, J, sh, pr, qov);
        __syncthreads();  // row sums are in srow
        if (tid < T) diag_unary<T, MODE, QUAL, UNI>(a, t, I, sh, pr, qbox, qsum);
    } else if ((J + 1) * T > a.N) {
        pair_tile<T, MODE, 1, QUAL, UNI>(a, t, I, J, sh, pr, qov);
    } else {
        pair_tile<T, MODE, 0, QUAL, UNI>(a, t, I, J, sh, pr, qov);
    }

    if constexpr (APPLY_U) {
        // FP64 primal partial: wave butterfly, then the four waves in order
        pr = wsum64d(pr);
        if ((tid & 63) == 0) sh.sred[tid >> 6][0] = pr;
    }
    if constexpr (QUAL) {
        qov = wmax64d(qov);
        qbox = wmax64d(qbox);
        qsum = wsum64d(qsum);
        if ((tid & 63) == 0) {
            sh.sred[tid >> 6][1] = qov;
            sh.sred[tid >> 6][2] = qbox;
        }
    }
    __syncthreads();
    if constexpr (PART) {
        // P[q][K][c]: the T partials of one block are contiguous
        const int N = a.N, nb = a.nb;
        if (I != J) {
            if (tid < T) {
                const int i = I * T + tid;
                if (i < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q) a.P[((long)q * nb + J) * N + i] = sh.srow[q][tid];
            } else if (tid < 2 * T) {
                const int jl = tid - T, j = J * T + jl;
                if (j < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                        a.P[((long)q * nb + I) * N + j] =
                            sh.scol[0][q][jl] + sh.scol[1][q][jl] + sh.scol[2][q][jl] + sh.scol[3][q][jl];
            }
        } else if (tid < T) {
            const int c = I * T + tid;
            if (c < N)
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    a.P[((long)q * nb + I) * N + c] =
                        sh.srow[q][tid] + (sh.scol[0][q][tid] + sh.scol[1][q][tid] + sh.scol[2][q][tid] + sh.scol[3][q][tid]);
        }
    }
    if constexpr (APPLY_U) {
        if (tid == 0) *rp = (sh.sred[0][0] + sh.sred[1][0]) + (sh.sred[2][0] + sh.sred[3][0]);
    }
    if constexpr (QUAL) {
        // radii come from diag_unary's lanes (< T <= 64: wave 0), whose
        // butterfly sum thread 0 holds
        if (tid == 0) {
            double o = sh.sred[0][1], b = sh.sred[0][2];
            for (int w = 1; w < 4; ++w) {
                o = sh.sred[w][1] > o ? sh.sred[w][1] : o;
                b = sh.sred[w][2] > b ? sh.sred[w][2] : b;
            }
            q3[0] = o;
            q3[1] = b;
            q3[2] = qsum;
        }
    }
}


// ---- one wavefront per 32 x 32 tile --------------------------------------
//
// The same pass for tile side 32 done by a single wavefront, so the waves of
// a workgroup work on different tiles with no workgroup barrier between
// them (the batched solver hands tiles to waves).  Lane l owns columns
// 4 (l % 8) .. +3 of rows 4 (l / 8) .. +3: the per-lane work of the
// 64-circle workgroup tile.  Row and column partials meet in the wave's own
// LDS scratch.
constexpr int WT = 32;
constexpr int WTT = WT * WT;

struct WaveShared {
    float srow[5][WT];
    float scol[5][WT];
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

template <int MODE, int VAR, bool QUAL, bool UNI>
__device__ __forceinline__ void wave_pair_tile(const DArgs& a, const int t, const int I, const int J, WaveShared& ws,
                                               double& pr, double& qov)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH);
    constexpr bool PROX = (MODE <= DM_LEGACY_X);
    constexpr bool NBUF = (MODE == DM_FUSED_NBUF || MODE == DM_LEGACY_X);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z);
    constexpr bool PART_M = (MODE == DM_LEGACY_Z);
    constexpr bool WRITE_U = APPLY_U;
    constexpr bool WRITE_X = PROX;
    constexpr bool NEED_X = APPLY_U || (VAR == 2 && WRITE_X);
    constexpr bool NEED_U = APPLY_U || (PROX && !NBUF) || (PART && !PART_M);
    constexpr bool NEED_A = APPLY_U;

    const int N = a.N;
    const long S = a.stride;
    const int lane = threadIdx.x & 63;
    const int rg = lane >> 3;         // rows 4 rg .. 4 rg + 3
    const int c0 = (lane & 7) * 4;    // columns c0 .. c0 + 3
    const int base = t * (2 * WTT);

    float zj[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = min(J * WT + c0 + c, N - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) zj[c][k] = a.Z[3 * j + k];
    }
    float cacc[4][5];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 5; ++q) cacc[c][q] = 0.0f;

#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
        const int il = rg * 4 + r;
        const int i = I * WT + il;
        const int ic = min(i, N - 1);
        float zi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) zi[k] = a.Z[3 * ic + k];
        const int pa = base + il * WT + c0;
        const int pb = pa + WTT;
        float racc[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        float prow = 0.0f;
        bool on = true;
        if constexpr (VAR == 1) on = J * WT + c0 < N;
        if constexpr (VAR == 2) on = (c0 + 3 > il) && (J * WT + max(c0, il + 1) < N);
        if (on) {
            float4 xa[3], xb[3], ua[3], ub[3], na[3], nbv[3], ma[3], mb[3];
            float4 ra, rb, aa, ab;
            if constexpr (UNI) {
                ra = rb = make_float4(a.rho_u, a.rho_u, a.rho_u, a.rho_u);
                if constexpr (NEED_A) aa = ab = make_float4(a.alpha_u, a.alpha_u, a.alpha_u, a.alpha_u);
            } else {
                ra = ld4(a.rho + pa);
                rb = ld4(a.rho + pb);
                if constexpr (NEED_A) { aa = ld4(a.alpha + pa); ab = ld4(a.alpha + pb); }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (NEED_X) { xa[k] = ld4((a.X + k * S) + pa); xb[k] = ld4((a.X + k * S) + pb); }
                if constexpr (NEED_U) { ua[k] = ld4((a.U + k * S) + pa); ub[k] = ld4((a.U + k * S) + pb); }
                if constexpr (NBUF) { na[k] = ld4((a.Nb + k * S) + pa); nbv[k] = ld4((a.Nb + k * S) + pb); }
                if constexpr (PART_M) { ma[k] = ld4((a.Mb + k * S) + pa); mb[k] = ld4((a.Mb + k * S) + pb); }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int jl = c0 + c;
                bool vm = true;
                if constexpr (VAR >= 1) vm = (i < N) && (J * WT + jl < N);
                if constexpr (VAR == 2) vm = vm && (il < jl);
                if constexpr (QUAL) {
                    const double dx = (double)zi[0] - (double)zj[c][0], dy = (double)zi[1] - (double)zj[c][1];
                    const double ov = (double)zi[2] + (double)zj[c][2] - sqrt(dx * dx + dy * dy);
                    qov = (vm && ov > qov) ? ov : qov;
                }
                float xA[3], xB[3], uA[3], uB[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if constexpr (NEED_U) { uA[k] = cmp(ua[k], c); uB[k] = cmp(ub[k], c); }
                }
                if constexpr (APPLY_U) {
                    const float alA = cmp(aa, c), alB = cmp(ab, c);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float dA = cmp(xa[k], c) - zi[k];
                        const float dB = cmp(xb[k], c) - zj[c][k];
                        const float nuA = u_update(uA[k], cmp(xa[k], c), zi[k], alA, dA);
                        const float nuB = u_update(uB[k], cmp(xb[k], c), zj[c][k], alB, dB);
                        uA[k] = vm ? nuA : uA[k];
                        uB[k] = vm ? nuB : uB[k];
                        prow += vm ? dA * dA + dB * dB : 0.0f;
                    }
                }
                if constexpr (PROX) {
                    float nA[3], nB[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        nA[k] = NBUF ? cmp(na[k], c) : zi[k] - uA[k];
                        nB[k] = NBUF ? cmp(nbv[k], c) : zj[c][k] - uB[k];
                        xA[k] = nA[k];
                        xB[k] = nB[k];
                    }
                    const float rA = cmp(ra, c), rB = cmp(rb, c);
                    if (vm)
                        pair_project(nA[0], nA[1], nA[2], nB[0], nB[1], nB[2], rA, rA, rB, rB,
                                     xA[0], xA[1], xA[2], xB[0], xB[1], xB[2]);
                }
                if constexpr (PART) {
                    const float rA = cmp(ra, c), rB = cmp(rb, c);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float mA = PART_M ? cmp(ma[k], c) : xA[k] + uA[k];
                        const float mB = PART_M ? cmp(mb[k], c) : xB[k] + uB[k];
                        racc[k] += vm ? rA * mA : 0.0f;
                        cacc[c][k] += vm ? rB * mB : 0.0f;
                    }
                    racc[3] += vm ? rA : 0.0f;
                    racc[4] += vm ? rA * rA : 0.0f;
                    cacc[c][3] += vm ? rB : 0.0f;
                    cacc[c][4] += vm ? rB * rB : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if constexpr (WRITE_X) {
                        if constexpr (VAR == 2) {
                            cmp(xa[k], c) = vm ? xA[k] : cmp(xa[k], c);
                            cmp(xb[k], c) = vm ? xB[k] : cmp(xb[k], c);
                        } else {
                            cmp(xa[k], c) = xA[k];
                            cmp(xb[k], c) = xB[k];
                        }
                    }
                    if constexpr (WRITE_U) { cmp(ua[k], c) = uA[k]; cmp(ub[k], c) = uB[k]; }
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (WRITE_X) { st4((a.X + k * S) + pa, xa[k]); st4((a.X + k * S) + pb, xb[k]); }
                if constexpr (WRITE_U) { st4((a.U + k * S) + pa, ua[k]); st4((a.U + k * S) + pb, ub[k]); }
            }
        }  // on
        if constexpr (APPLY_U) pr += (double)prow;
        if constexpr (PART) {
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const float v = xsum_cg<8>(racc[q]);
                if ((lane & 7) == 0) ws.srow[q][il] = v;
            }
        }
    }
    if constexpr (PART) {
        // column sums across the 8 row groups of the wave
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                float v = cacc[c][q];
                v += __shfl_xor(v, 8, 64);
                v += __shfl_xor(v, 16, 64);
                v += __shfl_xor(v, 32, 64);
                if (lane < 8) ws.scol[q][c0 + c] = v;
            }
    }
}

// One 32 x 32 tile by one wavefront: pairs, the diagonal's unary factors,
// the block partials, the FP64 primal partial (*rp) and, with QUAL, the
// quality partials q3 = {max pair overlap, max container violation, sum r}.
template <int MODE, bool QUAL, bool UNI>
__device__ __forceinline__ void wave_tile_pass(const DArgs& a, const int t, const int I, const int J, WaveShared& ws,
                                               double* rp, double* q3)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z);
    const int lane = threadIdx.x & 63;
    double pr = 0.0, qov = 0.0, qbox = 0.0, qsum = 0.0;
    if (I == J) wave_pair_tile<MODE, 2, QUAL, UNI>(a, t, I, J, ws, pr, qov);
    else if ((J + 1) * WT > a.N) wave_pair_tile<MODE, 1, QUAL, UNI>(a, t, I, J, ws, pr, qov);
    else wave_pair_tile<MODE, 0, QUAL, UNI>(a, t, I, J, ws, pr, qov);
    wave_sync();
    if (I == J && lane < WT) {
        const UnaryAcc r = diag_unary_body<WT, MODE, QUAL, UNI>(a.X, a.U, a.rho, a.alpha, a.Nb, a.Mb, a.Z, a.aexp,
                                                                 a.stride, a.N, a.rho_u, a.alpha_u, t, I,
                                                                 &ws.srow[0][0], lane);
        pr += r.pr;
        qbox = r.qbox > qbox ? r.qbox : qbox;
        qsum += r.qsum;
    }
    wave_sync();
    if constexpr (PART) {
        const int N = a.N, nb = a.nb;
        if (I != J) {
            if (lane < WT) {
                const int i = I * WT + lane;
                if (i < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q) a.P[((long)q * nb + J) * N + i] = ws.srow[q][lane];
            } else {
                const int jl = lane - WT, j = J * WT + jl;
                if (j < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q) a.P[((long)q * nb + I) * N + j] = ws.scol[q][jl];
            }
        } else if (lane < WT) {
            const int c = I * WT + lane;
            if (c < N)
#pragma unroll
                for (int q = 0; q < 5; ++q) a.P[((long)q * nb + I) * N + c] = ws.srow[q][lane] + ws.scol[q][lane];
        }
    }
    if constexpr (APPLY_U) {
        pr = wsum64d(pr);
        if (lane == 0) *rp = pr;
    }
    if constexpr (QUAL) {
        qov = wmax64d(qov);
        qbox = wmax64d(qbox);
        qsum = wsum64d(qsum);
        if (lane == 0) {
            q3[0] = qov;
            q3[1] = qbox;
            q3[2] = qsum;
        }
    }
    wave_sync();  // the wave's scratch is reused by its next tile
}

}  // namespace pdm

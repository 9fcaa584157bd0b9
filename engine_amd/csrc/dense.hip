// dense.hip -- dense pair-tile layout for the all-pairs unit-square packing.
//
// A graph whose factors are exactly: one box and one radius factor per
// circle and one pair factor for every unordered pair of circles (the
// unit-square packing, dim 3) is stored as tiles of DT x DT circle pairs.
// Tile (I, J), I <= J, holds two sub-blocks of DT*DT edge slots:
//   A[il][jl] = the edge of circle i = I*DT+il in pair (i, j),
//   B[il][jl] = the edge of circle j = J*DT+jl in the same pair,
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
//   k_dense_zreduce: Z_c = sum rho M / sum rho over the circle's DT-block
//     partials in fixed order (atomic-free, deterministic), and the FP64
//     dual residual.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "paradmm_gpu.h"
#include "prox_device.h"

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

namespace pdm {

constexpr int DT = 64;        // circles per tile side
constexpr int DTT = DT * DT;  // edge slots per sub-block

// minimum waves per SIMD requested for the tile kernel (register budget)
#ifndef PDM_PAIRTILE_WAVES
#define PDM_PAIRTILE_WAVES 2
#endif

enum {
    DM_FUSED = 0,          // U update, N = Z - U, prox, partials      (steady state)
    DM_FUSED_CURRENT = 1,  // N = Z - U (U already current), prox, partials
    DM_FUSED_NBUF = 2,     // N from the N buffer, prox, partials       (after upload)
    DM_LEGACY_X = 3,       // N from the N buffer, prox                 (updateX)
    DM_LEGACY_Z = 4,       // partials of the stored M                   (updateZ)
    DM_FINISH = 5          // U update only (+ primal residual)
};

struct DenseState {
    int N = 0, nb = 0, nt = 0;
    int2* d_tiles = nullptr;  // [nt] (I, J)
    float* d_P = nullptr;     // [5][N][nb] partial sums
    double* d_Rp = nullptr;   // [nt] primal residual partials
    double* d_Rd = nullptr;   // [N] dual residual partials
    float* d_aexp = nullptr;  // [N] radius exponent per circle
    bool u_pending = false;   // U still lacks the update of the last iteration
};

struct DArgs {
    int N, nb;
    long stride;
    float* X;  // [3][stride]
    float* U;  // [3][stride]
    const float* rho;
    const float* alpha;
    const float* Nb;  // [3][stride]
    const float* Mb;  // [3][stride]
    const float* Z;   // [N][3]
    float* P;
    double* Rp;
    const float* aexp;
    const int2* tiles;
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

__device__ __forceinline__ float xsum16(float v)
{
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Body of one tile.  VAR 0: interior tile (no masks), 1: tile touching the
// padded end (circle masks), 2: diagonal tile (masks, unary diagonal).
template <int MODE, int VAR>
__device__ __forceinline__ void pair_tile(const DArgs& a, const int t, const int I, const int J,
                                          float (&srow)[5][DT], float (&scol)[4][5][DT], double& pr)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH);
    constexpr bool PROX = (MODE <= DM_LEGACY_X);
    constexpr bool NBUF = (MODE == DM_FUSED_NBUF || MODE == DM_LEGACY_X);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z);
    constexpr bool PART_M = (MODE == DM_LEGACY_Z);
    constexpr bool WRITE_U = APPLY_U;
    constexpr bool WRITE_X = PROX;
    // a diagonal tile's vector stores also cover its unary slots (processed
    // afterwards by diag_unary, which reads their old X) and its padding:
    // X is loaded there so masked slots are written back unchanged
    constexpr bool NEED_X = APPLY_U || (VAR == 2 && WRITE_X);
    constexpr bool NEED_U = APPLY_U || (PROX && !NBUF) || (PART && !PART_M);
    constexpr bool NEED_A = APPLY_U;

    const int N = a.N;
    const long S = a.stride;
    const int tid = threadIdx.x;
    const int rg = tid >> 4;      // row group: rows 4*rg .. 4*rg+3
    const int c0 = (tid & 15) * 4;  // first of this thread's 4 columns
    const int base = t * (2 * DTT);  // slots fit in 31 bits (checked at finalize)

    float zj[4][3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = min(J * DT + c0 + c, N - 1);
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
        const int i = I * DT + il;
        const int ic = min(i, N - 1);
        float zi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) zi[k] = a.Z[3 * ic + k];
        const int pa = base + il * DT + c0;  // 32-bit lane offsets off uniform bases
        const int pb = pa + DTT;

        float4 xa[3], xb[3], ua[3], ub[3], na[3], nbv[3], ma[3], mb[3];
        float4 ra, rb, aa, ab;
        ra = ld4(a.rho + pa);
        rb = ld4(a.rho + pb);
        if constexpr (NEED_A) { aa = ld4(a.alpha + pa); ab = ld4(a.alpha + pb); }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (NEED_X) { xa[k] = ld4((a.X + k * S) + pa); xb[k] = ld4((a.X + k * S) + pb); }
            if constexpr (NEED_U) { ua[k] = ld4((a.U + k * S) + pa); ub[k] = ld4((a.U + k * S) + pb); }
            if constexpr (NBUF) { na[k] = ld4((a.Nb + k * S) + pa); nbv[k] = ld4((a.Nb + k * S) + pb); }
            if constexpr (PART_M) { ma[k] = ld4((a.Mb + k * S) + pa); mb[k] = ld4((a.Mb + k * S) + pb); }
        }

        float racc[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // squared primal differences of this row's 4 pairs (<= 24 terms) in
        // fp32; the accumulation across rows, threads and tiles is FP64
        float prow = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int jl = c0 + c;
            // pair slots only; the diagonal's unary factors run in diag_unary
            bool vm = true;
            if constexpr (VAR >= 1) vm = (i < N) && (J * DT + jl < N);
            if constexpr (VAR == 2) vm = vm && (il < jl);
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
                    const float nuA = uA[k] + alA * dA;
                    const float nuB = uB[k] + alB * dB;
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
                    if constexpr (VAR == 2) {  // masked slots keep their X
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
        if constexpr (APPLY_U) pr += (double)prow;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if constexpr (WRITE_X) { st4((a.X + k * S) + pa, xa[k]); st4((a.X + k * S) + pb, xb[k]); }
            if constexpr (WRITE_U) { st4((a.U + k * S) + pa, ua[k]); st4((a.U + k * S) + pb, ub[k]); }
        }
        if constexpr (PART) {
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const float v = xsum16(racc[q]);
                if ((tid & 15) == 0) srow[q][il] = v;
            }
        }
    }
    if constexpr (PART) {
        // column sums: across the 4 row groups of the wave, then across waves
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                float v = cacc[c][q];
                v += __shfl_xor(v, 16, 64);
                v += __shfl_xor(v, 32, 64);
                if ((tid & 63) < 16) scol[tid >> 6][q][c0 + c] = v;
            }
    }
}

// The unary factors of a diagonal tile: box edge A[c][c] and radius edge
// B[c][c] of circle c = I*DT + tid, one lane per circle, run after the pair
// sweep; their partial sums are added to the circle's row sum.
template <int MODE>
__device__ __noinline__ void diag_unary(const DArgs& a, const int t, const int I, float (&srow)[5][DT], double& pr)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH);
    constexpr bool PROX = (MODE <= DM_LEGACY_X);
    constexpr bool NBUF = (MODE == DM_FUSED_NBUF || MODE == DM_LEGACY_X);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z);
    constexpr bool PART_M = (MODE == DM_LEGACY_Z);
    const int cl = threadIdx.x;
    const int c = I * DT + cl;
    if (c >= a.N) return;
    const long S = a.stride;
    const long p[2] = {(long)t * (2 * DTT) + (long)cl * DT + cl, (long)t * (2 * DTT) + DTT + (long)cl * DT + cl};
    float z[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) z[k] = a.Z[3 * c + k];
    float part[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float prs = 0.0f;
#pragma unroll
    for (int side = 0; side < 2; ++side) {  // 0: box (A), 1: radius (B)
        const long q = p[side];
        const float r = a.rho[q];
        float u[3], x[3], n[3], m[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = a.U[k * S + q];
            if constexpr (APPLY_U) x[k] = a.X[k * S + q];
        }
        if constexpr (APPLY_U) {
            const float al = a.alpha[q];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float d = x[k] - z[k];
                u[k] += al * d;
                prs += d * d;
                a.U[k * S + q] = u[k];
            }
        }
        if constexpr (PROX) {
#pragma unroll
            for (int k = 0; k < 3; ++k) n[k] = NBUF ? a.Nb[k * S + q] : z[k] - u[k];
            if (side == 0) {
                box3(x[0], x[1], x[2], n[0], n[1], n[2]);
            } else {
                x[0] = n[0];
                x[1] = n[1];
                x[2] = max_alpha_radius(n[2], r, a.aexp[c]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) a.X[k * S + q] = x[k];
        }
        if constexpr (PART) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                m[k] = PART_M ? a.Mb[k * S + q] : x[k] + u[k];
                part[k] += r * m[k];
            }
            part[3] += r;
            part[4] += r * r;
        }
    }
    if constexpr (APPLY_U) pr += (double)prs;
    if constexpr (PART) {
#pragma unroll
        for (int k = 0; k < 5; ++k) srow[k][cl] += part[k];
    }
}

template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PDM_PAIRTILE_WAVES))) void k_pairtile(DArgs a)
{
    constexpr bool APPLY_U = (MODE == DM_FUSED || MODE == DM_FINISH);
    constexpr bool PART = (MODE <= DM_FUSED_NBUF || MODE == DM_LEGACY_Z);
    __shared__ float srow[5][DT];
    __shared__ float scol[4][5][DT];
    __shared__ double sred[4];

    const int t = blockIdx.x;
    const int2 ij = a.tiles[t];
    const int I = ij.x, J = ij.y;
    const int tid = threadIdx.x;
    double pr = 0.0;
    if (I == J) {
        pair_tile<MODE, 2>(a, t, I, J, srow, scol, pr);
        __syncthreads();  // row sums are in srow
        if (tid < DT) diag_unary<MODE>(a, t, I, srow, pr);
    } else if ((J + 1) * DT > a.N) {
        pair_tile<MODE, 1>(a, t, I, J, srow, scol, pr);
    } else {
        pair_tile<MODE, 0>(a, t, I, J, srow, scol, pr);
    }

    if constexpr (APPLY_U) {
        // FP64 primal partial: wave butterfly, then the four waves in order
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) pr += __shfl_xor(pr, o, 64);
        if ((tid & 63) == 0) sred[tid >> 6] = pr;
    }
    __syncthreads();
    if constexpr (PART) {
        const int N = a.N, nb = a.nb;
        if (I != J) {
            if (tid < DT) {
                const int i = I * DT + tid;
                if (i < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q) a.P[((long)q * N + i) * nb + J] = srow[q][tid];
            } else if (tid < 2 * DT) {
                const int jl = tid - DT, j = J * DT + jl;
                if (j < N)
#pragma unroll
                    for (int q = 0; q < 5; ++q)
                        a.P[((long)q * N + j) * nb + I] = scol[0][q][jl] + scol[1][q][jl] + scol[2][q][jl] + scol[3][q][jl];
            }
        } else if (tid < DT) {
            const int c = I * DT + tid;
            if (c < N)
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    a.P[((long)q * N + c) * nb + I] =
                        srow[q][tid] + (scol[0][q][tid] + scol[1][q][tid] + scol[2][q][tid] + scol[3][q][tid]);
        }
    }
    if constexpr (APPLY_U) {
        if (tid == 0) a.Rp[t] = (sred[0] + sred[1]) + (sred[2] + sred[3]);
    }
}

__device__ __forceinline__ float wsum64(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Z_c = sum rho M / sum rho over the circle's nb block partials; one
// wavefront per circle, lanes over blocks in fixed order.
__global__ __launch_bounds__(256) void k_dense_zreduce(int N, int nb, const float* __restrict__ P, float* __restrict__ Z,
                                                       float* __restrict__ Zprev, double* __restrict__ Rd)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= N) return;
    float s[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int K = lane; K < nb; K += 64)
#pragma unroll
        for (int q = 0; q < 5; ++q) s[q] += P[((long)q * N + c) * nb + K];
#pragma unroll
    for (int q = 0; q < 5; ++q) s[q] = wsum64(s[q]);
    double dz2 = 0.0;
    float z[3], zo[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        z[k] = s[k] / s[3];
        zo[k] = Z[3 * c + k];
        const double d = (double)(z[k] - zo[k]);
        dz2 += d * d;
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Z[3 * c + k] = z[k];
            Zprev[3 * c + k] = zo[k];
        }
        Rd[c] = (double)s[4] * dz2;
    }
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

// ------------------------------------------------------------------ host side

static inline long tile_index(int I, int J, int nb) { return (long)I * nb - (long)I * (I - 1) / 2 + (J - I); }

int dense_try_finalize(DevGraph* s, graph* G, const std::vector<int>& fkind)
{
    if (G->num_dims != 3) return -1;
    const int N = G->num_vars, E = G->num_edges, F = G->num_opts;
    if (N < 2) return -1;
    const long npairs = (long)N * (N - 1) / 2;
    if ((long)F != npairs + 2L * N || (long)E != 2 * npairs + 2L * N) return -1;
    std::vector<int> box(N, -1), rad(N, -1);
    std::vector<float> aexp(N, 1.0f);
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

    const int nb = (N + DT - 1) / DT;
    const int nt = (int)((long)nb * (nb + 1) / 2);
    const long stride = (long)nt * 2 * DTT;
    if (stride > 0x7fffffffL) {
        pdm_set_error(PARADMM_ERR_SHAPE, "dense layout too large (%ld slots)", stride);
        return PARADMM_ERR_SHAPE;
    }
    std::vector<int> pos(E, -1);
    for (int f : pair_f) {
        const int* el = G->operatorix_to_edgelistix[f];
        const int i = ev[el[0]], j = ev[el[1]];
        const long b = tile_index(i / DT, j / DT, nb) * 2 * DTT + (long)(i % DT) * DT + (j % DT);
        pos[el[0]] = (int)b;
        pos[el[1]] = (int)(b + DTT);
    }
    for (int c = 0; c < N; ++c) {
        const long b = tile_index(c / DT, c / DT, nb) * 2 * DTT + (long)(c % DT) * DT + (c % DT);
        pos[box[c]] = (int)b;
        pos[rad[c]] = (int)(b + DTT);
    }
    std::vector<int2> tiles(nt);
    for (int I = 0, t = 0; I < nb; ++I)
        for (int J = I; J < nb; ++J, ++t) tiles[t] = make_int2(I, J);

    DenseState* d = new DenseState();
    s->dense = d;
    d->N = N;
    d->nb = nb;
    d->nt = nt;
    s->stride = stride;
    s->npos = stride;
    RC(dmalloc(s, (void**)&s->d_pos, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&s->d_edge_var_api, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&d->d_tiles, sizeof(int2) * (size_t)nt));
    RC(dmalloc(s, (void**)&d->d_aexp, sizeof(float) * (size_t)N));
    CK(hipMemcpy(s->d_pos, pos.data(), sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_edge_var_api, ev, sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(d->d_tiles, tiles.data(), sizeof(int2) * (size_t)nt, hipMemcpyHostToDevice));
    CK(hipMemcpy(d->d_aexp, aexp.data(), sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
    return 0;
}

int dense_alloc_state(DevGraph* s)
{
    DenseState* d = s->dense;
    RC(dmalloc(s, (void**)&d->d_P, sizeof(float) * 5 * (size_t)d->N * d->nb));
    RC(dmalloc(s, (void**)&d->d_Rp, sizeof(double) * (size_t)d->nt));
    RC(dmalloc(s, (void**)&d->d_Rd, sizeof(double) * (size_t)d->N));
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
    default: return "pairtile_finish";
    }
}

static int launch_pairtile(DevGraph* s, int mode)
{
    DArgs a = dargs(s);
    const unsigned nt = (unsigned)s->dense->nt;
    prof_begin(s, mode_name(mode));
    switch (mode) {
    case DM_FUSED: k_pairtile<DM_FUSED><<<nt, 256, 0, s->stream>>>(a); break;
    case DM_FUSED_CURRENT: k_pairtile<DM_FUSED_CURRENT><<<nt, 256, 0, s->stream>>>(a); break;
    case DM_FUSED_NBUF: k_pairtile<DM_FUSED_NBUF><<<nt, 256, 0, s->stream>>>(a); break;
    case DM_LEGACY_X: k_pairtile<DM_LEGACY_X><<<nt, 256, 0, s->stream>>>(a); break;
    case DM_LEGACY_Z: k_pairtile<DM_LEGACY_Z><<<nt, 256, 0, s->stream>>>(a); break;
    default: k_pairtile<DM_FINISH><<<nt, 256, 0, s->stream>>>(a); break;
    }
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

static int launch_zreduce(DevGraph* s)
{
    DenseState* d = s->dense;
    prof_begin(s, "zreduce");
    k_dense_zreduce<<<(unsigned)((d->N + 3) / 4), 256, 0, s->stream>>>(d->N, d->nb, d->d_P, s->d_Z, s->d_Zprev, d->d_Rd);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

int dense_iteration(DevGraph* s, int from_n)
{
    DenseState* d = s->dense;
    const int mode = from_n ? DM_FUSED_NBUF : (d->u_pending ? DM_FUSED : DM_FUSED_CURRENT);
    RC(launch_pairtile(s, mode));
    RC(launch_zreduce(s));
    d->u_pending = true;
    return PARADMM_OK;
}

bool dense_steady(DevGraph* s) { return s->dense && s->dense->u_pending; }

void dense_reset(DevGraph* s)
{
    if (s->dense) s->dense->u_pending = false;
}

int dense_finish(DevGraph* s)
{
    DenseState* d = s->dense;
    if (!d || !d->u_pending) return PARADMM_OK;
    RC(launch_pairtile(s, DM_FINISH));
    d->u_pending = false;
    return PARADMM_OK;
}

int dense_legacy_xpass(DevGraph* s) { return launch_pairtile(s, DM_LEGACY_X); }

int dense_legacy_zpass(DevGraph* s)
{
    RC(launch_pairtile(s, DM_LEGACY_Z));
    return launch_zreduce(s);
}

int dense_residual_reduce(DevGraph* s)
{
    RC(dense_finish(s));
    DenseState* d = s->dense;
    k_dense_res_final<<<1, 256, 0, s->stream>>>(d->nt, d->d_Rp, d->N, d->d_Rd, s->d_res);
    CK(hipGetLastError());
    return PARADMM_OK;
}

}  // namespace pdm

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

int dense_iteration(DevGraph* s, int from_n, int mphase)
{
    DenseState* d = s->dense;
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
    }
}

int dense_finish(DevGraph* s)
{
    DenseState* d = s->dense;
    if (!d || !d->u_pending) return PARADMM_OK;
    if (d->m_state) {
        pdm_set_error(PARADMM_ERR_STATE, "dense: finish inside the M state");
        return PARADMM_ERR_STATE;
    }
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

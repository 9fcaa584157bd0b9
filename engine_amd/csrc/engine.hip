// engine.hip -- device state, layout finalisation, kernels and the C ABI of
// the MI355X parADMM engine.
//
// One iteration of the reference (core/paradmm.c:736-741) is
//   X = prox_f(N) per factor, M = X + U, Z_v = sum rho M / sum rho,
//   U += alpha (X - Z_v), N = Z_v - U.
// Device state is structure-of-arrays fp32: X[k][p], U[k][p] (k < num_dims,
// p the device position of an edge), rho[p], alpha[p], Z[v][k].
//
// Two layouts:
//  * generic (this file): edges stored variable-major (CSR order), so the Z
//    consensus is a contiguous, atomic-free segmented reduction, one
//    wavefront per variable, with M = X + U and the U update fused into it;
//    N = Z - U is recomputed inside the X pass, which runs one thread per
//    factor per kind (static dispatch, no function pointers).
//  * dense pair tiles (dense.hip): the all-pairs unit-square packing.
//
// Legacy update calls (legacy.c) run single passes with host coherence;
// paradmm_gpu_iterate runs fused iterations captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "paradmm_gpu.h"
#include "prox_device.h"

using namespace pdm;

// ------------------------------------------------------------------ errors

static thread_local char g_err[1024];

extern "C" const char* paradmm_last_error(void) { return g_err; }

extern "C" void pdm_set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int n = snprintf(g_err, sizeof g_err, "[%d] ", code);
    vsnprintf(g_err + n, sizeof g_err - n, fmt, ap);
    va_end(ap);
}

extern "C" void pdm_legacy_fail(int code)
{
    fprintf(stderr, "paradmm: %s\n", g_err);
    const char* e = getenv("PARADMM_NO_EXIT");
    if (e && e[0] == '1') return;
    (void)code;
    exit(EXIT_FAILURE);
}

namespace pdm {
int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    pdm_set_error(PARADMM_ERR_HIP, "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
    return PARADMM_ERR_HIP;
}
}  // namespace pdm

// ------------------------------------------------------------------ kind registry

static std::mutex g_kind_mu;
static std::unordered_map<void*, int> g_kind_of;

static void register_builtins_locked()
{
    static bool done = false;
    if (done) return;
    done = true;
    g_kind_of[(void*)&paradmm_prox_disk_collision_2d] = PARADMM_KIND_PAIR_2D;
    g_kind_of[(void*)&paradmm_prox_disk_plane_2d] = PARADMM_KIND_WALL_2D;
    g_kind_of[(void*)&paradmm_prox_max_radius_2d] = PARADMM_KIND_RADIUS_2D;
    g_kind_of[(void*)&paradmm_prox_disk_collision_3d] = PARADMM_KIND_PAIR_3D;
    g_kind_of[(void*)&paradmm_prox_box_3d] = PARADMM_KIND_BOX_3D;
    g_kind_of[(void*)&paradmm_prox_max_radius_3d] = PARADMM_KIND_RADIUS_3D;
}

// The reference demo's operators (packing/host.c:31, 67, 83), found by name
// through the dynamic symbol table when the caller exports its symbols.
static int kind_by_symbol(void* fn)
{
    Dl_info info;
    if (!dladdr(fn, &info) || !info.dli_sname || info.dli_saddr != fn) return PARADMM_KIND_NONE;
    const char* s = info.dli_sname;
    if (!strcmp(s, "disk_collision_variable_radius")) return PARADMM_KIND_PAIR_2D;
    if (!strcmp(s, "disk_plane_collision_variable_radius")) return PARADMM_KIND_WALL_2D;
    if (!strcmp(s, "maximize_sum_of_squared_disk_radii_alpha")) return PARADMM_KIND_RADIUS_2D;
    return PARADMM_KIND_NONE;
}

extern "C" int paradmm_register_prox_kind(paradmm_prox_fn fn, int kind)
{
    if (!fn || kind <= PARADMM_KIND_NONE || kind >= PARADMM_KIND_COUNT) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_register_prox_kind: bad callback or kind %d", kind);
        return PARADMM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_kind_mu);
    register_builtins_locked();
    g_kind_of[(void*)fn] = kind;
    return PARADMM_OK;
}

extern "C" int paradmm_prox_kind_of(paradmm_prox_fn fn)
{
    std::lock_guard<std::mutex> lk(g_kind_mu);
    register_builtins_locked();
    auto it = g_kind_of.find((void*)fn);
    if (it != g_kind_of.end()) return it->second;
    int k = kind_by_symbol((void*)fn);
    if (k != PARADMM_KIND_NONE) g_kind_of[(void*)fn] = k;
    return k;
}

namespace pdm {
const KindInfo& kind_info(int kind)
{
    // degree (0 = any, each edge independent), dims, float params per item
    static const KindInfo t[PARADMM_KIND_COUNT] = {
        {0, 0, 0, "none"},
        {4, 2, 0, "disk_collision_2d"},
        {2, 2, 4, "disk_plane_2d"},
        {0, 2, 1, "max_radius_2d"},
        {2, 3, 0, "disk_collision_3d"},
        {0, 3, 0, "box_3d"},
        {0, 3, 1, "max_radius_3d"},
    };
    return t[kind];
}
}  // namespace pdm

// ------------------------------------------------------------------ kernels (generic layout)

namespace {

template <int K> struct KT;
template <> struct KT<PARADMM_KIND_PAIR_2D> { enum { DEG = 4, D = 2, P = 0 }; };
template <> struct KT<PARADMM_KIND_WALL_2D> { enum { DEG = 2, D = 2, P = 4 }; };
template <> struct KT<PARADMM_KIND_RADIUS_2D> { enum { DEG = 1, D = 2, P = 1 }; };
template <> struct KT<PARADMM_KIND_PAIR_3D> { enum { DEG = 2, D = 3, P = 0 }; };
template <> struct KT<PARADMM_KIND_BOX_3D> { enum { DEG = 1, D = 3, P = 0 }; };
template <> struct KT<PARADMM_KIND_RADIUS_3D> { enum { DEG = 1, D = 3, P = 1 }; };

template <int K, int DEG, int D, int P>
__device__ __forceinline__ void apply_prox(float (&x)[DEG][D], const float (&n)[DEG][D], const float (&rho)[DEG],
                                           const float (&prm)[P > 0 ? P : 1])
{
    if constexpr (K == PARADMM_KIND_PAIR_2D) prox_pair2(x, n, rho);
    else if constexpr (K == PARADMM_KIND_WALL_2D) prox_wall2(x, n, rho, prm);
    else if constexpr (K == PARADMM_KIND_RADIUS_2D) prox_radius2(x, n, rho, prm);
    else if constexpr (K == PARADMM_KIND_PAIR_3D) prox_pair3(x, n, rho);
    else if constexpr (K == PARADMM_KIND_BOX_3D) prox_box3(x, n);
    else prox_radius3(x, n, rho, prm);
}

struct XArgs {
    int nitems;
    const int* items;     // [DEG][nitems] device edge positions
    const float* params;  // [nitems][P]
    const int* edge_var;  // device position -> variable
    const float* Z;       // [V][D]
    const float* U;       // [D][stride]
    const float* Nbuf;    // [D][stride]
    const float* rho;     // [stride]
    float* X;             // [D][stride]
    long stride;
    int from_n;           // 1: N from the N buffer (legacy pass / first iteration after upload)
};

// X pass: one thread per factor (per edge for unary kinds).  N = Z - U is
// formed on the fly unless the N buffer is authoritative.
template <int K>
__device__ __forceinline__ void xpass_item(const XArgs& a, const int i)
{
    constexpr int DEG = KT<K>::DEG, D = KT<K>::D, P = KT<K>::P;
    int e[DEG];
    float n[DEG][D], x[DEG][D], rho[DEG], prm[P > 0 ? P : 1];
#pragma unroll
    for (int s = 0; s < DEG; ++s) e[s] = a.items[(long)s * a.nitems + i];
#pragma unroll
    for (int s = 0; s < DEG; ++s) rho[s] = a.rho[e[s]];
    if (a.from_n) {
#pragma unroll
        for (int s = 0; s < DEG; ++s)
#pragma unroll
            for (int k = 0; k < D; ++k) n[s][k] = a.Nbuf[k * a.stride + e[s]];
    } else {
#pragma unroll
        for (int s = 0; s < DEG; ++s) {
            const int v = a.edge_var[e[s]];
#pragma unroll
            for (int k = 0; k < D; ++k) n[s][k] = a.Z[v * D + k] - a.U[k * a.stride + e[s]];
        }
    }
#pragma unroll
    for (int p = 0; p < (P > 0 ? P : 1); ++p) prm[p] = P > 0 ? a.params[(long)i * P + p] : 0.0f;
    apply_prox<K, DEG, D, P>(x, n, rho, prm);
#pragma unroll
    for (int s = 0; s < DEG; ++s)
#pragma unroll
        for (int k = 0; k < D; ++k) a.X[k * a.stride + e[s]] = x[s][k];
}

template <int K>
__global__ __launch_bounds__(256) void k_xpass(XArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < a.nitems) xpass_item<K>(a, i);
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct ZArgs {
    int V;
    const int* var_off;  // [V+1] CSR offsets into device positions
    const float* X;
    float* U;
    const float* M;      // legacy pass only
    const float* rho;
    const float* alpha;
    float* Z;
    float* Zprev;
    double* res_var;     // [V][2] primal, dual partials
    long stride;
};

// Fused Z pass: one wavefront per variable.  Sweep 1: sum rho, rho^2 and
// rho (X + U) over the variable's contiguous edge segment; Z = ratio.
// Sweep 2: U += alpha (X - Z) and the FP64 primal residual partial.  The
// sums are per-lane in edge order then a fixed xor butterfly: deterministic.
template <int D>
__global__ __launch_bounds__(256) void k_zpass_fused(ZArgs a)
{
    const int lane = threadIdx.x & 63;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= a.V) return;
    const int b = a.var_off[v], end = a.var_off[v + 1];
    float sr = 0.0f, sr2 = 0.0f, sm[D];
#pragma unroll
    for (int k = 0; k < D; ++k) sm[k] = 0.0f;
    for (int p = b + lane; p < end; p += 64) {
        const float r = a.rho[p];
        sr += r;
        sr2 += r * r;
#pragma unroll
        for (int k = 0; k < D; ++k) sm[k] += r * (a.X[k * a.stride + p] + a.U[k * a.stride + p]);
    }
    sr = wave_sum(sr);
    sr2 = wave_sum(sr2);
    float z[D], zold[D];
    double dz2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        z[k] = wave_sum(sm[k]) / sr;
        zold[k] = a.Z[v * D + k];
        const double t = (double)(z[k] - zold[k]);
        dz2 += t * t;
    }
    double pr = 0.0;
    for (int p = b + lane; p < end; p += 64) {
        const float al = a.alpha[p];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float t = a.X[k * a.stride + p] - z[k];
            a.U[k * a.stride + p] += al * t;
            pr += (double)t * (double)t;
        }
    }
    pr = wave_sum(pr);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            a.Z[v * D + k] = z[k];
            a.Zprev[v * D + k] = zold[k];
        }
        a.res_var[2 * v] = pr;
        a.res_var[2 * v + 1] = (double)sr2 * dz2;
    }
}

// Legacy Z pass (core/paradmm.c:637-672): Z from the stored M.
// Most edges per thread that k_zpass_fused_wg keeps in registers (X and U
// of each: 2 C D floats, about 96 registers); the launch takes the
// smallest of C = 8, 16, .. that covers the longest variable.
template <int D>
struct ZWG {
    static constexpr int C = D == 1 ? 32 : D == 2 ? 24 : 16;
};

// The fused Z pass with one workgroup per variable, its edges' X and U held
// in registers between the two sweeps (the Z sums, then the U update with
// the new Z): X and U are read once (32 B per edge at D = 2, against 48 B
// for k_zpass_fused, which reads them again).  For variables of at most
// 256 C edges.  The sums run per thread over edges b + tid + 256 c, then
// per wavefront and over the four wavefronts in a fixed order: runs are
// bitwise reproducible.
template <int D, int C>
__global__ __launch_bounds__(256) void k_zpass_fused_wg(ZArgs a)
{
    __shared__ float red[4][D + 2];
    __shared__ double redd[4];
    const int v = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int b = a.var_off[v], end = a.var_off[v + 1];
    const long S = a.stride;
    float xs[C][D], us[C][D];
    float sr = 0.0f, sr2 = 0.0f, sm[D];
#pragma unroll
    for (int k = 0; k < D; ++k) sm[k] = 0.0f;
    // every slot loads (a slot past the variable's edges reads edge 0, E >= 1,
    // and takes 0 for its values), so no branch sits around the loads and
    // the C slots' loads are in flight together; the sums add +0 for such a
    // slot, which leaves them unchanged (a sum started at +0 never holds -0)
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int p = b + tid + 256 * c;
        const bool in = p < end;
        const int pc = in ? p : 0;
        const float rl = a.rho[pc];
        float xl[D], ul[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            xl[k] = a.X[k * S + pc];
            ul[k] = a.U[k * S + pc];
        }
        const float r = in ? rl : 0.0f;
        sr += r;
        sr2 += r * r;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            xs[c][k] = in ? xl[k] : 0.0f;
            us[c][k] = in ? ul[k] : 0.0f;
            sm[k] += r * (xs[c][k] + us[c][k]);
        }
    }
    sr = wave_sum(sr);
    sr2 = wave_sum(sr2);
#pragma unroll
    for (int k = 0; k < D; ++k) sm[k] = wave_sum(sm[k]);
    if (lane == 0) {
        red[wv][0] = sr;
        red[wv][1] = sr2;
#pragma unroll
        for (int k = 0; k < D; ++k) red[wv][2 + k] = sm[k];
    }
    __syncthreads();
    const float tr = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    const float tr2 = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
    float z[D], zold[D];
    double dz2 = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        z[k] = ((red[0][2 + k] + red[1][2 + k]) + (red[2][2 + k] + red[3][2 + k])) / tr;
        zold[k] = a.Z[v * D + k];
        const double t = (double)(z[k] - zold[k]);
        dz2 += t * t;
    }
    double pr = 0.0;
    float al[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int p = b + tid + 256 * c;
        al[c] = a.alpha[p < end ? p : 0];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int p = b + tid + 256 * c;
        if (p < end) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float t = xs[c][k] - z[k];
                a.U[k * S + p] = us[c][k] + al[c] * t;
                pr += (double)t * (double)t;
            }
        }
    }
    pr = wave_sum(pr);
    if (lane == 0) redd[wv] = pr;
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            a.Z[v * D + k] = z[k];
            a.Zprev[v * D + k] = zold[k];
        }
        a.res_var[2 * v] = (redd[0] + redd[1]) + (redd[2] + redd[3]);
        a.res_var[2 * v + 1] = (double)tr2 * dz2;
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_zpass_legacy(ZArgs a)
{
    const int lane = threadIdx.x & 63;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= a.V) return;
    const int b = a.var_off[v], end = a.var_off[v + 1];
    float sr = 0.0f, sm[D];
#pragma unroll
    for (int k = 0; k < D; ++k) sm[k] = 0.0f;
    for (int p = b + lane; p < end; p += 64) {
        const float r = a.rho[p];
        sr += r;
#pragma unroll
        for (int k = 0; k < D; ++k) sm[k] += r * a.M[k * a.stride + p];
    }
    sr = wave_sum(sr);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float z = wave_sum(sm[k]) / sr;
        if (lane == 0) {
            a.Zprev[v * D + k] = a.Z[v * D + k];
            a.Z[v * D + k] = z;
        }
    }
}

// M = X + U over all D*stride entries (core/paradmm.c:625-634)
__global__ __launch_bounds__(256) void k_mpass(long n, const float* __restrict__ X, const float* __restrict__ U,
                                               float* __restrict__ M)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) M[i] = X[i] + U[i];
}

struct EArgs {
    int E;
    const int* pos;       // api edge -> device position
    const int* edge_var;  // api edge -> variable
    const float* X;
    float* U;
    float* Nbuf;
    float* M;
    const float* Z;
    const float* alpha;
    long stride;
};

// U += alpha (X - Z_v) (core/paradmm.c:696-715), one thread per edge
template <int D>
__global__ __launch_bounds__(256) void k_upass(EArgs a)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.E) return;
    const int p = a.pos[e], v = a.edge_var[e];
    const float al = a.alpha[p];
#pragma unroll
    for (int k = 0; k < D; ++k) a.U[k * a.stride + p] += al * (a.X[k * a.stride + p] - a.Z[v * D + k]);
}

// N = Z_v - U (core/paradmm.c:718-734)
template <int D>
__global__ __launch_bounds__(256) void k_npass(EArgs a)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.E) return;
    const int p = a.pos[e], v = a.edge_var[e];
#pragma unroll
    for (int k = 0; k < D; ++k) a.Nbuf[k * a.stride + p] = a.Z[v * D + k] - a.U[k * a.stride + p];
}

// M of the last fused iteration, X + U_old with U_old = U - alpha (X - Z)
template <int D>
__global__ __launch_bounds__(256) void k_mrecon(EArgs a)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.E) return;
    const int p = a.pos[e], v = a.edge_var[e];
    const float al = a.alpha[p];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float x = a.X[k * a.stride + p], u = a.U[k * a.stride + p];
        a.M[k * a.stride + p] = x + (u - al * (x - a.Z[v * D + k]));
    }
}

// host AoS fp64 [E][D] <-> device SoA fp32 [D][stride] through the position map
__global__ __launch_bounds__(256) void k_edges_in(int E, int D, const double* __restrict__ aos,
                                                  float* __restrict__ soa, long stride, const int* __restrict__ pos)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    const int p = pos[e];
    for (int k = 0; k < D; ++k) soa[k * stride + p] = (float)aos[(long)e * D + k];
}

__global__ __launch_bounds__(256) void k_edges_out(int E, int D, const float* __restrict__ soa, long stride,
                                                   double* __restrict__ aos, const int* __restrict__ pos)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    const int p = pos[e];
    for (int k = 0; k < D; ++k) aos[(long)e * D + k] = (double)soa[k * stride + p];
}

__global__ __launch_bounds__(256) void k_flat_in(long n, const double* __restrict__ src, float* __restrict__ dst)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ __launch_bounds__(256) void k_flat_out(long n, const float* __restrict__ src, double* __restrict__ dst)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

}  // namespace

// Deterministic final reduction of [n][2] FP64 partials: one workgroup,
// fixed strided assignment then a fixed LDS tree.
__global__ __launch_bounds__(256) void pdm_k_res_final(int n, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double sp[256], sq[256];
    double p = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        p += part[2 * i];
        q += part[2 * i + 1];
    }
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

// ------------------------------------------------------------------ state management

static std::mutex g_state_mu;
static std::unordered_map<graph*, DevGraph*> g_state;

static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

namespace pdm {

void free_dev(DevGraph* s)
{
    if (!s) return;
    // nothing may still run on the stream when its graphs and buffers go
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& r : s->runs) {
        if (r.exec) (void)hipGraphExecDestroy(r.exec);
        if (r.graph) (void)hipGraphDestroy(r.graph);
    }
    s->runs.clear();
    if (s->host) {
        if (s->host->h_buf) (void)hipHostFree(s->host->h_buf);
        delete s->host;
        s->host = nullptr;
    }
    for (void* p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    for (hipEvent_t e : s->prof_pool) (void)hipEventDestroy(e);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    dense_free(s);
    delete s;
}

int dmalloc(DevGraph* s, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        pdm_set_error(PARADMM_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return PARADMM_ERR_NOMEM;
    }
    s->allocs.push_back(*p);
    return PARADMM_OK;
}

}  // namespace pdm

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

// Builds the generic CSR layout and the per-kind factor tables.
// The row structure of the dim-2 all-pairs packing (DevGraph::rows2d_n):
// variables 2c (position) and 2c + 1 (radius) of circles c < nc, one
// PAIR_2D factor (2i, 2i + 1, 2j, 2j + 1) for every i < j, and each of
// these edges at its row position.  The reference demo builds exactly this
// (packing/host.c:256-343: radius factors, then the walls, then the pairs
// i < j in order), so its variables list their pair edges in the order of
// the other circle.  Anything else keeps the per-factor pass.
// PARADMM_PAIR_ROWS=0 turns it off.
static int detect_pair2d_rows(DevGraph* s, graph* G, const std::vector<int>& fkind, const std::vector<int>& pos)
{
    if (const char* e = getenv("PARADMM_PAIR_ROWS"))
        if (e[0] == '0') return PARADMM_OK;
    if (G->num_dims != 2 || G->num_vars % 2) return PARADMM_OK;
    const int nc = G->num_vars / 2;
    if (nc < 2) return PARADMM_OK;
    long npairs = 0;
    for (int f = 0; f < G->num_opts; ++f) npairs += fkind[f] == PARADMM_KIND_PAIR_2D;
    if (npairs != (long)nc * (nc - 1) / 2) return PARADMM_OK;
    // a pair listed twice would need two edges at one position: the position
    // map is one-to-one, so the row checks below also rule that out
    std::vector<long> rowP(nc, -1), rowR(nc, -1);
    auto set_row = [](std::vector<long>& row, int c, long v) {
        if (row[c] < 0) row[c] = v;
        return row[c] == v;
    };
    for (int f = 0; f < G->num_opts; ++f) {
        if (fkind[f] != PARADMM_KIND_PAIR_2D) continue;
        if (G->operatorix_to_numedges[f] != 4) return PARADMM_OK;
        const int* el = G->operatorix_to_edgelistix[f];
        const int v0 = G->edgeix_to_varlistix[el[0]], v1 = G->edgeix_to_varlistix[el[1]];
        const int v2 = G->edgeix_to_varlistix[el[2]], v3 = G->edgeix_to_varlistix[el[3]];
        if (v0 % 2 || v1 != v0 + 1 || v2 % 2 || v3 != v2 + 1) return PARADMM_OK;
        const int i = v0 / 2, j = v2 / 2;
        if (i >= j) return PARADMM_OK;
        if (!set_row(rowP, i, (long)pos[el[0]] - (j - 1)) || !set_row(rowR, i, (long)pos[el[1]] - (j - 1)) ||
            !set_row(rowP, j, (long)pos[el[2]] - i) || !set_row(rowR, j, (long)pos[el[3]] - i))
            return PARADMM_OK;
    }
    std::vector<int> rp(nc), rr(nc);
    for (int c = 0; c < nc; ++c) {
        if (rowP[c] < 0 || rowR[c] < 0) return PARADMM_OK;
        rp[c] = (int)rowP[c];
        rr[c] = (int)rowR[c];
    }
    RC(dmalloc(s, (void**)&s->d_rowP, sizeof(int) * (size_t)nc));
    RC(dmalloc(s, (void**)&s->d_rowR, sizeof(int) * (size_t)nc));
    CK(hipMemcpy(s->d_rowP, rp.data(), sizeof(int) * (size_t)nc, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_rowR, rr.data(), sizeof(int) * (size_t)nc, hipMemcpyHostToDevice));
    s->rows2d_n = nc;
    return PARADMM_OK;
}

static int finalize_generic(DevGraph* s, graph* G, const std::vector<int>& fkind)
{
    const int E = G->num_edges, V = G->num_vars, F = G->num_opts, D = G->num_dims;
    std::vector<int> pos(E, -1), edge_var_dev(E), var_off(V + 1);
    int p = 0;
    for (int v = 0; v < V; ++v) {
        var_off[v] = p;
        for (int i = 0; i < G->varix_to_numedges[v]; ++i) {
            const int e = G->varix_to_edgelistix[v][i];
            if (e < 0 || e >= E || pos[e] >= 0) {
                pdm_set_error(PARADMM_ERR_STATE, "graph inconsistent: variable %d lists edge %d", v, e);
                return PARADMM_ERR_STATE;
            }
            pos[e] = p;
            edge_var_dev[p] = v;
            ++p;
        }
    }
    var_off[V] = p;
    int maxdeg = 0;
    for (int v = 0; v < V; ++v) maxdeg = std::max(maxdeg, var_off[v + 1] - var_off[v]);
    // edges per thread of k_zpass_fused_wg (0: the wavefront-per-variable
    // pass): the smallest of 8, 16, 24 .. ZWG<D>::C that covers maxdeg
    const int cmax = D == 1 ? ZWG<1>::C : D == 2 ? ZWG<2>::C : D == 3 ? ZWG<3>::C : 0;
    s->zpass_wg = 0;
    for (int c = 8; c <= cmax && !s->zpass_wg; c += 8)
        if (maxdeg <= 256 * c) s->zpass_wg = c;
    if (const char* ze = getenv("PARADMM_ZPASS_WG"))
        if (ze[0] == '0') s->zpass_wg = 0;
    if (p != E) {
        pdm_set_error(PARADMM_ERR_STATE, "graph inconsistent: %d of %d edges listed by variables", p, E);
        return PARADMM_ERR_STATE;
    }
    s->stride = E;
    s->npos = E;
    RC(dmalloc(s, (void**)&s->d_pos, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&s->d_edge_var_api, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&s->d_edge_var_dev, sizeof(int) * (size_t)E));
    RC(dmalloc(s, (void**)&s->d_var_off, sizeof(int) * (size_t)(V + 1)));
    CK(hipMemcpy(s->d_pos, pos.data(), sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_edge_var_api, G->edgeix_to_varlistix, sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_edge_var_dev, edge_var_dev.data(), sizeof(int) * (size_t)E, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->d_var_off, var_off.data(), sizeof(int) * (size_t)(V + 1), hipMemcpyHostToDevice));

    // per-kind item tables
    for (int K = 1; K < PARADMM_KIND_COUNT; ++K) {
        const KindInfo& ki = kind_info(K);
        std::vector<int> items0;  // [item][deg] row-major, transposed on upload
        std::vector<float> prm;
        int deg = ki.deg ? ki.deg : 1;
        int n = 0;
        for (int f = 0; f < F; ++f) {
            if (fkind[f] != K) continue;
            const int fd = G->operatorix_to_numedges[f];
            const double* fp = (const double*)G->list_of_operators_param_pointers[f];
            if (ki.nparams > 0 && !fp) {
                pdm_set_error(PARADMM_ERR_ARG, "factor %d (%s) needs %d parameters, got NULL", f, ki.name, ki.nparams);
                return PARADMM_ERR_ARG;
            }
            if (ki.deg) {
                for (int sidx = 0; sidx < deg; ++sidx) items0.push_back(pos[G->operatorix_to_edgelistix[f][sidx]]);
                for (int q = 0; q < ki.nparams; ++q) prm.push_back((float)fp[q]);
                ++n;
            } else {
                for (int sidx = 0; sidx < fd; ++sidx) {
                    items0.push_back(pos[G->operatorix_to_edgelistix[f][sidx]]);
                    for (int q = 0; q < ki.nparams; ++q) prm.push_back((float)fp[q]);
                    ++n;
                }
            }
        }
        if (n == 0) continue;
        std::vector<int> items((size_t)n * deg);
        for (int i = 0; i < n; ++i)
            for (int sidx = 0; sidx < deg; ++sidx) items[(size_t)sidx * n + i] = items0[(size_t)i * deg + sidx];
        KindTable kt;
        kt.kind = K;
        kt.nitems = n;
        RC(dmalloc(s, (void**)&kt.d_items, sizeof(int) * items.size()));
        CK(hipMemcpy(kt.d_items, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice));
        if (!prm.empty()) {
            RC(dmalloc(s, (void**)&kt.d_params, sizeof(float) * prm.size()));
            CK(hipMemcpy(kt.d_params, prm.data(), sizeof(float) * prm.size(), hipMemcpyHostToDevice));
        }
        s->kinds.push_back(kt);
    }
    RC(detect_pair2d_rows(s, G, fkind, pos));
    // factors whose callback runs on the host
    HostProx hp;
    std::vector<int> hpos;
    for (int f = 0; f < F; ++f) {
        if (fkind[f] >= 0) continue;
        const int fd = G->operatorix_to_numedges[f];
        hp.fn.push_back((void*)G->list_of_operators_functions_pointers[f]);
        hp.params.push_back(G->list_of_operators_param_pointers[f]);
        hp.first.push_back((int)hpos.size());
        hp.deg.push_back(fd);
        if (fd > hp.maxdeg) hp.maxdeg = fd;
        for (int q = 0; q < fd; ++q) hpos.push_back(pos[G->operatorix_to_edgelistix[f][q]]);
    }
    if (!hpos.empty()) {
        HostProx* h = new HostProx(std::move(hp));
        s->host = h;
        h->nedges = (int)hpos.size();
        const size_t nb = sizeof(float) * (size_t)h->nedges * (2 * D + 1);
        RC(dmalloc(s, (void**)&h->d_pos, sizeof(int) * hpos.size()));
        RC(dmalloc(s, (void**)&h->d_buf, nb));
        CK(hipHostMalloc((void**)&h->h_buf, nb, hipHostMallocDefault));
        CK(hipMemcpy(h->d_pos, hpos.data(), sizeof(int) * hpos.size(), hipMemcpyHostToDevice));
    }
    return PARADMM_OK;
}

namespace pdm {
// device operator kind of every factor, with degree / dimension checks
int pdm_factor_kinds(graph* G, std::vector<int>& fkind, bool allow_host)
{
    const int F = G->num_opts, D = G->num_dims;
    fkind.assign(F, 0);
    for (int f = 0; f < F; ++f) {
        paradmm_prox_fn fn = (paradmm_prox_fn)G->list_of_operators_functions_pointers[f];
        const int K = paradmm_prox_kind_of(fn);
        if (K == PARADMM_KIND_NONE && allow_host && fn) {
            fkind[f] = -1;
            continue;
        }
        if (K == PARADMM_KIND_NONE) {
            Dl_info info;
            const char* nm = (dladdr((void*)fn, &info) && info.dli_sname) ? info.dli_sname : "?";
            pdm_set_error(PARADMM_ERR_KIND,
                          "factor %d: callback %p (%s) is not a device prox kind and host callbacks are off "
                          "(PARADMM_HOST_PROX=0); use the paradmm_prox_* operators, register it with "
                          "paradmm_register_prox_kind, or link the caller with -rdynamic",
                          f, (void*)fn, nm);
            return PARADMM_ERR_KIND;
        }
        const KindInfo& ki = kind_info(K);
        const int fd = G->operatorix_to_numedges[f];
        if (ki.dims != D || (ki.deg && fd != ki.deg) || fd < 1) {
            pdm_set_error(PARADMM_ERR_SHAPE, "factor %d (%s): degree %d / num_dims %d do not fit (need degree %d, dims %d)",
                          f, ki.name, fd, D, ki.deg, ki.dims);
            return PARADMM_ERR_SHAPE;
        }
        fkind[f] = K;
    }

    return PARADMM_OK;
}
}  // namespace pdm

static int finalize(graph* G, DevGraph** out)
{
    const int E = G->num_edges, V = G->num_vars, F = G->num_opts, D = G->num_dims;
    if (D < 1 || D > 4) {
        pdm_set_error(PARADMM_ERR_SHAPE, "num_dims %d not supported (1..4)", D);
        return PARADMM_ERR_SHAPE;
    }
    if (E <= 0 || V <= 0 || F <= 0) {
        pdm_set_error(PARADMM_ERR_STATE, "graph is empty");
        return PARADMM_ERR_STATE;
    }
    std::vector<int> fkind;
    const char* hp = getenv("PARADMM_HOST_PROX");
    RC(pdm_factor_kinds(G, fkind, !(hp && hp[0] == '0')));

    DevGraph* s = new DevGraph();
    s->G = G;
    s->E = E; s->V = V; s->F = F; s->D = D;
    int rc = PARADMM_OK;
#define FAIL_IF(x) do { rc = (x); if (rc) { free_dev(s); return rc; } } while (0)
    {
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { FAIL_IF(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__)); }
        e = hipEventCreate(&s->ev0);
        if (e == hipSuccess) e = hipEventCreate(&s->ev1);
        if (e != hipSuccess) { FAIL_IF(hip_fail(e, "hipEventCreate", __FILE__, __LINE__)); }
    }
    const char* force = getenv("PARADMM_LAYOUT");
    bool want_dense = !(force && !strcmp(force, "generic"));
    s->layout = 0;
    if (want_dense) {
        rc = dense_try_finalize(s, G, fkind);
        if (rc < 0) rc = PARADMM_OK;          // structure does not match: generic
        else if (rc) { free_dev(s); return rc; }
        else s->layout = 2;
    }
    if (s->layout == 0) {
        FAIL_IF(finalize_generic(s, G, fkind));
        s->layout = 1;
    }
    // factor type of every api edge (the learned rho policy's feature)
    {
        std::vector<unsigned char> et(E, 0);
        for (int f = 0; f < F; ++f) {
            const int K = fkind[f];
            const unsigned char t = (K == PARADMM_KIND_PAIR_2D || K == PARADMM_KIND_PAIR_3D) ? 0
                                    : (K == PARADMM_KIND_RADIUS_2D || K == PARADMM_KIND_RADIUS_3D) ? 2 : 1;
            for (int q = 0; q < G->operatorix_to_numedges[f]; ++q) et[G->operatorix_to_edgelistix[f][q]] = t;
        }
        FAIL_IF(dmalloc(s, (void**)&s->d_edge_type_api, (size_t)E));
        hipError_t e = hipMemcpy(s->d_edge_type_api, et.data(), (size_t)E, hipMemcpyHostToDevice);
        if (e != hipSuccess) FAIL_IF(hip_fail(e, "hipMemcpy", __FILE__, __LINE__));
    }
    // state buffers
    const size_t se = (size_t)D * s->stride;
    FAIL_IF(dmalloc(s, (void**)&s->d_X, sizeof(float) * se));
    FAIL_IF(dmalloc(s, (void**)&s->d_U, sizeof(float) * se));
    FAIL_IF(dmalloc(s, (void**)&s->d_rho, sizeof(float) * s->stride));
    FAIL_IF(dmalloc(s, (void**)&s->d_alpha, sizeof(float) * s->stride));
    FAIL_IF(dmalloc(s, (void**)&s->d_Z, sizeof(float) * (size_t)V * D));
    FAIL_IF(dmalloc(s, (void**)&s->d_Zprev, sizeof(float) * (size_t)V * D));
    FAIL_IF(dmalloc(s, (void**)&s->d_res, sizeof(double) * 4));
    if (s->layout == 1) FAIL_IF(dmalloc(s, (void**)&s->d_res_part, sizeof(double) * 2 * (size_t)V));
    {
        // padding slots of the dense layout must hold finite values
        hipError_t e = hipMemsetAsync(s->d_X, 0, sizeof(float) * se, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->d_U, 0, sizeof(float) * se, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->d_rho, 0, sizeof(float) * s->stride, s->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->d_alpha, 0, sizeof(float) * s->stride, s->stream);
        if (e != hipSuccess) FAIL_IF(hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__));
    }
    if (s->layout == 2) FAIL_IF(dense_alloc_state(s));
#undef FAIL_IF
    *out = s;
    return PARADMM_OK;
}

namespace pdm {

int ensure_aux(DevGraph* s)
{
    // N, M buffers and the host staging area: only legacy passes and host
    // synchronisation need them.
    const size_t se = (size_t)s->D * s->stride;
    if (!s->d_N) {
        RC(dmalloc(s, (void**)&s->d_N, sizeof(float) * se));
        CK(hipMemsetAsync(s->d_N, 0, sizeof(float) * se, s->stream));
    }
    if (!s->d_M) {
        RC(dmalloc(s, (void**)&s->d_M, sizeof(float) * se));
        CK(hipMemsetAsync(s->d_M, 0, sizeof(float) * se, s->stream));
    }
    if (!s->d_stage) {
        // host transfers use D*max(E,V) doubles; the warm start keeps a copy
        // of X (D*stride floats, larger when the dense layout pads) here too
        size_t n = (size_t)s->D * (size_t)(s->E > s->V ? s->E : s->V);
        const size_t nx = ((size_t)s->D * (size_t)s->stride + 1) / 2;
        if (nx > n) n = nx;
        RC(dmalloc(s, (void**)&s->d_stage, sizeof(double) * n));
    }
    return PARADMM_OK;
}

}  // namespace pdm

// ------------------------------------------------------------------ host <-> device

enum { A_X = 1, A_M = 2, A_N = 4, A_U = 8, A_Z = 16, A_RHO = 32, A_ALPHA = 64 };

static int upload(DevGraph* s, int which)
{
    graph* G = s->G;
    RC(ensure_aux(s));
    const int E = s->E, D = s->D;
    struct { int bit; double* host; float* dev; int d; } edge_arrays[] = {
        {A_X, G->X, s->d_X, D}, {A_M, G->M, s->d_M, D}, {A_N, G->N, s->d_N, D}, {A_U, G->U, s->d_U, D},
        {A_RHO, G->RHOS, s->d_rho, 1}, {A_ALPHA, G->ALPHAS, s->d_alpha, 1}};
    for (auto& a : edge_arrays) {
        if (!(which & a.bit)) continue;
        if (!a.host) {
            pdm_set_error(PARADMM_ERR_STATE, "host array needed by the pass is not allocated (initialise the graph first)");
            return PARADMM_ERR_STATE;
        }
        CK(hipMemcpyAsync(s->d_stage, a.host, sizeof(double) * (size_t)a.d * E, hipMemcpyHostToDevice, s->stream));
        k_edges_in<<<nblk(E), 256, 0, s->stream>>>(E, a.d, s->d_stage, a.dev, s->stride, s->d_pos);
        CK(hipGetLastError());
    }
    if (which & A_Z) {
        if (!G->Z) {
            pdm_set_error(PARADMM_ERR_STATE, "Z is not allocated (initialise the graph first)");
            return PARADMM_ERR_STATE;
        }
        const long n = (long)s->V * D;
        CK(hipMemcpyAsync(s->d_stage, G->Z, sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
        k_flat_in<<<nblk(n), 256, 0, s->stream>>>(n, s->d_stage, s->d_Z);
        CK(hipGetLastError());
    }
    if (which & (A_RHO | A_ALPHA)) s->rho_changed = true;
    // one common value over every edge (compared in fp32, the device type)
    auto common = [E](const double* h, float& v) {
        v = (float)h[0];
        for (int e = 1; e < E; ++e)
            if ((float)h[e] != v) return false;
        return true;
    };
    if (which & A_RHO) s->rho_uni = common(G->RHOS, s->rho_u);
    if (which & A_ALPHA) s->alpha_uni = common(G->ALPHAS, s->alpha_u);
    CK(hipStreamSynchronize(s->stream));  // the staging area is reused
    if (which & (A_RHO | A_ALPHA)) RC(sync_uniform(s));
    return PARADMM_OK;
}

namespace pdm {
int sync_uniform(DevGraph* s)
{
    float* d = dense_uni_ptr(s);
    if (!d) return PARADMM_OK;
    const float v[2] = {s->rho_u, s->alpha_u};
    CK(hipStreamSynchronize(s->stream));
    CK(hipMemcpy(d, v, sizeof v, hipMemcpyHostToDevice));
    return PARADMM_OK;
}
}  // namespace pdm

static int download(DevGraph* s, int which)
{
    graph* G = s->G;
    RC(ensure_aux(s));
    const int E = s->E, D = s->D;
    struct { int bit; double* host; float* dev; int d; } edge_arrays[] = {
        {A_X, G->X, s->d_X, D}, {A_M, G->M, s->d_M, D}, {A_N, G->N, s->d_N, D}, {A_U, G->U, s->d_U, D},
        {A_RHO, G->RHOS, s->d_rho, 1}, {A_ALPHA, G->ALPHAS, s->d_alpha, 1}};
    for (auto& a : edge_arrays) {
        if (!(which & a.bit)) continue;
        if (!a.host) {
            pdm_set_error(PARADMM_ERR_STATE, "host array to receive a pass output is not allocated");
            return PARADMM_ERR_STATE;
        }
        k_edges_out<<<nblk(E), 256, 0, s->stream>>>(E, a.d, a.dev, s->stride, s->d_stage, s->d_pos);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(a.host, s->d_stage, sizeof(double) * (size_t)a.d * E, hipMemcpyDeviceToHost, s->stream));
        CK(hipStreamSynchronize(s->stream));
    }
    if (which & A_Z) {
        if (!G->Z) {
            pdm_set_error(PARADMM_ERR_STATE, "Z is not allocated");
            return PARADMM_ERR_STATE;
        }
        const long n = (long)s->V * D;
        k_flat_out<<<nblk(n), 256, 0, s->stream>>>(n, s->d_Z, s->d_stage);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(G->Z, s->d_stage, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        CK(hipStreamSynchronize(s->stream));
    }
    return PARADMM_OK;
}

// ------------------------------------------------------------------ passes

// The pair X pass of the dim-2 all-pairs packing whose edges lie in rows
// (DevGraph::rows2d_n): one workgroup per tile of T x T circle pairs
// (i in block I, j in block J, i < j).  The first circle's edges of the
// tile's pairs are row segments (rowP[i] + j - 1, rowR[i] + j - 1:
// consecutive j, read and written with lanes along j); the second's are
// the transposed block (rowP[j] + i, rowR[j] + i: consecutive i), read with
// lanes along i into LDS and written back the same way.  The prox and its
// inputs are those of k_xpass<PAIR_2D>, so the results are the same bits.
struct XRowArgs {
    int nc, nb;
    const int *rowP, *rowR;
    const float* Z;     // [V][2]: circle c's pos at 4c, rad at 4c + 2
    const float* U;
    const float* Nbuf;
    const float* rho;
    float* X;
    long stride;
    int from_n;
    // the graph's wall and radius factors, taken by wblk and rblk
    // workgroups after the pair tiles' (their edges are not the pairs')
    XArgs wall, rad;
    int wblk, rblk;
};

template <int T, bool FROM_N>
__global__ __launch_bounds__(256) void k_xpass_pair2d_rows(XRowArgs a)
{
    // second circle's N (pos 2, rad 2) and rho (pos, rad), [.][jl][il]; each
    // thread overwrites its own entry's N with the second circle's X
    __shared__ float bin[6][T][T + 1];
    constexpr int R = T * T / 256;  // pairs per thread in each phase
    static_assert(T * T % 256 == 0 && 256 % T == 0, "whole rows of the tile per pass of the workgroup");
    const int tid = threadIdx.x;
    // tile t of the upper triangle of nb x nb blocks.  Workgroups are dealt
    // to the 8 XCDs in turn (blockIdx % 8), each with its own L2; each XCD
    // takes a contiguous run of tiles instead, so the tiles (I, J), (I, J + 1)
    // that share the cache lines where their row segments meet share an L2
    const int nt = a.nb * (a.nb + 1) / 2, chunk = (nt + 7) / 8;
    if ((int)blockIdx.x >= 8 * chunk) {  // the wall and radius factors, one per thread
        const int b = (int)blockIdx.x - 8 * chunk;
        if (b < a.wblk) {
            if (b * 256 + tid < a.wall.nitems) xpass_item<PARADMM_KIND_WALL_2D>(a.wall, b * 256 + tid);
        } else if (b - a.wblk < a.rblk) {
            const int i = (b - a.wblk) * 256 + tid;
            if (i < a.rad.nitems) xpass_item<PARADMM_KIND_RADIUS_2D>(a.rad, i);
        }
        return;
    }
    if (a.nc < 2) return;  // no pairs (the stand-in pair below is (0, 1))
    const int t = (int)(blockIdx.x % 8) * chunk + (int)(blockIdx.x / 8);
    if (t >= nt) return;
    int I = (int)((2.0f * a.nb + 1.0f - sqrtf((2.0f * a.nb + 1.0f) * (2.0f * a.nb + 1.0f) - 8.0f * (float)t)) * 0.5f);
    I = max(0, min(I, a.nb - 1));
    while (I > 0 && I * a.nb - I * (I - 1) / 2 > t) --I;
    while (I + 1 < a.nb && (I + 1) * a.nb - (I + 1) * I / 2 <= t) ++I;
    const int J = I + (t - (I * a.nb - I * (I - 1) / 2));
    const long S = a.stride;
    // A thread's R pairs of a phase are loaded before any is used, and a
    // pair outside the tile's triangle or past nc reads the stand-in pair
    // (0, 1) instead (its values are not used): no branch around a load, so
    // the R pairs' loads are in flight together.
    // The second circles' inputs, lanes along i.
    float in[R][6];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int idx = tid + 256 * m;
        const int jl = idx / T, il = idx % T;
        const int i = I * T + il, j = J * T + jl;
        const bool ok = i < a.nc && j < a.nc && i < j;
        const int ie = ok ? i : 0, je = ok ? j : 1;
        const long p = a.rowP[je] + ie, q = a.rowR[je] + ie;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            in[m][k] = FROM_N ? a.Nbuf[k * S + p] : a.Z[4L * je + k] - a.U[k * S + p];
            in[m][2 + k] = FROM_N ? a.Nbuf[k * S + q] : a.Z[4L * je + 2 + k] - a.U[k * S + q];
        }
        in[m][4] = a.rho[p];
        in[m][5] = a.rho[q];
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int idx = tid + 256 * m;
        const int jl = idx / T, il = idx % T;
#pragma unroll
        for (int f = 0; f < 6; ++f) bin[f][jl][il] = in[m][f];
    }
    __syncthreads();
    // the first circles' edges, lanes along j; the prox
    float fn[R][4], fr[R][2];
    long fp[R], fq[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int idx = tid + 256 * m;
        const int il = idx / T, jl = idx % T;
        const int i = I * T + il, j = J * T + jl;
        const bool ok = i < a.nc && j < a.nc && i < j;
        const int ie = ok ? i : 0, je = ok ? j : 1;
        fp[m] = a.rowP[ie] + je - 1;
        fq[m] = a.rowR[ie] + je - 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            fn[m][k] = FROM_N ? a.Nbuf[k * S + fp[m]] : a.Z[4L * ie + k] - a.U[k * S + fp[m]];
            fn[m][2 + k] = FROM_N ? a.Nbuf[k * S + fq[m]] : a.Z[4L * ie + 2 + k] - a.U[k * S + fq[m]];
        }
        fr[m][0] = a.rho[fp[m]];
        fr[m][1] = a.rho[fq[m]];
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int idx = tid + 256 * m;
        const int il = idx / T, jl = idx % T;
        const int i = I * T + il, j = J * T + jl;
        if (i >= a.nc || j >= a.nc || i >= j) continue;
        float n[4][2], x[4][2], rho[4];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            n[0][k] = fn[m][k];
            n[1][k] = fn[m][2 + k];
            n[2][k] = bin[k][jl][il];
            n[3][k] = bin[2 + k][jl][il];
        }
        rho[0] = fr[m][0];
        rho[1] = fr[m][1];
        rho[2] = bin[4][jl][il];
        rho[3] = bin[5][jl][il];
        prox_pair2(x, n, rho);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            a.X[k * S + fp[m]] = x[0][k];
            a.X[k * S + fq[m]] = x[1][k];
            bin[k][jl][il] = x[2][k];
            bin[2 + k][jl][il] = x[3][k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const int idx = tid + 256 * m;
        const int jl = idx / T, il = idx % T;
        const int i = I * T + il, j = J * T + jl;
        if (i >= a.nc || j >= a.nc || i >= j) continue;
        const long p = a.rowP[j] + i, q = a.rowR[j] + i;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            a.X[k * S + p] = bin[k][jl][il];
            a.X[k * S + q] = bin[2 + k][jl][il];
        }
    }
}

constexpr int XROW_T = 32;

static XArgs xargs(DevGraph* s, const KindTable& kt, int from_n)
{
    XArgs a;
    a.nitems = kt.nitems;
    a.items = kt.d_items;
    a.params = kt.d_params;
    a.edge_var = s->d_edge_var_dev;
    a.Z = s->d_Z;
    a.U = s->d_U;
    a.Nbuf = s->d_N;
    a.rho = s->d_rho;
    a.X = s->d_X;
    a.stride = s->stride;
    a.from_n = from_n;
    return a;
}

// The pair tiles, then (wall, rad: their kind tables, or null) the wall and
// radius factors in the same launch: their latency hides in the pair
// pass's tail instead of taking two launches of their own.
static int launch_x_pair2d_rows(DevGraph* s, int from_n, const KindTable* wall, const KindTable* rad)
{
    XRowArgs a;
    a.nc = s->rows2d_n;
    a.nb = (a.nc + XROW_T - 1) / XROW_T;
    a.rowP = s->d_rowP;
    a.rowR = s->d_rowR;
    a.Z = s->d_Z;
    a.U = s->d_U;
    a.Nbuf = s->d_N;
    a.rho = s->d_rho;
    a.X = s->d_X;
    a.stride = s->stride;
    a.from_n = from_n;
    a.wall = wall ? xargs(s, *wall, from_n) : XArgs{};
    a.rad = rad ? xargs(s, *rad, from_n) : XArgs{};
    a.wblk = wall ? (int)nblk(wall->nitems) : 0;
    a.rblk = rad ? (int)nblk(rad->nitems) : 0;
    const unsigned nt = (unsigned)(a.nb * (a.nb + 1) / 2);
    const unsigned grid = 8 * ((nt + 7) / 8) + (unsigned)(a.wblk + a.rblk);
    prof_begin(s, "disk_collision_2d_rows");
    if (from_n) k_xpass_pair2d_rows<XROW_T, true><<<grid, 256, 0, s->stream>>>(a);
    else k_xpass_pair2d_rows<XROW_T, false><<<grid, 256, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

template <int K>
static int launch_x_kind(DevGraph* s, const KindTable& kt, int from_n)
{
    const XArgs a = xargs(s, kt, from_n);
    prof_begin(s, kind_info(K).name);
    k_xpass<K><<<nblk(kt.nitems), 256, 0, s->stream>>>(a);
    prof_end(s);
    CK(hipGetLastError());
    return PARADMM_OK;
}

// Host-callback factors: gather (N, X, rho) of their edges, run the
// callbacks in FP64 on the host in factor order, scatter (X, rho) back.
__global__ __launch_bounds__(256) void k_host_gather(int n, const int* __restrict__ pos, const int* __restrict__ edge_var,
                                                     const float* __restrict__ Z, const float* __restrict__ U,
                                                     const float* __restrict__ Nbuf, const float* __restrict__ X,
                                                     const float* __restrict__ rho, int from_n, int D, long stride,
                                                     float* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = pos[i];
    const int v = edge_var[p];
    float* o = out + (long)i * (2 * D + 1);
    for (int k = 0; k < D; ++k) {
        o[k] = from_n ? Nbuf[k * stride + p] : Z[v * D + k] - U[k * stride + p];
        o[D + k] = X[k * stride + p];
    }
    o[2 * D] = rho[p];
}

__global__ __launch_bounds__(256) void k_host_scatter(int n, const int* __restrict__ pos, int D, long stride,
                                                      const float* __restrict__ in, float* __restrict__ X,
                                                      float* __restrict__ rho)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = pos[i];
    const float* o = in + (long)i * (2 * D + 1);
    for (int k = 0; k < D; ++k) X[k * stride + p] = o[D + k];
    if (rho) rho[p] = o[2 * D];
}

static int run_host_prox(DevGraph* s, int from_n)
{
    HostProx* h = s->host;
    const int D = s->D, n = h->nedges, W = 2 * D + 1;
    prof_begin(s, "host_prox_gather");
    k_host_gather<<<nblk(n), 256, 0, s->stream>>>(n, h->d_pos, s->d_edge_var_dev, s->d_Z, s->d_U, s->d_N, s->d_X,
                                                  s->d_rho, from_n, D, s->stride, h->d_buf);
    prof_end(s);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h->h_buf, h->d_buf, sizeof(float) * (size_t)n * W, hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    std::vector<double> x((size_t)h->maxdeg * D), nn((size_t)h->maxdeg * D), r(h->maxdeg);
    bool rho_changed = false;
    for (size_t f = 0; f < h->fn.size(); ++f) {
        const int deg = h->deg[f];
        float* b = h->h_buf + (size_t)h->first[f] * W;
        for (int q = 0; q < deg; ++q) {
            for (int k = 0; k < D; ++k) {
                nn[(size_t)q * D + k] = b[(size_t)q * W + k];
                x[(size_t)q * D + k] = b[(size_t)q * W + D + k];
            }
            r[q] = b[(size_t)q * W + 2 * D];
        }
        ((paradmm_prox_fn)h->fn[f])(x.data(), nn.data(), r.data(), deg, D, h->params[f]);
        for (int q = 0; q < deg; ++q) {
            for (int k = 0; k < D; ++k) b[(size_t)q * W + D + k] = (float)x[(size_t)q * D + k];
            const float rn = (float)r[q];
            rho_changed = rho_changed || rn != b[(size_t)q * W + 2 * D];
            b[(size_t)q * W + 2 * D] = rn;
        }
    }
    CK(hipMemcpyAsync(h->d_buf, h->h_buf, sizeof(float) * (size_t)n * W, hipMemcpyHostToDevice, s->stream));
    prof_begin(s, "host_prox_scatter");
    k_host_scatter<<<nblk(n), 256, 0, s->stream>>>(n, h->d_pos, D, s->stride, h->d_buf, s->d_X,
                                                   rho_changed ? s->d_rho : nullptr);
    prof_end(s);
    CK(hipGetLastError());
    if (rho_changed) {
        s->rho_device_written = true;
        s->rho_changed = true;
        s->rho_uni = false;
    }
    return PARADMM_OK;
}

static int launch_xpass_generic(DevGraph* s, int from_n)
{
    // with the pairs in rows, the wall and radius factors go with the pair tiles
    const KindTable *wall = nullptr, *rad = nullptr;
    if (s->rows2d_n)
        for (const KindTable& kt : s->kinds) {
            if (kt.kind == PARADMM_KIND_WALL_2D) wall = &kt;
            if (kt.kind == PARADMM_KIND_RADIUS_2D) rad = &kt;
        }
    for (const KindTable& kt : s->kinds) {
        switch (kt.kind) {
        case PARADMM_KIND_PAIR_2D:
            if (s->rows2d_n) RC(launch_x_pair2d_rows(s, from_n, wall, rad));
            else RC(launch_x_kind<PARADMM_KIND_PAIR_2D>(s, kt, from_n));
            break;
        case PARADMM_KIND_WALL_2D:
            if (!wall) RC(launch_x_kind<PARADMM_KIND_WALL_2D>(s, kt, from_n));
            break;
        case PARADMM_KIND_RADIUS_2D:
            if (!rad) RC(launch_x_kind<PARADMM_KIND_RADIUS_2D>(s, kt, from_n));
            break;
        case PARADMM_KIND_PAIR_3D: RC(launch_x_kind<PARADMM_KIND_PAIR_3D>(s, kt, from_n)); break;
        case PARADMM_KIND_BOX_3D: RC(launch_x_kind<PARADMM_KIND_BOX_3D>(s, kt, from_n)); break;
        case PARADMM_KIND_RADIUS_3D: RC(launch_x_kind<PARADMM_KIND_RADIUS_3D>(s, kt, from_n)); break;
        default: break;
        }
    }
    if (s->host) RC(run_host_prox(s, from_n));
    return PARADMM_OK;
}

static ZArgs zargs(DevGraph* s)
{
    ZArgs a;
    a.V = s->V;
    a.var_off = s->d_var_off;
    a.X = s->d_X;
    a.U = s->d_U;
    a.M = s->d_M;
    a.rho = s->d_rho;
    a.alpha = s->d_alpha;
    a.Z = s->d_Z;
    a.Zprev = s->d_Zprev;
    a.res_var = s->d_res_part;
    a.stride = s->stride;
    return a;
}

static EArgs eargs(DevGraph* s)
{
    EArgs a;
    a.E = s->E;
    a.pos = s->d_pos;
    a.edge_var = s->d_edge_var_api;
    a.X = s->d_X;
    a.U = s->d_U;
    a.Nbuf = s->d_N;
    a.M = s->d_M;
    a.Z = s->d_Z;
    a.alpha = s->d_alpha;
    a.stride = s->stride;
    return a;
}

#define DISPATCH_D(D, KERNEL, GRID, ARGS)                                                      \
    do {                                                                                       \
        switch (D) {                                                                           \
        case 1: KERNEL<1><<<GRID, 256, 0, s->stream>>>(ARGS); break;                           \
        case 2: KERNEL<2><<<GRID, 256, 0, s->stream>>>(ARGS); break;                           \
        case 3: KERNEL<3><<<GRID, 256, 0, s->stream>>>(ARGS); break;                           \
        default: KERNEL<4><<<GRID, 256, 0, s->stream>>>(ARGS); break;                          \
        }                                                                                      \
        CK(hipGetLastError());                                                                 \
    } while (0)

namespace pdm {

int launch_mrecon(DevGraph* s)
{
    EArgs a = eargs(s);
    DISPATCH_D(s->D, k_mrecon, nblk(s->E), a);
    return PARADMM_OK;
}

int launch_npass(DevGraph* s)
{
    EArgs a = eargs(s);
    DISPATCH_D(s->D, k_npass, nblk(s->E), a);
    return PARADMM_OK;
}

}  // namespace pdm

// One legacy pass on the device (no host traffic).
static int device_pass(DevGraph* s, int pass)
{
    switch (pass) {
    case PDM_PASS_X:
        if (s->layout == 2) return dense_legacy_xpass(s);
        return launch_xpass_generic(s, 1);
    case PDM_PASS_M: {
        const long n = (long)s->D * s->stride;
        k_mpass<<<nblk(n), 256, 0, s->stream>>>(n, s->d_X, s->d_U, s->d_M);
        CK(hipGetLastError());
        return PARADMM_OK;
    }
    case PDM_PASS_Z: {
        if (s->layout == 2) return dense_legacy_zpass(s);
        ZArgs a = zargs(s);
        DISPATCH_D(s->D, k_zpass_legacy, nblk(s->V, 4), a);
        return PARADMM_OK;
    }
    case PDM_PASS_U: {
        EArgs a = eargs(s);
        DISPATCH_D(s->D, k_upass, nblk(s->E), a);
        return PARADMM_OK;
    }
    case PDM_PASS_N: return launch_npass(s);
    }
    return PARADMM_OK;
}

// One fused iteration: X pass (N = Z - U inline, or from the N buffer), then
// the Z pass with M and the U update fused and residual partials.
// Dense layout with alpha = 1 on every edge: the middle iterations of a run
// keep M = X + U in the X array and do not store U (DENSE_M_*).
static bool mrun_ok(const DevGraph* s) { return s->layout == 2 && s->alpha_uni && s->alpha_u == 1.0f; }

// place of iteration i of a run of n in the M state sequence
static int mphase_of(const DevGraph* s, int i, int n)
{
    // contact state: the last iteration of a run also sums the residuals
    if (s->layout == 2 && dense_in_contact(const_cast<DevGraph*>(s))) return i == n - 1 ? DENSE_PLAIN : DENSE_CONTACT;
    if (!mrun_ok(s) || n < 2) return DENSE_PLAIN;
    return i == 0 ? DENSE_M_ENTER : i == n - 1 ? DENSE_M_EXIT : DENSE_M_MID;
}

static int fused_iteration(DevGraph* s, int from_n, int mphase = DENSE_PLAIN)
{
    if (s->layout == 2) return dense_iteration(s, from_n, mphase);
    RC(launch_xpass_generic(s, from_n));
    ZArgs a = zargs(s);
    prof_begin(s, "zpass_fused");
    if (s->zpass_wg) {
        const unsigned g = (unsigned)s->V;
        const int c = s->zpass_wg;
        if (s->D == 2) {
            if (c <= 8) k_zpass_fused_wg<2, 8><<<g, 256, 0, s->stream>>>(a);
            else if (c <= 16) k_zpass_fused_wg<2, 16><<<g, 256, 0, s->stream>>>(a);
            else k_zpass_fused_wg<2, ZWG<2>::C><<<g, 256, 0, s->stream>>>(a);
        } else if (s->D == 3) {
            if (c <= 8) k_zpass_fused_wg<3, 8><<<g, 256, 0, s->stream>>>(a);
            else k_zpass_fused_wg<3, ZWG<3>::C><<<g, 256, 0, s->stream>>>(a);
        } else {
            if (c <= 8) k_zpass_fused_wg<1, 8><<<g, 256, 0, s->stream>>>(a);
            else if (c <= 16) k_zpass_fused_wg<1, 16><<<g, 256, 0, s->stream>>>(a);
            else k_zpass_fused_wg<1, ZWG<1>::C><<<g, 256, 0, s->stream>>>(a);
        }
        CK(hipGetLastError());
    } else {
        DISPATCH_D(s->D, k_zpass_fused, nblk(s->V, 4), a);
    }
    prof_end(s);
    return PARADMM_OK;
}

// kernel variant a captured run depends on: the uniform rho/alpha kernels,
// the M state (dense, alpha = 1), the contact state with its start parity
// (RUN_PARITY: runs that differ only there are both kept) and whether its
// kernels read per-edge rho (a policy's raised edges leave the uniform-rho
// flag set, so that bit alone does not tell)
constexpr int RUN_PARITY = 8, RUN_CONTACT_RHO_EDGE = 16;

static bool in_contact(const DevGraph* s)
{
    return s->layout == 2 && dense_in_contact(const_cast<DevGraph*>(s));
}

static int run_variant(const DevGraph* s)
{
    const bool contact = in_contact(s);
    const int par = contact && dense_contact_parity(const_cast<DevGraph*>(s)) ? RUN_PARITY : 0;
    const int re = contact && dense_contact_rho_per_edge(const_cast<DevGraph*>(s)) ? RUN_CONTACT_RHO_EDGE : 0;
    return ((s->rho_uni && s->alpha_uni) ? 1 : 0) | (mrun_ok(s) ? 2 : 0) | (contact ? 4 : 0) | par | re;
}

// The captured run of n iterations for the current variant (recorded on
// first use; runs of another variant are dropped; at most 4 are kept).
// Recording needs the steady state (U one update behind), which it also
// leaves.
static int captured_run(DevGraph* s, int n, hipGraphExec_t* out)
{
    const int variant = run_variant(s);
    // A run launched earlier may still execute on the stream: wait for it
    // before destroying any executable graph (launches are asynchronous).
    bool synced = false;
    auto drain = [&]() -> int {
        if (!synced) {
            CK(hipStreamSynchronize(s->stream));
            synced = true;
        }
        return PARADMM_OK;
    };
    for (auto it = s->runs.begin(); it != s->runs.end();) {
        if ((it->variant | RUN_PARITY) != (variant | RUN_PARITY)) {
            RC(drain());
            CK(hipGraphExecDestroy(it->exec));
            CK(hipGraphDestroy(it->graph));
            it = s->runs.erase(it);
        } else if (it->iters == n && it->variant == variant) {
            *out = it->exec;
            return PARADMM_OK;
        } else {
            ++it;
        }
    }
    if (s->runs.size() >= 4) {  // the oldest run other than the main length
        auto victim = s->runs.begin();
        if (victim->iters == s->graph_iters && s->runs.size() > 1) ++victim;
        RC(drain());
        CK(hipGraphExecDestroy(victim->exec));
        CK(hipGraphDestroy(victim->graph));
        s->runs.erase(victim);
    }
    DevGraph::CapturedRun r;
    r.iters = n;
    r.variant = variant;
    // recording runs nothing: the contact parity is advanced when the run is launched
    const bool contact = in_contact(s);
    const int par0 = contact ? dense_contact_parity(s) : 0;
    CK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = PARADMM_OK;
    for (int i = 0; i < n && !rc; ++i) rc = fused_iteration(s, 0, mphase_of(s, i, n));
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &gr);
    if (contact) dense_contact_set_parity(s, par0);
    if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    r.graph = gr;
    e = hipGraphInstantiate(&r.exec, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(gr);
        return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    s->runs.push_back(r);
    *out = r.exec;
    return PARADMM_OK;
}

// Brings the state to the one captured runs start from (eager iterations
// if needed); returns the iterations that took.
static int steady_start(DevGraph* s, int iters, int* used)
{
    *used = 0;
    if (iters > 0 && s->n_from_buffer) {
        RC(fused_iteration(s, 1));
        s->n_from_buffer = false;
        ++*used;
    }
    if (s->layout == 2 && !dense_steady(s) && iters > *used) {
        // captured runs assume a pending U update
        RC(fused_iteration(s, 0));
        ++*used;
    }
    if (s->layout == 2) {
        // one rho and alpha = 1: the contact state (entering it runs an iteration)
        if (dense_in_contact(s) && !dense_contact_ok(s)) RC(dense_contact_leave(s));
        if (iters > *used && !dense_in_contact(s) && dense_contact_ok(s)) {
            RC(dense_contact_enter(s));
            ++*used;
        }
    }
    return PARADMM_OK;
}

static int run_fused(DevGraph* s, int iters)
{
    if (iters <= 0) return PARADMM_OK;
    int used = 0;
    RC(steady_start(s, iters, &used));
    iters -= used;
    // host-callback factors synchronise with the host every iteration: no capture
    if (!s->host && iters >= 2) {
        const int GI = s->graph_iters;
        hipGraphExec_t ex = nullptr;
        const bool contact = in_contact(s);
        if (iters >= GI) {
            RC(captured_run(s, GI, &ex));
            while (iters >= GI) {
                CK(hipGraphLaunch(ex, s->stream));
                if (contact) dense_contact_set_parity(s, dense_contact_parity(s) ^ (GI & 1));
                iters -= GI;
            }
        }
        if (iters >= 2) {
            RC(captured_run(s, iters, &ex));
            CK(hipGraphLaunch(ex, s->stream));
            if (contact) dense_contact_set_parity(s, dense_contact_parity(s) ^ (iters & 1));
            iters = 0;
        }
    }
    for (int i = 0; i < iters; ++i) RC(fused_iteration(s, 0));
    s->iters_done = true;
    return PARADMM_OK;
}

// ------------------------------------------------------------------ state lookup

static DevGraph* lookup(graph* G)
{
    std::lock_guard<std::mutex> lk(g_state_mu);
    auto it = g_state.find(G);
    return it == g_state.end() ? nullptr : it->second;
}

extern "C" void pdm_release_graph(graph* G)
{
    DevGraph* s = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_state_mu);
        auto it = g_state.find(G);
        if (it != g_state.end()) {
            s = it->second;
            g_state.erase(it);
        }
    }
    if (s) {
        (void)hipStreamSynchronize(s->stream);
        free_dev(s);
    }
}

extern "C" void pdm_host_written(graph* G)
{
    DevGraph* s = lookup(G);
    if (s) s->host_newer = true;
}

// Device state for G, finalised and current with respect to the graph
// structure.  Uploads everything when the host was written since.  The
// dense contact state is left (per-edge X and U restored) unless the caller
// works on it (keep_contact).
static int get_state(graph* G, DevGraph** out, bool upload_if_new, bool keep_contact = false)
{
    if (!G) {
        pdm_set_error(PARADMM_ERR_ARG, "graph is NULL");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s = lookup(G);
    if (s && (s->E != G->num_edges || s->V != G->num_vars || s->F != G->num_opts || s->D != G->num_dims)) {
        pdm_release_graph(G);  // structure changed since finalisation
        s = nullptr;
    }
    if (!s) {
        RC(finalize(G, &s));
        s->host_newer = true;
        std::lock_guard<std::mutex> lk(g_state_mu);
        g_state[G] = s;
    }
    if (upload_if_new && s->host_newer) {
        RC(upload(s, A_X | A_N | A_U | A_Z | A_RHO | A_ALPHA));
        s->host_newer = false;
        s->dev_newer = false;
        s->n_from_buffer = true;
        s->m_valid = false;
        if (s->layout == 2) dense_reset(s);
    }
    if (!keep_contact && s->layout == 2) RC(dense_contact_leave(s));
    *out = s;
    return PARADMM_OK;
}

namespace pdm {
int pdm_get_state(graph* G, DevGraph** out, bool upload) { return get_state(G, out, upload); }
}  // namespace pdm

// Brings the host arrays up to date with a device-resident run.
static int sync_down(DevGraph* s)
{
    if (!s->dev_newer) return PARADMM_OK;
    RC(ensure_aux(s));
    if (s->layout == 2) RC(dense_finish(s));
    if (!s->m_valid) RC(launch_mrecon(s));
    RC(launch_npass(s));
    int which = A_X | A_M | A_N | A_U | A_Z;
    if (s->rho_device_written) which |= A_RHO;
    RC(download(s, which));
    s->dev_newer = false;
    s->m_valid = true;
    return PARADMM_OK;
}

// ------------------------------------------------------------------ legacy entry points

extern "C" int pdm_legacy_passes(graph* G, int passes)
{
    DevGraph* s;
    RC(get_state(G, &s, false));
    RC(sync_down(s));
    RC(ensure_aux(s));
    // inputs: what the passes read that an earlier pass of this call does not produce
    static const struct { int pass, reads, writes; } tab[] = {
        {PDM_PASS_X, A_N | A_RHO, A_X},
        {PDM_PASS_M, A_X | A_U, A_M},
        {PDM_PASS_Z, A_M | A_RHO | A_Z, A_Z},
        {PDM_PASS_U, A_X | A_Z | A_ALPHA | A_U, A_U},
        {PDM_PASS_N, A_Z | A_U, A_N},
    };
    int in = 0, produced = 0;
    for (auto& t : tab) {
        if (!(passes & t.pass)) continue;
        in |= t.reads & ~produced;
        produced |= t.writes;
    }
    RC(upload(s, in));
    for (auto& t : tab)
        if (passes & t.pass) RC(device_pass(s, t.pass));
    RC(download(s, produced));
    // only this call's inputs reached the device: the host stays authoritative
    s->host_newer = true;
    s->dev_newer = false;
    if (s->layout == 2 && (passes & (PDM_PASS_Z | PDM_PASS_U))) s->dense_unary_stale = true;
    return PARADMM_OK;
}

extern "C" int pdm_legacy_iterations(graph* G, int maxiter)
{
    DevGraph* s;
    RC(get_state(G, &s, false));
    RC(sync_down(s));
    RC(upload(s, A_X | A_N | A_U | A_Z | A_RHO | A_ALPHA));
    s->host_newer = false;
    s->n_from_buffer = true;
    s->m_valid = false;
    if (s->layout == 2) dense_reset(s);
    RC(run_fused(s, maxiter));
    s->dev_newer = maxiter > 0;
    RC(sync_down(s));
    s->host_newer = true;
    return PARADMM_OK;
}

extern "C" int pdm_legacy_warm_start(graph* G, int maxiter)
{
    // initialize_X_N_Z_M_U_optimal (core/paradmm.c:331-370): per sweep
    //   Xit = prox(N);  U = Xit - N;  N = X - U   with X fixed.
    DevGraph* s;
    RC(get_state(G, &s, false));
    RC(upload(s, A_X | A_N | A_U | A_Z | A_RHO | A_ALPHA));
    RC(warm_start_device(s, maxiter));
    RC(download(s, A_X | A_M | A_N | A_U));
    s->host_newer = true;
    s->dev_newer = false;
    return PARADMM_OK;
}

// ------------------------------------------------------------------ extension API

extern "C" int paradmm_gpu_prepare(graph* G)
{
    DevGraph* s;
    return get_state(G, &s, true);
}

extern "C" int paradmm_gpu_iterate(graph* G, int iters)
{
    if (iters < 0) {
        pdm_set_error(PARADMM_ERR_ARG, "iters < 0");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(get_state(G, &s, true, true));
    RC(run_fused(s, iters));
    if (iters > 0) {
        s->dev_newer = true;
        s->m_valid = false;
    }
    CK(hipStreamSynchronize(s->stream));
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_sync_to_host(graph* G)
{
    DevGraph* s;
    RC(get_state(G, &s, false));
    return sync_down(s);
}

extern "C" int pdm_sync_if_device_newer(graph* G)
{
    DevGraph* s = lookup(G);
    if (!s || !s->dev_newer) return PARADMM_OK;
    return sync_down(s);
}

extern "C" int paradmm_gpu_sync_from_host(graph* G)
{
    DevGraph* s;
    RC(get_state(G, &s, false));
    s->host_newer = true;
    return get_state(G, &s, true);
}

extern "C" int paradmm_gpu_read(graph* G, int which, double* out)
{
    if (!out || which < PARADMM_READ_Z || which > PARADMM_READ_RHO) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_read: which is PARADMM_READ_Z, _X, _U or _RHO and out not NULL");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(get_state(G, &s, true, which == PARADMM_READ_Z));
    RC(ensure_aux(s));
    if (which == PARADMM_READ_Z) {
        const long n = (long)s->V * s->D;
        k_flat_out<<<nblk(n), 256, 0, s->stream>>>(n, s->d_Z, s->d_stage);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(out, s->d_stage, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        CK(hipStreamSynchronize(s->stream));
        return PARADMM_OK;
    }
    // the dense layout holds U one update behind between iterations
    if (which == PARADMM_READ_U && s->layout == 2) RC(dense_finish(s));
    const int d = which == PARADMM_READ_RHO ? 1 : s->D;
    const float* src = which == PARADMM_READ_X ? s->d_X : which == PARADMM_READ_U ? s->d_U : s->d_rho;
    k_edges_out<<<nblk(s->E), 256, 0, s->stream>>>(s->E, d, src, s->stride, s->d_stage, s->d_pos);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(out, s->d_stage, sizeof(double) * (size_t)d * s->E, hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_release(graph* G)
{
    pdm_release_graph(G);
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_layout(graph* G)
{
    DevGraph* s = lookup(G);
    return s ? s->layout : 0;
}

extern "C" int paradmm_gpu_contact_info(graph* G, int* in_contact, int* tile_side, long* listed_pairs)
{
    DevGraph* s = lookup(G);
    int cs = 0;
    long n = 0;
    if (s && s->layout == 2) RC(dense_contact_info(s, &cs, &n));
    if (in_contact) *in_contact = cs > 0;
    if (tile_side) *tile_side = cs;
    if (listed_pairs) *listed_pairs = n;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_residuals(graph* G, double* primal2, double* dual2)
{
    DevGraph* s = lookup(G);
    if (!s || !s->iters_done) {
        pdm_set_error(PARADMM_ERR_STATE, "no device iteration has run on this graph");
        return PARADMM_ERR_STATE;
    }
    RC(residual_reduce(s));
    double h[2];
    CK(hipMemcpyAsync(h, s->d_res, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    if (primal2) *primal2 = h[0];
    if (dual2) *dual2 = h[1];
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_square_quality(graph* G, double* max_overlap, double* max_box, double* sum_r)
{
    DevGraph* s;
    RC(get_state(G, &s, true, true));
    double q[3];
    RC(square_quality(s, q));
    if (max_overlap) *max_overlap = q[0];
    if (max_box) *max_box = q[1];
    if (sum_r) *sum_r = q[2];
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_solve(graph* G, const paradmm_solve_opts* o, paradmm_solve_result* r)
{
    if (!o || !r || o->max_iters < 1 || o->check_every < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_solve: bad options");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(get_state(G, &s, true, true));
    if (o->tol_feas >= 0 && s->D != 3) {
        pdm_set_error(PARADMM_ERR_SHAPE, "feasibility check needs the dim-3 packing");
        return PARADMM_ERR_SHAPE;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const double scale = 1.0 / ((double)s->D * (double)s->E);
    int it = 0, since_policy = 0;
    r->converged = 0;
    r->max_overlap = r->max_box = r->objective = -1.0;
    while (it < o->max_iters) {
        int n = o->check_every < o->max_iters - it ? o->check_every : o->max_iters - it;
        if (o->policy_every > 0 && s->rho_policy) {
            // never run past a policy point inside a chunk
            const int left = o->policy_every - since_policy;
            if (left < n) n = left;
        }
        RC(run_fused(s, n));
        it += n;
        since_policy += n;
        RC(residual_reduce(s));
        double h[2];
        CK(hipMemcpyAsync(h, s->d_res, sizeof(h), hipMemcpyDeviceToHost, s->stream));
        CK(hipStreamSynchronize(s->stream));
        r->primal_rms = sqrt(h[0] * scale);
        r->dual_rms = sqrt(h[1] * scale);
        // a non-finite residual: the iterate has lost a value (diverged),
        // more iterations cannot bring it back
        if (!std::isfinite(r->primal_rms) || !std::isfinite(r->dual_rms)) break;
        if (r->primal_rms <= o->tol_primal && r->dual_rms <= o->tol_dual) {
            bool ok = true;
            if (o->tol_feas >= 0) {
                double q[3];
                RC(square_quality(s, q));
                r->max_overlap = q[0];
                r->max_box = q[1];
                r->objective = q[2];
                ok = q[0] <= o->tol_feas && q[1] <= o->tol_feas;
            }
            if (ok) {
                r->converged = 1;
                break;
            }
        }
        // the policy sets rho for the iterations that follow (none after the
        // last chunk); the residuals above do not depend on it
        if (o->policy_every > 0 && s->rho_policy && since_policy >= o->policy_every && it < o->max_iters) {
            s->policy_in_solve = true;
            const int prc = s->rho_policy(s, (float)it / (float)o->max_iters);
            s->policy_in_solve = false;
            RC(prc);
            since_policy = 0;
        }
    }
    if (o->tol_feas >= 0 && !r->converged) {
        double q[3];
        RC(square_quality(s, q));
        r->max_overlap = q[0];
        r->max_box = q[1];
        r->objective = q[2];
    }
    r->iterations = it;
    r->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->dev_newer = true;
    s->m_valid = false;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_time_iterations(graph* G, int iters, double* ms)
{
    DevGraph* s;
    RC(get_state(G, &s, true, true));
    // reach the steady state and record the runs outside the timed region
    int used = 0;
    RC(steady_start(s, 2, &used));
    if (!s->host) {
        hipGraphExec_t ex = nullptr;
        if (iters >= s->graph_iters) RC(captured_run(s, s->graph_iters, &ex));
        if (iters % s->graph_iters >= 2) RC(captured_run(s, iters % s->graph_iters, &ex));
    }
    CK(hipStreamSynchronize(s->stream));
    CK(hipEventRecord(s->ev0, s->stream));
    RC(run_fused(s, iters));
    CK(hipEventRecord(s->ev1, s->stream));
    CK(hipEventSynchronize(s->ev1));
    float t = 0.0f;
    CK(hipEventElapsedTime(&t, s->ev0, s->ev1));
    if (ms) *ms = t;
    s->dev_newer = true;
    s->m_valid = false;
    return PARADMM_OK;
}

namespace pdm {

// a new record with two events of the pool (false: no event left)
static bool prof_record(DevGraph* s, const char* name)
{
    if (s->prof_used + 2 > s->prof_pool.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return false;
            s->prof_pool.push_back(e);
        }
    }
    ProfRec r;
    r.name = name;
    r.a = s->prof_pool[s->prof_used++];
    r.b = s->prof_pool[s->prof_used++];
    s->prof_recs.push_back(r);
    return true;
}

void prof_begin(DevGraph* s, const char* name)
{
    if (!s->prof || !prof_record(s, name)) return;
    (void)hipEventRecord(s->prof_recs.back().a, s->stream);
}

void prof_end(DevGraph* s)
{
    if (!s->prof || s->prof_recs.empty()) return;
    (void)hipEventRecord(s->prof_recs.back().b, s->stream);
}

void prof_dispatch(DevGraph* s, const char* name, hipEvent_t& start, hipEvent_t& stop)
{
    start = stop = nullptr;
    if (!s->prof || !prof_record(s, name)) return;
    start = s->prof_recs.back().a;
    stop = s->prof_recs.back().b;
}

}  // namespace pdm

// Runs iters eager iterations (no graph) with events around every kernel and
// returns, per distinct kernel name, the mean duration in ms.  names receives
// max_kernels NUL-terminated strings of 64 bytes.
extern "C" int paradmm_gpu_profile_kernels(graph* G, int iters, int max_kernels, char* names, double* avg_ms,
                                           int* count, int* nkernels)
{
    DevGraph* s;
    RC(get_state(G, &s, true, true));
    int used = 0;
    RC(steady_start(s, 2, &used));
    s->prof = true;
    s->prof_recs.clear();
    s->prof_used = 0;
    int rc = PARADMM_OK;
    // the kernel sequence of a captured run of iters iterations
    for (int i = 0; i < iters && !rc; ++i) rc = fused_iteration(s, 0, mphase_of(s, i, iters));
    // with a rho policy attached: one application as a solve makes it
    // (iterations follow it), then one more iteration
    if (!rc && s->rho_policy && iters > 0) {
        s->policy_in_solve = true;
        rc = s->rho_policy(s, 0.5f);
        s->policy_in_solve = false;
        if (!rc) rc = fused_iteration(s, 0, DENSE_PLAIN);
    }
    s->prof = false;
    RC(rc);
    CK(hipStreamSynchronize(s->stream));
    std::vector<std::string> nm;
    std::vector<double> tot;
    std::vector<int> cnt;
    for (const ProfRec& r : s->prof_recs) {
        float t = 0.0f;
        CK(hipEventElapsedTime(&t, r.a, r.b));
        size_t j = 0;
        while (j < nm.size() && nm[j] != r.name) ++j;
        if (j == nm.size()) { nm.push_back(r.name); tot.push_back(0.0); cnt.push_back(0); }
        tot[j] += t;
        cnt[j] += 1;
    }
    s->prof_recs.clear();
    s->prof_used = 0;
    const int k = (int)nm.size() < max_kernels ? (int)nm.size() : max_kernels;
    for (int j = 0; j < k; ++j) {
        snprintf(names + 64 * j, 64, "%s", nm[j].c_str());
        avg_ms[j] = tot[j] / cnt[j];
        if (count) count[j] = cnt[j];
    }
    if (nkernels) *nkernels = k;
    s->iters_done = true;
    s->dev_newer = true;
    s->m_valid = false;
    return PARADMM_OK;
}

namespace pdm {

int residual_reduce(DevGraph* s)
{
    if (s->layout == 2) return dense_residual_reduce(s);
    pdm_k_res_final<<<1, 256, 0, s->stream>>>(s->V, s->d_res_part, s->d_res);
    CK(hipGetLastError());
    return PARADMM_OK;
}

// initialize_X_N_Z_M_U_optimal sweeps: X (fixed) holds Zinit per edge,
// N = X - U.  Each sweep: Xit = prox(N) into the M buffer, U = Xit - N,
// N = X - U.  Finally M = X + U.
__global__ __launch_bounds__(256) void k_warm_step(long n, const float* __restrict__ X, const float* __restrict__ Xit,
                                                   float* __restrict__ U, float* __restrict__ N)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float u = Xit[i] - N[i];
    U[i] = u;
    N[i] = X[i] - u;
}

int warm_start_device(DevGraph* s, int maxiter)
{
    RC(ensure_aux(s));
    const long n = (long)s->D * s->stride;
    // the prox writes its output to d_X; keep the fixed X in the stage area
    float* Xfix = (float*)s->d_stage;
    CK(hipMemcpyAsync(Xfix, s->d_X, sizeof(float) * n, hipMemcpyDeviceToDevice, s->stream));
    for (int it = 0; it < maxiter; ++it) {
        if (s->layout == 2) RC(dense_legacy_xpass(s));
        else RC(launch_xpass_generic(s, 1));
        k_warm_step<<<nblk(n), 256, 0, s->stream>>>(n, Xfix, s->d_X, s->d_U, s->d_N);
        CK(hipGetLastError());
    }
    CK(hipMemcpyAsync(s->d_X, Xfix, sizeof(float) * n, hipMemcpyDeviceToDevice, s->stream));
    k_mpass<<<nblk(n), 256, 0, s->stream>>>(n, s->d_X, s->d_U, s->d_M);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s->stream));
    return PARADMM_OK;
}

}  // namespace pdm

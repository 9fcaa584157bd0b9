// mlp.hip -- the learned rho policy: a 6 -> 32 -> 32 -> 1 ReLU MLP whose
// forward and backward passes run on bf16 MFMA (v_mfma_f32_16x16x16_bf16)
// with fp32 accumulation, trained with Adam on fp32 master weights.
//
// No reference counterpart: the reference's rho hook (core/paradmm.c:564-566)
// is empty.  Parameter vector (1313 fp32): W1[32][6], b1[32], W2[32][32],
// b2[32], W3[32], b3.
//
// MFMA 16x16x16 operand layout (wave64, 4 elements per lane):
//   A[i][k]: lane i + 16*(k/4), element k%4      (i row, k reduction)
//   B[k][j]: lane j + 16*(k/4), element k%4      (j column)
//   D[i][j]: lane j + 16*(i/4), element i%4
// A wave pushes tiles of 16 rows through the network; activations move from
// the D layout of one layer to the A layout of the next through LDS.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <cmath>
#include <algorithm>
#include <random>
#include <vector>

#include "engine_internal.h"
#include "engine_state.h"
#include "mlp.h"
#include "mlp_device.h"
#include "paradmm_gpu.h"

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

namespace pdm {


// y = MLP(features) for n rows of 6 fp32 features.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ feat, int n,
                                                     const float* __restrict__ params, float* __restrict__ out)
{
    __shared__ MlpShared w;
    __shared__ TileScratch tsw[4];
    stage_weights(params, w);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    TileScratch& ts = tsw[wv];
    const long ntiles = ((long)n + 15) / 16;
    for (long tile = (long)blockIdx.x * 4 + wv; tile < ntiles; tile += (long)gridDim.x * 4) {
        const long r0 = tile * 16;
        for (int q = l; q < 256; q += 64) {
            const int i = q / 16, k = q % 16;
            const long row = r0 + i;
            ts.f[i][k] = (k < MLP_NIN && row < n) ? f2bf(feat[row * MLP_NIN + k]) : (short)0;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        v4f z1[2], z2[2];
        tile_forward(w, ts, z1, z2);
        if (l < 16 && r0 + l < n) out[r0 + l] = ts.y[l];
        __builtin_amdgcn_wave_barrier();
    }
}

// Partial gradient of 0.5 * sum((y - t)^2) / inv_b over this block's rows.
// grad_part[blockIdx.x][MLP_NPARAMS].
__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ feat, const float* __restrict__ target,
                                                  int n, float inv_b, const float* __restrict__ params,
                                                  float* __restrict__ grad_part, double* __restrict__ loss_part)
{
    __shared__ MlpShared w;
    __shared__ TileScratch tsw[4];
    __shared__ float red[4][MLP_NPARAMS_PAD];
    __shared__ double lred[4];
    stage_weights(params, w);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int li = l & 15, lq = l >> 4;
    TileScratch& ts = tsw[wv];
    v4f gW2[2][2], gW1[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        gW1[a] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 2; ++b) gW2[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float gb1[2] = {0.0f, 0.0f}, gb2[2] = {0.0f, 0.0f}, gW3[2] = {0.0f, 0.0f}, gb3 = 0.0f;
    double loss = 0.0;
    const long ntiles = ((long)n + 15) / 16;
    for (long tile = (long)blockIdx.x * 4 + wv; tile < ntiles; tile += (long)gridDim.x * 4) {
        const long r0 = tile * 16;
        for (int q = l; q < 256; q += 64) {
            const int i = q / 16, k = q % 16;
            const long row = r0 + i;
            ts.f[i][k] = (k < MLP_NIN && row < n) ? f2bf(feat[row * MLP_NIN + k]) : (short)0;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        v4f z1[2], z2[2];
        tile_forward(w, ts, z1, z2);
        // dy per row (zero for padding rows), broadcast through LDS
        if (l < 16) {
            const long row = r0 + l;
            float dy = 0.0f;
            if (row < n) {
                const float d = ts.y[l] - target[row];
                loss += 0.5 * (double)d * (double)d;
                dy = d * inv_b;
            }
            ts.y[l] = dy;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW3[u] += sum_rows dy h2[u]; db3 += sum dy.  Lane li owns u = nt*16 + li over
        // rows 4*lq .. 4*lq+3.
        if (l == 0)
#pragma unroll
            for (int i = 0; i < 16; ++i) gb3 += ts.y[i];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int u = nt * 16 + li;
            const float w3u = bf2f(w.w3[u]);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = 4 * lq + e;
                const float dy = ts.y[i];
                gW3[nt] += dy * bf2f(ts.h2[i][u]);
                const float dz = z2[nt][e] > 0.0f ? dy * w3u : 0.0f;
                gb2[nt] += dz;
                ts.dz[i][u] = f2bf(dz);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW2[mt*16+i][nt*16+j] += sum_rows dz2[row][.] h1[row][.]  (K = the 16 rows)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            v4s aa;
#pragma unroll
            for (int e = 0; e < 4; ++e) aa[e] = ts.dz[4 * lq + e][mt * 16 + li];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                v4s bb;
#pragma unroll
                for (int e = 0; e < 4; ++e) bb[e] = ts.h1[4 * lq + e][nt * 16 + li];
                gW2[mt][nt] = mfma16(aa, bb, gW2[mt][nt]);
            }
        }
        // dh1 = dz2 (rows x 32 out) * W2 (32 out x 32 in), then dz1
        v4f dh1[2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                v4s aa, bb;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    aa[e] = ts.dz[li][ks * 16 + 4 * lq + e];
                    bb[e] = w.w2[ks * 16 + 4 * lq + e][nt * 16 + li];
                }
                acc = mfma16(aa, bb, acc);
            }
            dh1[nt] = acc;
        }
        __builtin_amdgcn_wave_barrier();  // all lanes have read dz2 before it is overwritten
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int u = nt * 16 + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float dz = z1[nt][e] > 0.0f ? dh1[nt][e] : 0.0f;
                gb1[nt] += dz;
                ts.dz[4 * lq + e][u] = f2bf(dz);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW1[mt*16+i][k] += sum_rows dz1[row][.] f[row][k]
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            v4s aa, bb;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                aa[e] = ts.dz[4 * lq + e][mt * 16 + li];
                bb[e] = ts.f[4 * lq + e][li];
            }
            gW1[mt] = mfma16(aa, bb, gW1[mt]);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // gather this wave's partials into red[wv][...] (parameter order)
    float* g = red[wv];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int u = mt * 16 + 4 * lq + e, k = li;
            if (k < MLP_NIN) g[MLP_OFF_W1 + u * MLP_NIN + k] = gW1[mt][e];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) g[MLP_OFF_W2 + u * MLP_NH + nt * 16 + li] = gW2[mt][nt][e];
        }
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        // sum the four lane groups (rows) of each column
        float b1v = gb1[nt], b2v = gb2[nt], w3v = gW3[nt];
        b1v += __shfl_xor(b1v, 16, 64); b1v += __shfl_xor(b1v, 32, 64);
        b2v += __shfl_xor(b2v, 16, 64); b2v += __shfl_xor(b2v, 32, 64);
        w3v += __shfl_xor(w3v, 16, 64); w3v += __shfl_xor(w3v, 32, 64);
        if (lq == 0) {
            g[MLP_OFF_B1 + nt * 16 + li] = b1v;
            g[MLP_OFF_B2 + nt * 16 + li] = b2v;
            g[MLP_OFF_W3 + nt * 16 + li] = w3v;
        }
    }
    if (l == 0) g[MLP_OFF_B3] = gb3;
    // loss: lanes 0..15 hold per-row terms
    double lw = loss;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) lw += __shfl_xor(lw, o, 64);
    if (l == 0) lred[wv] = lw;
    __syncthreads();
    for (int q = threadIdx.x; q < MLP_NPARAMS; q += blockDim.x)
        grad_part[(long)blockIdx.x * MLP_NPARAMS + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (lred[0] + lred[1]) + (lred[2] + lred[3]);
}

// grad = sum over blocks (fixed order); loss likewise
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ grad,
                                                    const double* __restrict__ lpart, double* __restrict__ loss)
{
    for (int q = blockIdx.x * 256 + threadIdx.x; q < MLP_NPARAMS; q += gridDim.x * 256) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += part[(long)b * MLP_NPARAMS + q];
        grad[q] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += lpart[b];
        *loss = s;
    }
}

// Adam (Kingma & Ba 2015) on the fp32 master weights
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                                  float* __restrict__ v, float lr, float c1, float c2)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= MLP_NPARAMS) return;
    const float gq = g[q];
    const float mq = 0.9f * m[q] + 0.1f * gq;
    const float vq = 0.999f * v[q] + 0.001f * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq * c1) / (sqrtf(vq * c2) + 1e-8f);
}

// rows of a minibatch gathered by index
__global__ __launch_bounds__(256) void k_mlp_gather(const float* __restrict__ feat, const float* __restrict__ target,
                                                    const int* __restrict__ idx, int n, float* __restrict__ fo,
                                                    float* __restrict__ to)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const long s = idx[r];
#pragma unroll
    for (int k = 0; k < MLP_NIN; ++k) fo[(long)r * MLP_NIN + k] = feat[s * MLP_NIN + k];
    to[r] = target[s];
}

// One MFMA on known operands: D = A x B with A[i][k] = bf16(i - k/4),
// B[k][j] = bf16((k + 2 j) / 8); written row-major [16][16].
__global__ void k_mfma_layout_probe(float* __restrict__ out)
{
    const int l = threadIdx.x & 63, li = l & 15, lq = l >> 4;
    v4s a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = 4 * lq + e;
        a[e] = f2bf((float)li - (float)k * 0.25f);
        b[e] = f2bf((float)(k + 2 * li) * 0.125f);
    }
    const v4f d = mfma16(a, b, v4f{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
    for (int e = 0; e < 4; ++e) out[(4 * lq + e) * 16 + li] = d[e];
}

}  // namespace pdm

using namespace pdm;

struct paradmm_rho_mlp {
    float* d_p = nullptr;   // fp32 master parameters
    float* d_m = nullptr;   // Adam moments
    float* d_v = nullptr;
    float* d_g = nullptr;   // reduced gradient
    float* d_gpart = nullptr;
    double* d_lpart = nullptr;
    double* d_loss = nullptr;
    int gblocks = 0;        // capacity of the partial buffers
    int step = 0;
    hipStream_t stream = nullptr;
};

static int mlp_fail(hipError_t e, const char* what)
{
    pdm_set_error(PARADMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return PARADMM_ERR_HIP;
}

#define MCK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return mlp_fail(e_, #x); } while (0)

static int mlp_alloc(paradmm_rho_mlp* m)
{
    MCK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    MCK(hipMalloc(&m->d_p, sizeof(float) * MLP_NPARAMS));
    MCK(hipMalloc(&m->d_m, sizeof(float) * MLP_NPARAMS));
    MCK(hipMalloc(&m->d_v, sizeof(float) * MLP_NPARAMS));
    MCK(hipMalloc(&m->d_g, sizeof(float) * MLP_NPARAMS));
    MCK(hipMalloc(&m->d_loss, sizeof(double)));
    MCK(hipMemset(m->d_m, 0, sizeof(float) * MLP_NPARAMS));
    MCK(hipMemset(m->d_v, 0, sizeof(float) * MLP_NPARAMS));
    return PARADMM_OK;
}

static int mlp_ensure_blocks(paradmm_rho_mlp* m, int nblk)
{
    if (nblk <= m->gblocks) return PARADMM_OK;
    if (m->d_gpart) MCK(hipFree(m->d_gpart));
    if (m->d_lpart) MCK(hipFree(m->d_lpart));
    MCK(hipMalloc(&m->d_gpart, sizeof(float) * MLP_NPARAMS * (size_t)nblk));
    MCK(hipMalloc(&m->d_lpart, sizeof(double) * (size_t)nblk));
    m->gblocks = nblk;
    return PARADMM_OK;
}

extern "C" void paradmm_rho_mlp_destroy(paradmm_rho_mlp* m)
{
    if (!m) return;
    (void)hipFree(m->d_p); (void)hipFree(m->d_m); (void)hipFree(m->d_v); (void)hipFree(m->d_g);
    (void)hipFree(m->d_gpart); (void)hipFree(m->d_lpart); (void)hipFree(m->d_loss);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" int paradmm_rho_mlp_create(paradmm_rho_mlp** out, const float* params, int nparams, uint64_t seed)
{
    if (!out || (params && nparams != MLP_NPARAMS)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_create: need %d parameters", MLP_NPARAMS);
        return PARADMM_ERR_ARG;
    }
    paradmm_rho_mlp* m = new paradmm_rho_mlp();
    int rc = mlp_alloc(m);
    if (rc) { paradmm_rho_mlp_destroy(m); return rc; }
    std::vector<float> p(MLP_NPARAMS, 0.0f);
    if (params) {
        memcpy(p.data(), params, sizeof(float) * MLP_NPARAMS);
    } else {
        // He-normal weights, zero biases
        std::mt19937_64 gen(seed);
        std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / MLP_NIN)), n2(0.0f, std::sqrt(2.0f / MLP_NH)),
            n3(0.0f, std::sqrt(1.0f / MLP_NH));
        for (int q = 0; q < MLP_NH * MLP_NIN; ++q) p[MLP_OFF_W1 + q] = n1(gen);
        for (int q = 0; q < MLP_NH * MLP_NH; ++q) p[MLP_OFF_W2 + q] = n2(gen);
        for (int q = 0; q < MLP_NH; ++q) p[MLP_OFF_W3 + q] = n3(gen);
    }
    hipError_t e = hipMemcpy(m->d_p, p.data(), sizeof(float) * MLP_NPARAMS, hipMemcpyHostToDevice);
    if (e != hipSuccess) { paradmm_rho_mlp_destroy(m); return mlp_fail(e, "hipMemcpy"); }
    *out = m;
    return PARADMM_OK;
}

extern "C" int paradmm_rho_mlp_get_params(paradmm_rho_mlp* m, float* params, int nparams)
{
    if (!m || !params || nparams != MLP_NPARAMS) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_get_params: need %d floats", MLP_NPARAMS);
        return PARADMM_ERR_ARG;
    }
    MCK(hipMemcpy(params, m->d_p, sizeof(float) * MLP_NPARAMS, hipMemcpyDeviceToHost));
    return PARADMM_OK;
}

namespace pdm {
const float* mlp_device_params(paradmm_rho_mlp* m) { return m ? m->d_p : nullptr; }
}

extern "C" int paradmm_rho_mlp_forward(paradmm_rho_mlp* m, const float* feat, int n, float* out)
{
    if (!m || !feat || !out || n < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_forward: bad arguments");
        return PARADMM_ERR_ARG;
    }
    float *df = nullptr, *dy = nullptr;
    MCK(hipMalloc(&df, sizeof(float) * MLP_NIN * (size_t)n));
    hipError_t e = hipMalloc(&dy, sizeof(float) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(df, feat, sizeof(float) * MLP_NIN * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        const long ntiles = ((long)n + 15) / 16;
        const unsigned grid = (unsigned)std::min<long>((ntiles + 3) / 4, 4096);
        k_mlp_forward<<<grid, 256, 0, m->stream>>>(df, n, m->d_p, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(df);
    (void)hipFree(dy);
    if (e != hipSuccess) return mlp_fail(e, "paradmm_rho_mlp_forward");
    return PARADMM_OK;
}

// gradient of 0.5 * mean((y - t)^2) over rows already on the device
static int device_gradient(paradmm_rho_mlp* m, const float* dfeat, const float* dtarget, int n)
{
    const long ntiles = ((long)n + 15) / 16;
    const int nblk = (int)std::min<long>((ntiles + 3) / 4, 1024);
    RC(mlp_ensure_blocks(m, nblk));
    k_mlp_grad<<<nblk, 256, 0, m->stream>>>(dfeat, dtarget, n, 1.0f / (float)n, m->d_p, m->d_gpart, m->d_lpart);
    MCK(hipGetLastError());
    k_mlp_reduce<<<(MLP_NPARAMS + 255) / 256, 256, 0, m->stream>>>(m->d_gpart, nblk, m->d_g, m->d_lpart, m->d_loss);
    MCK(hipGetLastError());
    return PARADMM_OK;
}

extern "C" int paradmm_rho_mlp_gradient(paradmm_rho_mlp* m, const float* feat, const float* target, int n,
                                        float* grad, double* loss)
{
    if (!m || !feat || !target || !grad || n < 1) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_gradient: bad arguments");
        return PARADMM_ERR_ARG;
    }
    float *df = nullptr, *dt = nullptr;
    MCK(hipMalloc(&df, sizeof(float) * MLP_NIN * (size_t)n));
    MCK(hipMalloc(&dt, sizeof(float) * (size_t)n));
    int rc = PARADMM_OK;
    hipError_t e = hipMemcpyAsync(df, feat, sizeof(float) * MLP_NIN * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, target, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) rc = mlp_fail(e, "hipMemcpy");
    if (!rc) rc = device_gradient(m, df, dt, n);
    double l = 0.0;
    if (!rc) {
        e = hipMemcpyAsync(grad, m->d_g, sizeof(float) * MLP_NPARAMS, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) rc = mlp_fail(e, "hipMemcpy");
    }
    if (loss) *loss = l / n;
    (void)hipFree(df);
    (void)hipFree(dt);
    return rc;
}

extern "C" int paradmm_rho_mlp_train(paradmm_rho_mlp* m, const float* feat, const float* target, int n, int epochs,
                                     int batch, float lr, uint64_t seed, double* final_loss)
{
    if (!m || !feat || !target || n < 1 || epochs < 1 || batch < 16) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_train: bad arguments");
        return PARADMM_ERR_ARG;
    }
    float *df = nullptr, *dt = nullptr, *bf = nullptr, *bt = nullptr;
    int* didx = nullptr;
    MCK(hipMalloc(&df, sizeof(float) * MLP_NIN * (size_t)n));
    MCK(hipMalloc(&dt, sizeof(float) * (size_t)n));
    MCK(hipMalloc(&bf, sizeof(float) * MLP_NIN * (size_t)batch));
    MCK(hipMalloc(&bt, sizeof(float) * (size_t)batch));
    MCK(hipMalloc(&didx, sizeof(int) * (size_t)n));
    int rc = PARADMM_OK;
    hipError_t e = hipMemcpyAsync(df, feat, sizeof(float) * MLP_NIN * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, target, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) rc = mlp_fail(e, "hipMemcpy");
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::mt19937_64 gen(seed);
    double epoch_loss = 0.0;
    for (int ep = 0; ep < epochs && !rc; ++ep) {
        std::shuffle(idx.begin(), idx.end(), gen);
        e = hipMemcpyAsync(didx, idx.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->stream);
        if (e != hipSuccess) { rc = mlp_fail(e, "hipMemcpy"); break; }
        e = hipStreamSynchronize(m->stream);  // idx is reused next epoch
        if (e != hipSuccess) { rc = mlp_fail(e, "sync"); break; }
        epoch_loss = 0.0;
        for (int s0 = 0; s0 < n && !rc; s0 += batch) {
            const int nb = std::min(batch, n - s0);
            k_mlp_gather<<<(nb + 255) / 256, 256, 0, m->stream>>>(df, dt, didx + s0, nb, bf, bt);
            if ((e = hipGetLastError()) != hipSuccess) { rc = mlp_fail(e, "gather"); break; }
            rc = device_gradient(m, bf, bt, nb);
            if (rc) break;
            m->step += 1;
            const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->step));
            const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->step));
            k_mlp_adam<<<(MLP_NPARAMS + 255) / 256, 256, 0, m->stream>>>(m->d_p, m->d_g, m->d_m, m->d_v, lr, c1, c2);
            if ((e = hipGetLastError()) != hipSuccess) { rc = mlp_fail(e, "adam"); break; }
            if (ep == epochs - 1) {
                double l = 0.0;
                e = hipMemcpyAsync(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
                if (e != hipSuccess) { rc = mlp_fail(e, "loss"); break; }
                epoch_loss += l;
            }
        }
    }
    if (!rc) {
        e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) rc = mlp_fail(e, "sync");
    }
    if (final_loss) *final_loss = epoch_loss / n;
    (void)hipFree(df); (void)hipFree(dt); (void)hipFree(bf); (void)hipFree(bt); (void)hipFree(didx);
    return rc;
}

extern "C" int paradmm_rho_mlp_layout_probe(float* out256)
{
    if (!out256) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_rho_mlp_layout_probe: NULL");
        return PARADMM_ERR_ARG;
    }
    float* d = nullptr;
    MCK(hipMalloc(&d, sizeof(float) * 256));
    k_mfma_layout_probe<<<1, 64>>>(d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out256, d, sizeof(float) * 256, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return mlp_fail(e, "layout probe");
    return PARADMM_OK;
}

// ------------------------------------------------------------------ rho policy on a graph

// mode 3: an edge keeps its own network rho when that exceeds the graph's
// rho by more than this factor (log10; 10^0.5 = 3.2)
#define RHO_RAISE_LOG10 0.5f

namespace pdm {

struct PolicyArgs {
    int E, D;
    long stride;
    const int* pos;
    const int* edge_var;
    const unsigned char* type;
    const float* X;
    float* U;
    float* rho;
    const float* Z;
    const float* Zprev;
    float frac;
    float* feat_out;  // features only (training data) when non-null
    // device-slot order (policy application): variable and type of every
    // slot, variable -1 for padding slots
    long nslots;
    const int* slot_var;
    const unsigned char* slot_type;
    float* yout;  // per-slot network output (mode 3), or null
    int* nraise;  // edges given their own rho by the last application (mode 3)
    int* tflag;          // dense layout: launch-order tile list, .y carries TILE_RAISED (mode 3), or null
    const int* tile_pos; // layout tile -> launch position
    int tile_slots, ntiles;
};

// residuals of the edge in device slot p (coalesced in slot order)
__device__ __forceinline__ void slot_residuals(const PolicyArgs& a, long p, int v, float& r, float& s, float& rho)
{
    // no contraction: the contact state's own application (dense.hip)
    // computes the same inputs bit for bit
#pragma clang fp contract(off)
    float r2 = 0.0f, dz2 = 0.0f;
    for (int k = 0; k < a.D; ++k) {
        const float z = a.Z[v * a.D + k];
        const float d = a.X[k * a.stride + p] - z;
        const float q = z - a.Zprev[v * a.D + k];
        r2 += d * d;
        dz2 += q * q;
    }
    rho = a.rho[p];
    r = sqrtf(r2);
    s = rho * sqrtf(dz2);
}

// slot -> variable / type maps from the api edge maps
__global__ __launch_bounds__(256) void k_build_slot_maps(int E, const int* __restrict__ pos,
                                                         const int* __restrict__ edge_var,
                                                         const unsigned char* __restrict__ type,
                                                         int* __restrict__ slot_var, unsigned char* __restrict__ slot_type)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    const int p = pos[e];
    slot_var[p] = edge_var[e];
    slot_type[p] = type[e];
}

__device__ __forceinline__ void edge_residuals(const PolicyArgs& a, int e, float& r, float& s, float& rho, int& type,
                                               int& p)
{
#pragma clang fp contract(off)
    p = a.pos[e];
    const int v = a.edge_var[e];
    float r2 = 0.0f, dz2 = 0.0f;
    for (int k = 0; k < a.D; ++k) {
        const float z = a.Z[v * a.D + k];
        const float d = a.X[k * a.stride + p] - z;
        const float q = z - a.Zprev[v * a.D + k];
        r2 += d * d;
        dz2 += q * q;
    }
    rho = a.rho[p];
    r = sqrtf(r2);
    s = rho * sqrtf(dz2);
    type = a.type[e];
}

__global__ __launch_bounds__(256) void k_rho_features(PolicyArgs a)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.E) return;
    float r, s, rho, f[MLP_NIN];
    int type, p;
    edge_residuals(a, e, r, s, rho, type, p);
    edge_features(r, s, rho, type, a.frac, f);
#pragma unroll
    for (int k = 0; k < MLP_NIN; ++k) a.feat_out[(long)e * MLP_NIN + k] = f[k];
}

// Per-edge mode: rho_e <- 10^y_e, U_e <- U_e rho_old / rho_new.
__global__ __launch_bounds__(256) void k_rho_policy(PolicyArgs a, const float* __restrict__ params)
{
    __shared__ MlpShared w;
    __shared__ TileScratch tsw[4];
    __shared__ float yv[4][64];
    stage_weights(params, w);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    TileScratch& ts = tsw[wv];
    const long nchunks = (a.nslots + 63) / 64;
    for (long ch = (long)blockIdx.x * 4 + wv; ch < nchunks; ch += (long)gridDim.x * 4) {
        const long p = ch * 64 + l;
        const int v = p < a.nslots ? a.slot_var[p] : -1;
        const bool valid = v >= 0;
        float r = 0.0f, s = 0.0f, rho = 1.0f, f[MLP_NIN];
        int type = 0;
        if (valid) {
            slot_residuals(a, p, v, r, s, rho);
            type = a.slot_type[p];
        }
        edge_features(r, s, rho, type, a.frac, f);
        const float y = policy_wave_eval(w, ts, yv[wv], f, l);
        if (valid) {
            const float rn = exp10f(y);
            const float scale = rho / rn;
            for (int k = 0; k < a.D; ++k) a.U[k * a.stride + p] *= scale;
            a.rho[p] = rn;
        }
    }
}

// Per-type mode, pass 1: the network's log10 rho of every edge, summed per
// factor type (pair, container, radius) into this workgroup's FP64 partial
// slots part[block][3] and counts cnt[block][3].
__global__ __launch_bounds__(256) void k_rho_policy_type_sums(PolicyArgs a, const float* __restrict__ params,
                                                              double* __restrict__ part, int* __restrict__ cnt)
{
    __shared__ MlpShared w;
    __shared__ TileScratch tsw[4];
    __shared__ float yv[4][64];
    __shared__ double ssum[4][3];
    __shared__ int scnt[4][3];
    // mode 3: the raised-tile flags of the last application are cleared here
    // (the apply pass, launched after this one, sets the new ones)
    if (a.tflag)
        for (int k = blockIdx.x * 256 + threadIdx.x; k < a.ntiles; k += gridDim.x * 256)
            a.tflag[2 * k + 1] &= ~TILE_RAISED;
    stage_weights(params, w);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    TileScratch& ts = tsw[wv];
    double acc[3] = {0.0, 0.0, 0.0};
    int num[3] = {0, 0, 0};
    const long nchunks = (a.nslots + 63) / 64;
    for (long ch = (long)blockIdx.x * 4 + wv; ch < nchunks; ch += (long)gridDim.x * 4) {
        const long p = ch * 64 + l;
        const int v = p < a.nslots ? a.slot_var[p] : -1;
        const bool valid = v >= 0;
        float r = 0.0f, s = 0.0f, rho = 1.0f, f[MLP_NIN];
        int type = 0;
        if (valid) {
            slot_residuals(a, p, v, r, s, rho);
            type = a.slot_type[p];
        }
        edge_features(r, s, rho, type, a.frac, f);
        const float y = policy_wave_eval(w, ts, yv[wv], f, l);
        if (valid && a.yout) a.yout[p] = y;
        if (valid) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                acc[k] += k == type ? (double)y : 0.0;
                num[k] += k == type ? 1 : 0;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
            acc[k] += __shfl_xor(acc[k], o, 64);
            num[k] += __shfl_xor(num[k], o, 64);
        }
    }
    if (l == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ssum[wv][k] = acc[k];
            scnt[wv][k] = num[k];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        part[blockIdx.x * 3 + k] = (ssum[0][k] + ssum[1][k]) + (ssum[2][k] + ssum[3][k]);
        cnt[blockIdx.x * 3 + k] = scnt[0][k] + scnt[1][k] + scnt[2][k] + scnt[3][k];
    }
}

// Per-type mode, pass 2: mean log10 rho per type over the partials in fixed
// order; rho_type = 10^mean (types without edges keep 0 and are unused).
// global: one mean over all edges for every type.  hist (optional):
// running sums {sum_k mean_k, count} per type over the applications so far;
// the applied value is then the average of this and the earlier estimates.
__global__ __launch_bounds__(256) void k_rho_policy_type_final(int nblocks, const double* __restrict__ part,
                                                               const int* __restrict__ cnt,
                                                               float* __restrict__ rho_type, int global,
                                                               double* __restrict__ hist)
{
    // thread t sums the partials of blocks t, t + 256, ... (fixed order),
    // then a fixed-shape tree over the 256 threads
    __shared__ double ssum[3][256];
    __shared__ long scnt[3][256];
    const int t = threadIdx.x;
    double sum[3] = {0.0, 0.0, 0.0};
    long n[3] = {0, 0, 0};
    for (int b = t; b < nblocks; b += 256)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            sum[q] += part[b * 3 + q];
            n[q] += cnt[b * 3 + q];
        }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ssum[q][t] = sum[q];
        scnt[q][t] = n[q];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                ssum[q][t] += ssum[q][t + w];
                scnt[q][t] += scnt[q][t + w];
            }
        __syncthreads();
    }
    if (t >= 3) return;
    const int k = t;
    double S = 0.0;
    long Nn = 0;
    for (int q = 0; q < 3; ++q)
        if (global || q == k) {
            S += ssum[q][0];
            Nn += scnt[q][0];
        }
    if (!Nn) {
        rho_type[k] = 0.0f;
        return;
    }
    double m = S / (double)Nn;
    if (hist) {
        hist[2 * k] += m;
        hist[2 * k + 1] += 1.0;
        m = hist[2 * k] / hist[2 * k + 1];
    }
    rho_type[k] = exp10f((float)m);
}

// Per-type mode, pass 3: rho_e <- rho_type(e), U_e <- U_e rho_old / rho_new.
__global__ __launch_bounds__(256) void k_rho_policy_type_apply(PolicyArgs a, const float* __restrict__ rho_type)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= a.nslots || a.slot_var[p] < 0) return;
    const float rn = rho_type[a.slot_type[p]];
    const float scale = a.rho[p] / rn;
    for (int k = 0; k < a.D; ++k) a.U[k * a.stride + p] *= scale;
    a.rho[p] = rn;
}

// Graph mode with raises, pass 3: rho_e <- 10^y_e where the network asks for
// more than 10^raise_log times the graph's rho (an edge it singles out),
// else the graph's rho; U_e <- U_e rho_old / rho_new.  Counts the raised
// edges (one atomic per wavefront: the count does not depend on the order).
__global__ __launch_bounds__(256) void k_rho_policy_raise_apply(PolicyArgs a, const float* __restrict__ rho_type,
                                                                float raise_log)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < a.nslots && a.slot_var[p] >= 0;
    bool raised = false;
    if (valid) {
        const float g = rho_type[a.slot_type[p]];
        const float y = a.yout[p];
        raised = y - log10f(g) > raise_log;
        const float rn = raised ? exp10f(y) : g;
        const float scale = a.rho[p] / rn;
        for (int k = 0; k < a.D; ++k) a.U[k * a.stride + p] *= scale;
        a.rho[p] = rn;
        if (raised && a.tflag) atomicOr(a.tflag + 2L * a.tile_pos[p / a.tile_slots] + 1, TILE_RAISED);
    }
    const unsigned long long b = __ballot(raised);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(a.nraise, (int)__popcll(b));
}

// uniform rho multiplier with the scaled dual rescaled: rho *= f, U /= f
__global__ __launch_bounds__(256) void k_scale_rho(long n, int D, long stride, float f, float* __restrict__ rho,
                                                   float* __restrict__ U)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    rho[p] *= f;
    const float inv = 1.0f / f;
    for (int k = 0; k < D; ++k) U[k * stride + p] *= inv;
}

static PolicyArgs policy_args(DevGraph* s, float frac)
{
    PolicyArgs a;
    a.E = s->E;
    a.D = s->D;
    a.stride = s->stride;
    a.pos = s->d_pos;
    a.edge_var = s->d_edge_var_api;
    a.type = s->d_edge_type_api;
    a.X = s->d_X;
    a.U = s->d_U;
    a.rho = s->d_rho;
    a.Z = s->d_Z;
    a.Zprev = s->d_Zprev;
    a.frac = frac;
    a.feat_out = nullptr;
    a.nslots = s->stride;
    a.slot_var = s->d_slot_var;
    a.slot_type = s->d_slot_type;
    a.yout = nullptr;
    a.nraise = nullptr;
    a.tflag = nullptr;
    a.tile_pos = nullptr;
    a.tile_slots = 1;
    a.ntiles = 0;
    return a;
}

// slot maps for the policy kernels, built on first use
static int ensure_slot_maps(DevGraph* s)
{
    if (s->d_slot_var) return PARADMM_OK;
    RC(dmalloc(s, (void**)&s->d_slot_var, sizeof(int) * (size_t)s->stride));
    RC(dmalloc(s, (void**)&s->d_slot_type, (size_t)s->stride));
    CK(hipMemsetAsync(s->d_slot_var, 0xff, sizeof(int) * (size_t)s->stride, s->stream));
    CK(hipMemsetAsync(s->d_slot_type, 0, (size_t)s->stride, s->stream));
    k_build_slot_maps<<<(unsigned)((s->E + 255) / 256), 256, 0, s->stream>>>(s->E, s->d_pos, s->d_edge_var_api,
                                                                           s->d_edge_type_api, s->d_slot_var,
                                                                           s->d_slot_type);
    CK(hipGetLastError());
    return PARADMM_OK;
}

// the per-type rho of the last per-type / global application
static float* rho_type_ptr(DevGraph* s)
{
    double* part = (double*)s->d_policy_part + 6;
    return (float*)((int*)(part + 3 * 8192) + 3 * 8192);
}

// makes U current (dense layout) so it can be rescaled
static int current_u(DevGraph* s)
{
    if (s->layout == 2) RC(dense_finish(s));
    return PARADMM_OK;
}

// running sums [3][2] (FP64), partial sums [8192][3] (FP64), counts
// [8192][3], rho per type [3], raised count
static int ensure_policy_part(DevGraph* s)
{
    if (s->d_policy_part) return PARADMM_OK;
    RC(dmalloc(s, (void**)&s->d_policy_part, sizeof(double) * (6 + 3 * 8192) + sizeof(int) * 3 * 8192 + 16));
    CK(hipMemsetAsync(s->d_policy_part, 0, sizeof(double) * 6, s->stream));
    return PARADMM_OK;
}

static int apply_policy(DevGraph* s, float frac)
{
    paradmm_rho_mlp* m = (paradmm_rho_mlp*)s->policy_ctx;
    if (!m) return PARADMM_OK;
    if (s->policy_in_solve && s->layout == 2 && (s->policy_per_type == 2 || s->policy_per_type == 3) &&
        dense_contact_policy_ok(s)) {
        // inside the contact state: no per-edge pass over the whole graph
        RC(ensure_policy_part(s));
        RC(dense_contact_policy(s, mlp_device_params(m), frac, s->policy_per_type, s->policy_average,
                                (double*)s->d_policy_part));
        s->rho_device_written = true;
        s->rho_uni = false;
        s->policy_raised = 1;  // not counted here: rho is treated as per edge
        return PARADMM_OK;
    }
    RC(current_u(s));
    RC(ensure_slot_maps(s));
    PolicyArgs a = policy_args(s, frac);
    const long nchunks = (a.nslots + 63) / 64;
    // every workgroup stages the network's weights once, then loops over
    // chunks of 64 slots: about 4 workgroups per CU
    const unsigned grid = (unsigned)std::min<long>((nchunks + 3) / 4, 1024);
    prof_begin(s, "rho_policy");
    if (!s->policy_per_type) {
        k_rho_policy<<<grid, 256, 0, s->stream>>>(a, mlp_device_params(m));
    } else {
        RC(ensure_policy_part(s));
        double* hist = (double*)s->d_policy_part;
        double* part = hist + 6;
        int* cnt = (int*)(part + 3 * 8192);
        float* rho_type = (float*)(cnt + 3 * 8192);
        int* nraise = (int*)(rho_type + 3);
        const bool raises = s->policy_per_type == 3;
        if (raises) {
            if (!s->d_policy_y) RC(dmalloc(s, (void**)&s->d_policy_y, sizeof(float) * (size_t)a.nslots));
            a.yout = s->d_policy_y;
            a.nraise = nraise;
            CK(hipMemsetAsync(nraise, 0, sizeof(int), s->stream));
            int nt = 0, ts = 1;
            const int* tpos = nullptr;
            if (int2* tl = dense_tile_flags(s, &tpos, &nt, &ts)) {
                a.tflag = (int*)tl;
                a.tile_pos = tpos;
                a.tile_slots = ts;
                a.ntiles = nt;
            }
        }
        k_rho_policy_type_sums<<<grid, 256, 0, s->stream>>>(a, mlp_device_params(m), part, cnt);
        k_rho_policy_type_final<<<1, 256, 0, s->stream>>>((int)grid, part, cnt, rho_type, s->policy_per_type >= 2,
                                                         s->policy_average ? hist : nullptr);
        if (raises)
            k_rho_policy_raise_apply<<<(unsigned)((a.nslots + 255) / 256), 256, 0, s->stream>>>(a, rho_type,
                                                                                             RHO_RAISE_LOG10);
        else
            k_rho_policy_type_apply<<<(unsigned)((a.nslots + 255) / 256), 256, 0, s->stream>>>(a, rho_type);
    }
    prof_end(s);
    CK(hipGetLastError());
    s->rho_device_written = true;
    s->rho_uni = false;
    if (s->policy_per_type >= 2) {
        // one rho for every edge (no edge raised): the steady-state kernels
        // may use it as a scalar
        int nr = 0;
        CK(hipMemcpyAsync(&s->rho_u, rho_type_ptr(s), sizeof(float), hipMemcpyDeviceToHost, s->stream));
        if (s->policy_per_type == 3)
            CK(hipMemcpyAsync(&nr, rho_type_ptr(s) + 3, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        CK(hipStreamSynchronize(s->stream));
        s->policy_raised = nr;
        // the dense tile kernel takes the raised edges' tiles through the
        // per-edge streams (tile flags) and keeps the scalar rho elsewhere
        if (nr == 0 || a.tflag) {
            s->rho_uni = true;
            RC(sync_uniform(s));
        }
    }
    return PARADMM_OK;
}

}  // namespace pdm

extern "C" int paradmm_gpu_set_rho_policy(graph* G, paradmm_rho_mlp* m)
{
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    s->policy_ctx = m;
    s->rho_policy = m ? &apply_policy : nullptr;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_set_rho_policy_mode(graph* G, int mode, int average)
{
    if (mode < 0 || mode > 3) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_set_rho_policy_mode: mode 0, 1, 2 or 3");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    s->policy_per_type = mode;
    s->policy_average = average != 0;
    if (s->d_policy_part) CK(hipMemsetAsync(s->d_policy_part, 0, sizeof(double) * 6, s->stream));
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_apply_rho_policy(graph* G, paradmm_rho_mlp* m, float iter_frac)
{
    if (!m) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_apply_rho_policy: no network");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    void* keep = s->policy_ctx;
    s->policy_ctx = m;
    int rc = apply_policy(s, iter_frac);
    s->policy_ctx = keep;
    RC(rc);
    CK(hipStreamSynchronize(s->stream));
    s->dev_newer = true;
    s->m_valid = false;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_rho_features(graph* G, float iter_frac, float* feat)
{
    if (!feat) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_rho_features: NULL output");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    const size_t n = (size_t)s->E * MLP_NIN;
    float* d = nullptr;
    CK(hipMalloc(&d, sizeof(float) * n));
    PolicyArgs a = policy_args(s, iter_frac);
    a.feat_out = d;
    k_rho_features<<<(unsigned)((s->E + 255) / 256), 256, 0, s->stream>>>(a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(feat, d, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "paradmm_gpu_rho_features", __FILE__, __LINE__);
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_scale_rho(graph* G, double factor)
{
    if (!(factor > 0.0)) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_scale_rho: factor must be > 0");
        return PARADMM_ERR_ARG;
    }
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    RC(current_u(s));
    k_scale_rho<<<(unsigned)((s->stride + 255) / 256), 256, 0, s->stream>>>(s->stride, s->D, s->stride, (float)factor,
                                                                             s->d_rho, s->d_U);
    s->rho_u *= (float)factor;
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s->stream));
    RC(sync_uniform(s));
    s->rho_device_written = true;
    s->dev_newer = true;
    s->m_valid = false;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_state_save(graph* G)
{
    DevGraph* s;
    RC(pdm_get_state(G, &s, true));
    RC(current_u(s));
    const size_t se = (size_t)s->D * s->stride, sv = (size_t)s->D * s->V;
    const size_t total = 2 * se + s->stride + 2 * sv;
    if (!s->d_ckpt) RC(dmalloc(s, (void**)&s->d_ckpt, sizeof(float) * total));
    float* c = s->d_ckpt;
    CK(hipMemcpyAsync(c, s->d_X, sizeof(float) * se, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(c + se, s->d_U, sizeof(float) * se, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(c + 2 * se, s->d_rho, sizeof(float) * s->stride, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(c + 2 * se + s->stride, s->d_Z, sizeof(float) * sv, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(c + 2 * se + s->stride + sv, s->d_Zprev, sizeof(float) * sv, hipMemcpyDeviceToDevice, s->stream));
    CK(hipStreamSynchronize(s->stream));
    s->ckpt_n_from_buffer = s->n_from_buffer;
    s->ckpt_rho_uni = s->rho_uni;
    s->ckpt_rho_u = s->rho_u;
    s->ckpt_raised = s->policy_raised;
    return PARADMM_OK;
}

extern "C" int paradmm_gpu_state_restore(graph* G)
{
    DevGraph* s;
    RC(pdm_get_state(G, &s, false));
    if (!s->d_ckpt) {
        pdm_set_error(PARADMM_ERR_STATE, "no saved state");
        return PARADMM_ERR_STATE;
    }
    const size_t se = (size_t)s->D * s->stride, sv = (size_t)s->D * s->V;
    const float* c = s->d_ckpt;
    CK(hipMemcpyAsync(s->d_X, c, sizeof(float) * se, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(s->d_U, c + se, sizeof(float) * se, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(s->d_rho, c + 2 * se, sizeof(float) * s->stride, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(s->d_Z, c + 2 * se + s->stride, sizeof(float) * sv, hipMemcpyDeviceToDevice, s->stream));
    CK(hipMemcpyAsync(s->d_Zprev, c + 2 * se + s->stride + sv, sizeof(float) * sv, hipMemcpyDeviceToDevice, s->stream));
    CK(hipStreamSynchronize(s->stream));
    if (s->layout == 2) dense_reset(s);  // the saved U was current
    s->n_from_buffer = s->ckpt_n_from_buffer;
    // raised edges of the saved state may sit in other tiles than the
    // current flags say: the per-edge streams until the next application
    s->rho_uni = s->ckpt_rho_uni && s->ckpt_raised == 0;
    s->rho_u = s->ckpt_rho_u;
    s->policy_raised = s->ckpt_raised;
    RC(sync_uniform(s));
    s->host_newer = false;
    s->dev_newer = true;
    s->m_valid = false;
    s->rho_device_written = true;
    return PARADMM_OK;
}

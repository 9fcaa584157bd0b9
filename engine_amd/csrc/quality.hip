// quality.hip -- feasibility and objective of a unit-square packing read
// from Z (one (cx, cy, r) variable per circle), on the GPU:
//   max over pairs of (r_i + r_j - |c_i - c_j|)+,
//   max over circles of the container violation
//     max(r - cx, cx + r - 1, r - cy, cy + r - 1, -r)+,
//   sum of radii (the objective the radius factors maximise).
// These judge every solve (fixed or learned rho); the reference has none.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_internal.h"
#include "engine_state.h"
#include "paradmm_gpu.h"

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return pdm::hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define RC(x) do { int r_ = (x); if (r_) return r_; } while (0)

namespace pdm {

// one workgroup per circle i; lanes over j > i; FP64 from the fp32 Z
__global__ __launch_bounds__(256) void k_square_quality(int N, const float* __restrict__ Z, double* __restrict__ part)
{
    __shared__ double sm[256];
    const int i = blockIdx.x;
    const double cx = Z[3 * i], cy = Z[3 * i + 1], r = Z[3 * i + 2];
    double mo = 0.0;
    for (int j = i + 1 + threadIdx.x; j < N; j += 256) {
        const double dx = cx - (double)Z[3 * j], dy = cy - (double)Z[3 * j + 1];
        const double ov = r + (double)Z[3 * j + 2] - sqrt(dx * dx + dy * dy);
        mo = ov > mo ? ov : mo;
    }
    sm[threadIdx.x] = mo;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mb = 0.0;
        const double v[5] = {r - cx, cx + r - 1.0, r - cy, cy + r - 1.0, -r};
        for (int k = 0; k < 5; ++k) mb = v[k] > mb ? v[k] : mb;
        part[3 * i] = sm[0];
        part[3 * i + 1] = mb;
        part[3 * i + 2] = r;
    }
}

__global__ __launch_bounds__(256) void k_square_quality_final(int N, const double* __restrict__ part,
                                                              double* __restrict__ out)
{
    __shared__ double so[256], sb[256], ss[256];
    double o = 0.0, b = 0.0, s = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) {
        o = fmax(o, part[3 * i]);
        b = fmax(b, part[3 * i + 1]);
        s += part[3 * i + 2];
    }
    so[threadIdx.x] = o;
    sb[threadIdx.x] = b;
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            so[threadIdx.x] = fmax(so[threadIdx.x], so[threadIdx.x + k]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + k]);
            ss[threadIdx.x] += ss[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = so[0];
        out[1] = sb[0];
        out[2] = ss[0];
    }
}

int square_quality(DevGraph* s, double out[3])
{
    if (s->D != 3) {
        pdm_set_error(PARADMM_ERR_SHAPE, "packing quality needs num_dims 3 (cx, cy, r)");
        return PARADMM_ERR_SHAPE;
    }
    if (!s->d_quality) {
        RC(dmalloc(s, (void**)&s->d_quality, sizeof(double) * (3 * (size_t)s->V + 4)));
    }
    double* part = s->d_quality;
    double* res = s->d_quality + 3 * (size_t)s->V;
    k_square_quality<<<(unsigned)s->V, 256, 0, s->stream>>>(s->V, s->d_Z, part);
    CK(hipGetLastError());
    k_square_quality_final<<<1, 256, 0, s->stream>>>(s->V, part, res);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return PARADMM_OK;
}

}  // namespace pdm

// ---------------------------------------------------------------- no allocation
// paradmm_square_quality_noalloc keeps the per-circle partials of
// k_square_quality in this file-scope array instead of memory the library
// allocates: one copy per device, held until the process ends.  Host code
// never uses the name `scratch` as an address (on the host it is only a
// stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns, looked up once per device with that device
// current and kept in g_scratch.
static __device__ __align__(256) double scratch[3 * PARADMM_QUALITY_NOALLOC_MAX];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices];

static int scratch_of_current_device(double** out)
{
    int dev = -1;
    CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_square_quality_noalloc: device %d out of range", dev);
        return PARADMM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            pdm_set_error(PARADMM_ERR_HIP, "paradmm_square_quality_noalloc: no device address for the scratch array");
            return PARADMM_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return PARADMM_OK;
}

extern "C" int paradmm_square_quality_noalloc(const float* Z, int ncircles, double* out3, void* stream)
{
    if (!Z || !out3 || ncircles < 0) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_square_quality_noalloc: Z and out3 must be device pointers, "
                                       "ncircles >= 0");
        return PARADMM_ERR_ARG;
    }
    if (ncircles > PARADMM_QUALITY_NOALLOC_MAX) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_square_quality_noalloc: %d circles exceed the scratch limit of %d",
                      ncircles, PARADMM_QUALITY_NOALLOC_MAX);
        return PARADMM_ERR_ARG;
    }
    double* part = nullptr;
    RC(scratch_of_current_device(&part));
    hipStream_t st = static_cast<hipStream_t>(stream);
    // every partial the final pass reads ([0, 3 ncircles)) is written first:
    // nothing of an earlier call's contents is used
    if (ncircles > 0) {
        hipLaunchKernelGGL(pdm::k_square_quality, dim3((unsigned)ncircles), dim3(256), 0, st, ncircles, Z, part);
        CK(hipGetLastError());
    }
    hipLaunchKernelGGL(pdm::k_square_quality_final, dim3(1), dim3(256), 0, st, ncircles, part, out3);
    CK(hipGetLastError());
    return PARADMM_OK;
}

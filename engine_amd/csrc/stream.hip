// stream.hip -- a plain device-to-device copy kernel of known size: the
// practical HBM ceiling reported next to the nominal peak, and the
// calibration kernel for hardware byte counters (FETCH_SIZE / WRITE_SIZE).
#include <hip/hip_runtime.h>

#include "engine_internal.h"
#include "paradmm_gpu.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// four 16-byte loads in flight per thread, non-temporal (streamed once)
__global__ __launch_bounds__(256) void k_copy_stream(const f4* __restrict__ src, f4* __restrict__ dst, long n4)
{
    const long stride = (long)gridDim.x * 1024;
    for (long i = (long)blockIdx.x * 1024 + threadIdx.x; i < n4; i += stride) {
        f4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long j = i + 256 * u;
            if (j < n4) v[u] = __builtin_nontemporal_load(src + j);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long j = i + 256 * u;
            if (j < n4) __builtin_nontemporal_store(v[u], dst + j);
        }
    }
}

int fail(hipError_t e, const char* what)
{
    pdm_set_error(PARADMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return PARADMM_ERR_HIP;
}

}  // namespace

extern "C" int paradmm_gpu_copy_bandwidth(size_t bytes, int iters, double* gbs)
{
    if (bytes < 4096 || iters < 1 || !gbs) {
        pdm_set_error(PARADMM_ERR_ARG, "paradmm_gpu_copy_bandwidth: bad arguments");
        return PARADMM_ERR_ARG;
    }
    const long n4 = (long)(bytes / 16);
    f4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PARADMM_OK;
    hipError_t e;
    if ((e = hipMalloc(&a, n4 * 16)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b, n4 * 16)) != hipSuccess) { (void)hipFree(a); return fail(e, "hipMalloc"); }
    do {
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) { rc = fail(e, "stream"); break; }
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) { rc = fail(e, "event"); break; }
        if ((e = hipMemsetAsync(a, 0, n4 * 16, st)) != hipSuccess) { rc = fail(e, "memset"); break; }
        const unsigned grid = 256 * 8;  // 8 workgroups per CU, grid-stride
        k_copy_stream<<<grid, 256, 0, st>>>(a, b, n4);  // warm-up
        if ((e = hipEventRecord(e0, st)) != hipSuccess) { rc = fail(e, "record"); break; }
        for (int i = 0; i < iters; ++i) k_copy_stream<<<grid, 256, 0, st>>>(a, b, n4);
        if ((e = hipEventRecord(e1, st)) != hipSuccess) { rc = fail(e, "record"); break; }
        if ((e = hipEventSynchronize(e1)) != hipSuccess) { rc = fail(e, "sync"); break; }
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) { rc = fail(e, "elapsed"); break; }
        *gbs = 2.0 * (double)n4 * 16.0 * iters / (ms * 1e-3) / 1e9;  // read + write
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

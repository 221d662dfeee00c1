// A plain device-to-device copy of known size, for calibrating hardware byte counters
// (rocprofv3 FETCH_SIZE / WRITE_SIZE) against the kernels' own traffic.
#include "mx_internal.h"

namespace {

__global__ __launch_bounds__(256) void mx_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void mx_fill_kernel(uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = make_uint4((unsigned)i, 1u, 2u, 3u);
}

} // namespace

extern "C" int mx_copy_probe(mx_ctx* ctx, size_t bytes, double* kernel_ms) {
    if (!ctx || bytes < 16) return MX_E_INVALID;
    const size_t n = bytes / 16;
    if (hipSetDevice(ctx->device) != hipSuccess) return MX_E_HIP;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return MX_E_NOMEM;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return MX_E_NOMEM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int blocks = ctx->cu_count * 8;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mx_fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, a, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mx_copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (kernel_ms) *kernel_ms = ms;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) {
        ctx->set_err(std::string("copy probe: ") + hipGetErrorString(e));
        return MX_E_HIP;
    }
    return MX_OK;
}

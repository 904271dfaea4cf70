// calibrate.hip — throughput probe for the roofline of the limb kernels: every lane runs eight
// independent v_mad_u64_u32 chains (the 32x32->64 multiply-add the limb products are made of), so the
// measured rate is the device's sustained word-product issue rate.  No reference counterpart.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "limbforge_hip.h"

namespace {
__global__ void __launch_bounds__(256) mad64_probe(uint32_t* out, int iters, uint32_t seed)
{
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    uint32_t x[8];
    uint64_t acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        x[k] = gid * 8u + (uint32_t)k + seed;
        acc[k] = (uint64_t)k * 0x9E3779B97F4A7C15ull;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = (uint64_t)x[k] * (uint32_t)acc[k] + acc[k];
    }
    uint64_t r = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) r ^= acc[k];
    out[gid] = (uint32_t)r ^ (uint32_t)(r >> 32);
}
} // namespace

extern "C" int lf_calibrate_mad64(uint64_t blocks, int iters, double* products_per_second)
{
    if (!blocks || iters <= 0 || !products_per_second) return LF_INVALID_ARGUMENT;
    uint32_t* out = nullptr;
    if (hipMalloc(&out, blocks * 256 * 4) != hipSuccess) return LF_RUNTIME_ERROR;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(mad64_probe, dim3((unsigned)blocks), dim3(256), 0, nullptr, out, iters, 1u); // warm
    (void)hipEventRecord(e0, nullptr);
    hipLaunchKernelGGL(mad64_probe, dim3((unsigned)blocks), dim3(256), 0, nullptr, out, iters, 2u);
    (void)hipEventRecord(e1, nullptr);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return LF_RUNTIME_ERROR;
    *products_per_second = (double)blocks * 256.0 * 8.0 * (double)iters / (ms * 1e-3);
    return LF_OK;
}

// Reference copy for HBM traffic counters: `bytes` (a multiple of 16) copied once from one fresh buffer to
// another, 16 bytes per lane per iteration of a grid-stride loop.  A profiler's FETCH_SIZE / WRITE_SIZE for
// this kernel, against the known size, gives the scale of those counters.
namespace {
__global__ void __launch_bounds__(256) copy_probe(const uint4* src, uint4* dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) dst[i] = src[i];
}
} // namespace

extern "C" int lf_calibrate_copy(uint64_t bytes, double* bytes_per_second)
{
    if (!bytes || bytes % 16 || !bytes_per_second) return LF_INVALID_ARGUMENT;
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess) return LF_RUNTIME_ERROR;
    if (hipMalloc(&dst, bytes) != hipSuccess) {
        (void)hipFree(src);
        return LF_RUNTIME_ERROR;
    }
    (void)hipMemset(src, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint64_t n = bytes / 16;
    const unsigned blocks = 256u * 8u;
    (void)hipEventRecord(e0, nullptr);
    hipLaunchKernelGGL(copy_probe, dim3(blocks), dim3(256), 0, nullptr, (const uint4*)src, (uint4*)dst, n);
    (void)hipEventRecord(e1, nullptr);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (err != hipSuccess) return LF_RUNTIME_ERROR;
    *bytes_per_second = 2.0 * (double)bytes / (ms * 1e-3);
    return LF_OK;
}

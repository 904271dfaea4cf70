// layout.hip — conversion between the public array-of-structs element (engine.hpp:195: N limbs,
// exponent, sign, status = 4N+12 bytes; a complex value is two of them) and device limb planes.
// Each workgroup moves 64 values: the AoS block is contiguous, so it is read (or written) with fully
// coalesced dword accesses and transposed through LDS (row stride made odd, so the per-plane column
// reads are free of bank conflicts); every plane row of 64 values is one 256-byte access.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lf_internal.hpp"

namespace lf {

constexpr int TILE = 64;
constexpr int MAX_RECORD = 2 * (32 + 3); // words per complex 1024-bit value

__global__ void __launch_bounds__(256) aos_to_soa_kernel(const uint32_t* __restrict__ aos, uint32_t* __restrict__ soa,
                                                         uint64_t pitch, uint32_t count, int record)
{
    __shared__ uint32_t tile[TILE * (MAX_RECORD + 1)];
    const int stride = record | 1;
    const uint64_t first = (uint64_t)blockIdx.x * TILE;
    const int values = (int)min((uint64_t)TILE, (uint64_t)count - first);
    const int words = values * record;
    const uint32_t* src = aos + first * (uint64_t)record;
    for (int k = threadIdx.x; k < words; k += 256) {
        const int v = k / record, w = k - v * record;
        tile[v * stride + w] = src[k];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TILE * record; k += 256) {
        const int w = k / TILE, v = k - w * TILE;
        if (v < values) soa[(uint64_t)w * pitch + first + v] = tile[v * stride + w];
    }
}

__global__ void __launch_bounds__(256) soa_to_aos_kernel(const uint32_t* __restrict__ soa, uint32_t* __restrict__ aos,
                                                         uint64_t pitch, uint32_t count, int record)
{
    __shared__ uint32_t tile[TILE * (MAX_RECORD + 1)];
    const int stride = record | 1;
    const uint64_t first = (uint64_t)blockIdx.x * TILE;
    const int values = (int)min((uint64_t)TILE, (uint64_t)count - first);
    for (int k = threadIdx.x; k < TILE * record; k += 256) {
        const int w = k / TILE, v = k - w * TILE;
        if (v < values) tile[v * stride + w] = soa[(uint64_t)w * pitch + first + v];
    }
    __syncthreads();
    const int words = values * record;
    uint32_t* dst = aos + first * (uint64_t)record;
    for (int k = threadIdx.x; k < words; k += 256) {
        const int v = k / record, w = k - v * record;
        dst[k] = tile[v * stride + w];
    }
}

hipError_t launch_aos_to_soa(const uint32_t* aos, uint32_t* soa, uint64_t pitch, uint32_t count, int record, hipStream_t s)
{
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(lf::grid_blocks(count, TILE)), dim3(256), 0, s, aos, soa, pitch, count, record);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const uint32_t* soa, uint32_t* aos, uint64_t pitch, uint32_t count, int record, hipStream_t s)
{
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(lf::grid_blocks(count, TILE)), dim3(256), 0, s, soa, aos, pitch, count, record);
    return hipGetLastError();
}

} // namespace lf

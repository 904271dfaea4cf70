// lf_target.hpp — the compilation target of the arithmetic core (limb_core.hpp).
//
// The core is written once and built twice: as gfx950 device code inside the HIP kernels, and as plain
// C++17 for the host (LF_HOST_BUILD), where liblimbforge.so gives the public scalar functions of
// include/limbforge/core.hpp (add, mul, div, sqrt, fma, ... on the CPU) from the same source, so host and
// device results are the same bits by construction.  The host build defines the HIP function qualifiers
// away and gives the two AMDGPU builtins the core uses their documented semantics:
//   v_alignbit_b32(hi, lo, s) = low 32 bits of ({hi, lo} >> (s & 31)),   clz of a 32-bit word (32 for 0).
#pragma once
#ifdef LF_HOST_BUILD
#include <stdint.h>
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
namespace lf {
inline uint32_t host_alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
}
inline int host_clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
} // namespace lf
#define __builtin_amdgcn_alignbit(hi, lo, s) ::lf::host_alignbit((hi), (lo), (s))
#define __clz(x) ::lf::host_clz(x)
#else
#include <hip/hip_runtime.h>
#endif

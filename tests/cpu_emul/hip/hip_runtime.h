// Test-only stand-in for <hip/hip_runtime.h>: lets the device arithmetic core (limbforge_amd/csrc/
// limb_core.hpp) compile as plain C++ on the host, so tests/device_core_cpu.cpp can check its logic
// against the oracle without a GPU.  The two AMDGPU builtins it uses are given their documented
// semantics (v_alignbit_b32: ({hi,lo} >> (s & 31)) low 32 bits; clz of a 32-bit word).
#pragma once
#include <stdint.h>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
static inline uint32_t lf_emul_alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
}
static inline int lf_emul_clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
#define __builtin_amdgcn_alignbit(hi, lo, s) lf_emul_alignbit((hi), (lo), (s))
#define __clz(x) lf_emul_clz(x)

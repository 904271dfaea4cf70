// residue_reduce.hpp — the residue of one int8-MFMA output, (k11 G11 + kmid Gmid + k00 G00) mod m, in 32-bit
// integer and FP32 arithmetic only.
//
// The three accumulators are the products of the high and low int8 digits of the two operands' residues
// (|G11|, |G00| <= 2^30 and |Gmid| < 2^31 for K <= 65472 terms of at most 128 * 128).  The weights are
// 65536, 256 and 1 for the residue itself (residue_combine), or 65536 y, 256 y and y mod m for a constant y,
// taken in the balanced range |k| <= (m - 1) / 2 < 2^15 (residue_fold: the residue times y mod m, with y the
// CRT factor (M/m)^-1 mod m).  Either way |T| = |k11 G11 + kmid Gmid + k00 G00| < 2^47, so T is exact in 64-bit
// (three 32x32+64-bit signed multiply-adds).  With T = hi 2^32 + lo (hi signed, |hi| <= 2^15; lo unsigned)
// and c the balanced value of 2^32 mod m (|c| <= m/2 < 2^15), V = hi c + lo is congruent to T and lies in
// (-2^30, 2^32 + 2^30).  V is formed exactly modulo 2^32 in integers, and estimated in FP32 as
// fma(hi, c, (float)lo) (the product is exact, the conversion of lo errs by at most 2^7 and the one rounding
// by at most 2^8).  With 2^15 < m < 2^16 the estimate of V / m is off by less than 0.04 (those errors, then
// the reciprocal's and the product's relative 2^-24 each on |V / m| < 2^18), so q = rint(estimate) leaves
// V - q m, exact modulo 2^32 (q m through the 24-bit multiplier), in [-0.54 m, 0.54 m]; one correction puts
// it in [0, m).
#pragma once
#include "lf_target.hpp"
#include <stdint.h>

namespace lf {

#ifdef LF_HOST_BUILD
// v_mul_i32_i24: the low 32 bits of the product of the operands' low 24 bits, taken as signed
inline int residue_mul24(int a, int b)
{
    const int64_t x = (int64_t)(int32_t)((uint32_t)a << 8) >> 8, y = (int64_t)(int32_t)((uint32_t)b << 8) >> 8;
    return (int)(uint32_t)(uint64_t)(x * y);
}
#else
__device__ __forceinline__ int residue_mul24(int a, int b) { return __mul24(a, b); }
#endif

// T mod m in [0, m) for |T| < 2^47; inv = 1.0f / m, c = balanced 2^32 mod m
__device__ __forceinline__ uint32_t residue_of(int64_t T, int m, float inv, int c)
{
    const int hi = (int)(T >> 32);
    const uint32_t lo = (uint32_t)T;
    const int v = (int)(lo + (uint32_t)residue_mul24(hi, c));
    const float vf = __builtin_fmaf((float)hi, (float)c, (float)lo);
    int r = v - residue_mul24((int)__builtin_rintf(vf * inv), m);
    r += r < 0 ? m : 0;
    return (uint32_t)r;
}

// (65536 g11 + 256 gmid + g00) mod m in [0, m)
__device__ __forceinline__ uint32_t residue_combine(int g11, int gmid, int g00, int m, float inv, int c)
{
    return residue_of((int64_t)g11 * 65536 + ((int64_t)gmid * 256 + (int64_t)g00), m, inv, c);
}

// (k11 g11 + kmid gmid + k00 g00) mod m in [0, m), the weights balanced (|k| <= (m - 1) / 2)
__device__ __forceinline__ uint32_t residue_fold(int g11, int gmid, int g00, int k11, int kmid, int k00, int m, float inv, int c)
{
    return residue_of((int64_t)g11 * k11 + (int64_t)gmid * kmid + (int64_t)g00 * k00, m, inv, c);
}

} // namespace lf

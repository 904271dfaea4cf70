// dot2_core.hpp — RN(a*b + c*d + e) with one rounding, the arithmetic of the complex fused operations
// (complex_fma / complex_fms: each part is one rounding of its exact value, core.hpp:330-358).
//
// The three exact terms (two 2N-word products and the N-word addend) are sorted by scale and split into
// exponent clusters: a new cluster starts where the remaining terms lie more than 32N + 6 bits below the
// current cluster's lowest bit.  Each cluster is summed exactly in a two's-complement accumulator.  With S
// the first nonzero cluster sum and F everything after it, |F| lies far below S's last bit and rounding
// position, so RN(S + F) = RN(S + sign(F) * eps) and sign(F) is the sign of the next nonzero cluster sum.
// A cluster of the three terms spans at most 8N + 1 words, so the accumulators are runtime-indexed arrays
// of 8N + 4 words walked by rolled loops: the operations are off the hot path, and the rolled form keeps
// code size and compile time small at 1024 bits (the arrays live in private memory).
#pragma once
#include "limb_core.hpp"

namespace lf {

constexpr int dot2_words(int N) { return 8 * N + 4; }

struct Dot2Term {
    i64 scale;
    int sign;
    int which; // 0: a*b, 1: c*d, 2: the addend
};

// acc (W words, two's complement) += sign * x (xw words) * 2^shift
__device__ inline void dot2_accumulate(u32* acc, int W, const u32* x, int xw, i64 shift, int sign)
{
    const int w0 = int(shift / 32), s = int(shift % 32);
    u32 carry = sign > 0 ? 0u : 1u; // a - v = a + ~v + 1
#pragma unroll 1
    for (int i = w0; i < W; ++i) {
        const int k = i - w0;
        u32 v = k < xw ? x[k] << s : 0u;
        if (s && k >= 1 && k - 1 < xw) v |= x[k - 1] >> (32 - s);
        if (sign < 0) v = ~v;
        else if (k > xw && !carry) break;
        const u64 t = (u64)acc[i] + v + carry;
        acc[i] = (u32)t;
        carry = (u32)(t >> 32);
    }
}

// |acc| in place; returns the sign (1 or -1) of the two's-complement value, 0 when it is zero
__device__ inline int dot2_magnitude(u32* acc, int W)
{
    const bool neg = (acc[W - 1] >> 31) != 0;
    bool any = false;
    u32 carry = 1;
#pragma unroll 1
    for (int i = 0; i < W; ++i) {
        if (neg) {
            const u64 z = (u64)(u32)~acc[i] + carry;
            acc[i] = (u32)z;
            carry = (u32)(z >> 32);
        }
        any = any || acc[i] != 0;
    }
    return any ? (neg ? -1 : 1) : 0;
}

// RN(sign * (|x| + dir * eps) * 2^scale) to N words, x of W words, eps below every bit of x and below the
// rounding position (dir 0: no eps); the reference's range rule (core.hpp:81-99).
template <int N> __device__ inline Num<N> dot2_round(const u32* x, int W, i64 scale, int sign, int dir)
{
    int top = W - 1;
#pragma unroll 1
    while (top > 0 && !x[top]) --top;
    const i64 h = 32 * (i64)top + 31 - __clz(x[top]);
    i64 e = scale + h;
    if (e > EMAX || e < -(i64)EMAX - 1) return zero<N>(ST_OVERFLOW);
    const i64 shift = h - (32 * N - 1); // bit position of the result's last bit in x
    auto bits_at = [&](i64 p) -> u32 {   // 32 bits of x from bit p (bits outside x are zero)
        if (p <= -32) return 0u;
        if (p < 0) return x[0] << (-p);
        const int i = int(p / 32), s = int(p % 32);
        u32 v = i < W ? x[i] >> s : 0u;
        if (s && i + 1 < W) v |= x[i + 1] << (32 - s);
        return v;
    };
    Num<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.w[i] = bits_at(shift + 32 * i);
    bool round = false, sticky = false;
    if (shift > 0) {
        const i64 rb = shift - 1;
        round = ((x[rb / 32] >> (rb % 32)) & 1u) != 0;
        sticky = (rb % 32) && (x[rb / 32] & ((1u << (rb % 32)) - 1u));
#pragma unroll 1
        for (i64 i = 0; i < rb / 32 && !sticky; ++i) sticky = x[i] != 0;
    }
    // with eps: |x| + eps rounds up on any round bit; |x| - eps rounds up only above the midpoint
    const bool up = round && (dir > 0 || sticky || (dir == 0 && (r.w[0] & 1u)));
    if (up) {
        int i = 0;
#pragma unroll 1
        while (i < N && ++r.w[i] == 0) ++i;
        if (i == N) {
            r.w[N - 1] = 0x80000000u;
            ++e;
        }
    }
    if (e > EMAX || e < -(i64)EMAX) return zero<N>(ST_OVERFLOW);
    r.e = (int)e;
    r.s = sign;
    r.st = ST_OK;
    return r;
}

template <int N> __device__ inline Num<N> dot2_add(const Num<N>& a, const Num<N>& b, const Num<N>& c, const Num<N>& d, const Num<N>& e)
{
    const u32 st = a.st | b.st | c.st | d.st | e.st;
    if (st) return zero<N>(st);
    constexpr int W = dot2_words(N);
    u32 prod[2][2 * N];
    mantissa_product<N>(a.w, b.w, prod[0]);
    mantissa_product<N>(c.w, d.w, prod[1]);
    Dot2Term t[3];
    int count = 0;
    if (a.s && b.s) t[count++] = {(i64)a.e + b.e - 2 * (32 * N - 1), a.s * b.s, 0};
    if (c.s && d.s) t[count++] = {(i64)c.e + d.e - 2 * (32 * N - 1), c.s * d.s, 1};
    if (!count) return e; // no nonzero product: the addend itself
    if (e.s) t[count++] = {(i64)e.e - (32 * N - 1), e.s, 2};
    // sort by scale, largest first
#pragma unroll 1
    for (int i = 0; i < count; ++i)
#pragma unroll 1
        for (int j = i + 1; j < count; ++j)
            if (t[j].scale > t[i].scale) {
                const Dot2Term x = t[i];
                t[i] = t[j];
                t[j] = x;
            }
    constexpr i64 gap = 32 * N + 2 + 4, width = 64 * N; // ceil(log2(K + 1)) = 2 for K = 2; a term is below 2^(scale + 64N)
    u32 first[W], next[W];
    i64 first_scale = 0;
    int first_sign = 0, next_sign = 0;
    int first_words = 0;
#pragma unroll 1
    for (int i = 0; i < count && !next_sign;) {
        int j = i + 1;
        i64 low = t[i].scale;
#pragma unroll 1
        while (j < count && t[j].scale + width - 1 + gap >= low) {
            low = t[j].scale < low ? t[j].scale : low;
            ++j;
        }
        const int words = int((t[i].scale + width - low) / 32 + 3);
        u32* acc = first_sign ? next : first;
#pragma unroll 1
        for (int k = 0; k < words; ++k) acc[k] = 0;
#pragma unroll 1
        for (int q = i; q < j; ++q) {
            const u32* x = t[q].which == 2 ? e.w : prod[t[q].which];
            dot2_accumulate(acc, words, x, t[q].which == 2 ? N : 2 * N, t[q].scale - low, t[q].sign);
        }
        const int sg = dot2_magnitude(acc, words);
        if (sg) {
            if (!first_sign) {
                first_sign = sg;
                first_scale = low;
                first_words = words;
            } else
                next_sign = sg;
        }
        i = j;
    }
    if (!first_sign) return zero<N>(ST_OK);
    return dot2_round<N>(first, first_words, first_scale, first_sign, next_sign * first_sign);
}

// complex a*b + c: re = RN(a.re b.re + (-a.im) b.im + c.re), im = RN(a.re b.im + a.im b.re + c.im)
template <int N> __device__ inline Cpx<N> cfma(const Cpx<N>& a, const Cpx<N>& b, const Cpx<N>& c)
{
    return {dot2_add<N>(a.re, b.re, negate<N>(a.im), b.im, c.re), dot2_add<N>(a.re, b.im, a.im, b.re, c.im)};
}
template <int N> __device__ inline Cpx<N> cfms(const Cpx<N>& a, const Cpx<N>& b, const Cpx<N>& c)
{
    return cfma<N>(a, b, {negate<N>(c.re), negate<N>(c.im)});
}

} // namespace lf

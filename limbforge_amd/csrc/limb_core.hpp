// limb_core.hpp — register-resident multiprecision arithmetic for gfx950 (CDNA4), one value per lane.
//
// A value is sign * integer(w[0..N)) * 2^(e - (32N-1)) with the top bit of w[N-1] set when nonzero,
// exponent e in [-1e9, 1e9], sign in {-1, 0, 1} and a status bitmask {1 division by zero, 2 exponent
// overflow, 4 invalid} (the reference's number contract, include/limbforge/core.hpp:22-28).  Every
// operation returns the round-to-nearest-even result of the exact value (bit-identical to MPFR).
//
// Design for the wave64 VALU: all limb arrays are indexed at compile time only (every loop over words
// is unrolled), so limbs stay in VGPRs; variable shifts are barrel shifters (log2(W) select stages of
// v_cndmask plus one v_alignbit funnel pass) instead of runtime-indexed scratch; 32x32->64 products are
// v_mad_u64_u32 carry chains.  Loops whose trip count grows with N (division digits, square-root bits)
// stay rolled and move their windows as shift registers, which keeps code size within the instruction
// cache at 1024 bits.
#pragma once
#include "lf_target.hpp"
#include <stdint.h>

namespace lf {
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

constexpr int EMAX = 1000000000;
enum : u32 { ST_OK = 0, ST_DIV0 = 1, ST_OVERFLOW = 2, ST_INVALID = 4 };

template <int N> struct Num {
    u32 w[N];
    int e;
    int s;
    u32 st;
};

template <int N> __device__ __forceinline__ Num<N> zero(u32 st)
{
    Num<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.w[i] = 0;
    r.e = 0;
    r.s = 0;
    r.st = st;
    return r;
}

template <int N> __device__ __forceinline__ Num<N> negate(Num<N> a)
{
    a.s = -a.s;
    return a;
}

// ------------------------------------------------------------------ shifters
// x <<= 32*k words, k a runtime value in [0, W]; vacated words are zero.
template <int W> __device__ __forceinline__ void shl_words(u32 (&x)[W], int k)
{
#pragma unroll
    for (int s = 1; s < W; s <<= 1) {
        const bool on = (k & s) != 0;
#pragma unroll
        for (int i = W - 1; i >= 0; --i) x[i] = on ? (i >= s ? x[i - s] : 0u) : x[i];
    }
    if (k >= W) {
#pragma unroll
        for (int i = 0; i < W; ++i) x[i] = 0;
    }
}

// x <<= b bits, b a runtime value in [0, 31].
template <int W> __device__ __forceinline__ void shl_bits(u32 (&x)[W], int b)
{
    const u32 r = (u32)(32 - b) & 31u;
#pragma unroll
    for (int i = W - 1; i >= 0; --i) {
        u32 lo = i ? x[i - 1] : 0u;
        u32 v = __builtin_amdgcn_alignbit(x[i], lo, r);
        x[i] = b ? v : x[i];
    }
}

// x - y - bw (bw 0 or 1) in 32 bits, with the borrow out in bw.  Written with 32-bit compares, a chain of
// these becomes one v_sub_co / v_subb_co pair per word; the 64-bit form t = x - y - bw, bw = t >> 63
// costs about six instructions per word.
__device__ __forceinline__ u32 sub_borrow(u32 x, u32 y, u32& bw)
{
    const u32 d = x - y;
    const u32 r = d - bw;
    bw = (u32)(x < y) | (u32)(d < bw);
    return r;
}

// x + y + cy (cy 0 or 1) in 32 bits, with the carry out in cy: one v_add_co / v_addc_co pair per word in a
// chain, as sub_borrow is for differences.
__device__ __forceinline__ u32 add_carry(u32 x, u32 y, u32& cy)
{
    const u32 t = x + y;
    const u32 r = t + cy;
    cy = (u32)(t < x) | (u32)(r < t);
    return r;
}

// x >>= d bits for a runtime d < 32 (KMAX + 1) (the caller bounds d); returns whether a set bit was
// shifted out.  One select stage per bit of the word count d >> 5 up to KMAX; each stage ORs the words it
// drops into the sticky flag (about W/2 instructions over all stages, where a select tree over prefix ORs
// of the words costs one select per word per stage).  A right funnel shift by 0 bits is the identity, so
// the bit shift needs no select.
template <int W, int KMAX> __device__ __forceinline__ bool shr_sticky(u32 (&x)[W], u32 d)
{
    const u32 k = d >> 5, b = d & 31u;
    u32 lost = 0;
#pragma unroll
    for (int s = 1; s <= KMAX; s <<= 1) {
        const bool on = (k & (u32)s) != 0;
        u32 o = 0;
#pragma unroll
        for (int i = 0; i < s && i < W; ++i) o |= x[i];
        lost |= on ? o : 0u;
#pragma unroll
        for (int i = 0; i < W; ++i) x[i] = on ? (i + s < W ? x[i + s] : 0u) : x[i];
    }
    lost |= b ? x[0] << (32u - b) : 0u;
#pragma unroll
    for (int i = 0; i < W; ++i) x[i] = __builtin_amdgcn_alignbit(i + 1 < W ? x[i + 1] : 0u, x[i], b);
    return lost != 0;
}

// dst (W words) = src (S words) * 2^off for a runtime bit offset; bits that fall below bit 0 of dst
// are dropped and reported in the return value (nonzero iff any was set); bits above dst are dropped
// (callers size dst so that none are).
template <int W, int S> __device__ __forceinline__ bool place(u32 (&dst)[W], const u32 (&src)[S], i64 off)
{
    // Work in a (W+S)-word frame whose bit 32S is dst bit 0.
    constexpr int F = W + S;
    u32 t[F];
#pragma unroll
    for (int i = 0; i < F; ++i) t[i] = i < S ? src[i] : 0u;
    bool sticky = false;
    i64 f = off + 32 * S; // position of src bit 0 in the frame
    if (f < 0) {
        // every bit of src lies below dst
#pragma unroll
        for (int i = 0; i < S; ++i) sticky |= src[i] != 0;
#pragma unroll
        for (int i = 0; i < W; ++i) dst[i] = 0;
        return sticky;
    }
    int fw = (int)(f >> 5), fb = (int)(f & 31);
    if (f >= 32 * (i64)F) fw = F; // everything above the frame (callers avoid this)
    shl_bits(t, fb);
    // bit-shift spill of the top word of src is lost only when S words + shift exceed the frame,
    // which cannot happen for fw < F because t has W >= 1 spare words above src.
    shl_words(t, fw);
#pragma unroll
    for (int i = 0; i < S; ++i) sticky |= t[i] != 0;
#pragma unroll
    for (int i = 0; i < W; ++i) dst[i] = t[S + i];
    return sticky;
}

// Shift x left until bit 32W-1 is set; returns the number of leading zero bits, or -1 for zero.
template <int W> __device__ __forceinline__ int normalize(u32 (&x)[W])
{
    int lzw = 0;
    bool seen = false;
#pragma unroll
    for (int i = W - 1; i >= 0; --i) {
        seen = seen || x[i] != 0;
        lzw += seen ? 0 : 1;
    }
    if (lzw == W) return -1;
    // the word shifter (log2(W) select stages over W words) runs only on a branch: a value whose top word
    // is set, the usual case after an addition or an fma window, needs the bit shift alone
    if (lzw) shl_words(x, lzw);
    int b = __clz(x[W - 1]);
    shl_bits(x, b);
    return 32 * lzw + b;
}

// ------------------------------------------------------------------ rounding
// m += inc (0 or 1) for an N-word mantissa; returns the carry out of the top word (every word is then
// zero).  The increment passes word 0 only when that word is all ones, so the carry chain through the
// other words runs on a branch that lanes almost never take (done unconditionally, its N dependent
// carries cost about 2N instructions on every rounding).
template <int N> __device__ __forceinline__ u32 increment_mantissa(u32 (&m)[N], bool inc)
{
    const u32 w0 = m[0] + (inc ? 1u : 0u);
    u32 c = 0;
    if (inc && w0 == 0) {
        c = 1;
#pragma unroll
        for (int i = 1; i < N; ++i) {
            u64 t = (u64)m[i] + c;
            m[i] = (u32)t;
            c = (u32)(t >> 32);
        }
    }
    m[0] = w0;
    return c;
}

// The fields of a result whose rounded mantissa is in r.w and whose top bit has exponent msb (after
// rounding): zero with status st when kill (a zero or status operand), zero with st | overflow when msb
// lies outside [-1e9, 1e9] (the reference's range rule, core.hpp:81-99: a msb below -1e9-1 before
// rounding cannot be rescued by the carry, so this one check after rounding covers its two).  No early
// returns: a lane whose result is zero clears its words in place.  A return of zero<N>() on an early
// path makes the compiler materialize a whole zero result in other registers and copy the rounded one
// over it on every call.
template <int N> __device__ __forceinline__ void finish_rounded(Num<N>& r, i64 msb, int sign, u32 st, bool kill)
{
    const bool ovf = !kill && (msb > EMAX || msb < -(i64)EMAX);
    r.e = (int)msb;
    r.s = sign;
    r.st = ovf ? st | ST_OVERFLOW : st;
    if (kill || ovf) {
#pragma unroll
        for (int i = 0; i < N; ++i) r.w[i] = 0;
        r.e = 0;
        r.s = 0;
    }
}

// x (W words) is normalized (bit 32W-1 set) and represents sign * x * 2^(msb - (32W-1)).
// Round to nearest even to N words with the reference's range rule (core.hpp:81-99): the msb may be
// emin-1 before rounding (a carry can rescue it) and must lie in [-1e9, 1e9] after.  With kill the
// result is zero with status st.
template <int N, int W>
__device__ __forceinline__ Num<N> round_normalized(const u32 (&x)[W], i64 msb, int sign, u32 st, bool extra_sticky = false,
                                                   bool kill = false)
{
    Num<N> r;
    if constexpr (W <= N) {
#pragma unroll
        for (int i = 0; i < N; ++i) r.w[i] = i >= N - W ? x[i - (N - W)] : 0u;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) r.w[i] = x[W - N + i];
        const u32 g = x[W - N - 1];
        bool sticky = extra_sticky || (g << 1) != 0;
#pragma unroll
        for (int i = 0; i < W - N - 1; ++i) sticky = sticky || x[i] != 0;
        const bool inc = (g >> 31) && (sticky || (r.w[0] & 1u));
        const u32 c = increment_mantissa<N>(r.w, inc);
        r.w[N - 1] |= c << 31;
        msb += c;
    }
    finish_rounded<N>(r, msb, sign, st, kill);
    return r;
}

// Round the N-word mantissa m (top bit set) whose discarded tail has leading bit `round` and any
// other set bit `sticky` to nearest even; msb is the exponent of m's top bit (same range rule).
template <int N>
__device__ __forceinline__ Num<N> round_words(const u32 (&m)[N], u32 round, bool sticky, i64 msb, int sign, u32 st)
{
    Num<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.w[i] = m[i];
    const bool inc = round && (sticky || (m[0] & 1u));
    const u32 c = increment_mantissa<N>(r.w, inc);
    r.w[N - 1] |= c << 31;
    msb += c;
    finish_rounded<N>(r, msb, sign, st, false);
    return r;
}

// RN-even of sign * int(x) * 2^scale (the reference's pack for any word count, core.hpp:81-99).
// Only the words a rounding reads are shifted into place: after the (rarely needed) word shift, the N
// mantissa words and the guard word are funnel-shifted by the top word's leading zeros, and the sticky
// bit is the OR of the words below, which a shift does not change (x is clobbered).
template <int N, int W> __device__ __forceinline__ Num<N> pack(u32 (&x)[W], i64 scale, int sign, u32 st)
{
    if constexpr (W <= N + 1) {
        const int lz = normalize(x);
        return round_normalized<N, W>(x, scale + (32 * W - 1 - lz), sign, st, false, lz < 0);
    } else {
        // the count of leading zero words and the word shifter run on a branch: the top word is usually set
        int lzw = 0;
        bool kill = false;
        if (x[W - 1] == 0) {
            bool seen = false;
#pragma unroll
            for (int i = W - 1; i >= 0; --i) {
                seen = seen || x[i] != 0;
                lzw += seen ? 0 : 1;
            }
            kill = lzw == W;
            if (!kill) shl_words(x, lzw);
        }
        const int b = __clz(x[W - 1]) & 31;  // 0 when x is zero (killed)
        const u32 r = (u32)(32 - b) & 31u;
        u32 m[N + 1]; // m[0] = the guard word, m[1..N] the mantissa words of x << b
#pragma unroll
        for (int i = 0; i <= N; ++i) {
            const int k = W - 1 - N + i;
            m[i] = b ? __builtin_amdgcn_alignbit(x[k], x[k - 1], r) : x[k];
        }
        // bits of x << b below the guard word: all of x[0 .. W-N-3] and x[W-N-2] << b
        u32 low = x[W - N - 2] << b;
#pragma unroll
        for (int i = 0; i < W - N - 2; ++i) low |= x[i];
        const i64 msb = scale + (32 * W - 1 - (32 * lzw + b));
        return round_normalized<N, N + 1>(m, msb, sign, st, low != 0, kill);
    }
}

// ------------------------------------------------------------- comparisons
template <int N> __device__ __forceinline__ int compare_words(const u32 (&a)[N], const u32 (&b)[N])
{
    int cmp = 0;
#pragma unroll
    for (int i = N - 1; i >= 0; --i)
        if (cmp == 0 && a[i] != b[i]) cmp = a[i] > b[i] ? 1 : -1;
    return cmp;
}

// |a| vs |b| for values without status (core.hpp:100-105)
template <int N> __device__ __forceinline__ int magnitude_compare(const Num<N>& a, const Num<N>& b)
{
    if (!a.s) return b.s ? -1 : 0;
    if (!b.s) return 1;
    if (a.e != b.e) return a.e > b.e ? 1 : -1;
    return compare_words<N>(a.w, b.w);
}

template <int N> __device__ __forceinline__ void swap_if(bool c, Num<N>& a, Num<N>& b)
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u32 x = a.w[i], y = b.w[i];
        a.w[i] = c ? y : x;
        b.w[i] = c ? x : y;
    }
    int ae = a.e, be = b.e, as = a.s, bs = b.s;
    a.e = c ? be : ae;
    b.e = c ? ae : be;
    a.s = c ? bs : as;
    b.s = c ? as : bs;
}

// -------------------------------------------------------------- add / sub
// Exact aligned sum in 2N+2 words, one rounding (contract of core.hpp:116-134).
template <int N> __device__ __forceinline__ Num<N> add(Num<N> a, Num<N> b)
{
    const u32 st = a.st | b.st;
    if (st) return zero<N>(st);
    if (!a.s) return b;
    if (!b.s) return a;
    swap_if<N>(magnitude_compare<N>(a, b) < 0, a, b);
    const i64 gap = (i64)a.e - b.e;
    if (gap > 32 * N + 2) return a; // |b| below a quarter ulp of the representable a
    constexpr int W = 2 * N + 2;
    u32 x[W], y[W];
#pragma unroll
    for (int i = 0; i < W; ++i) x[i] = 0;
    // a at bit offset 32N+2 (compile time): words N.., shifted by 2 bits
#pragma unroll
    for (int i = 0; i < N; ++i) {
        x[N + i] |= a.w[i] << 2;
        x[N + i + 1] |= a.w[i] >> 30;
    }
    place<W, N>(y, b.w, 32 * N + 2 - gap); // exact: offset >= 0
    if (a.s == b.s) {
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            u64 t = (u64)x[i] + y[i] + c;
            x[i] = (u32)t;
            c = (u32)(t >> 32);
        }
    } else {
        u32 bw = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            x[i] = sub_borrow(x[i], y[i], bw);
        }
    }
    return pack<N, W>(x, (i64)a.e - (32 * N - 1) - (32 * N + 2), a.s, ST_OK);
}

template <int N> __device__ __forceinline__ Num<N> sub(Num<N> a, Num<N> b) { return add<N>(a, negate<N>(b)); }

// -------------------------------------------------------------------- mul
// From 8 limbs on, products are formed from 29-bit digits: a column of D <= 36 digit products sums to
// less than 36 * 2^58 < 2^64, so every digit product is one v_mad_u64_u32 accumulating into its column
// in place, with no carry chain and no register shuffling; one carry pass and a repack to 32-bit words
// follow.  (32-bit schoolbook rows need carry adds and pair moves around every multiply-add: about 4.7
// VALU instructions per word product at 32 limbs against 1.3 here.)
constexpr int DIGIT_BITS = 29;
constexpr u32 DIGIT_MASK = (1u << DIGIT_BITS) - 1;
template <int N> constexpr int digit_count() { return (32 * N + DIGIT_BITS - 1) / DIGIT_BITS; }
template <int N> constexpr bool use_digits() { return N >= 8; }

template <int N> __device__ __forceinline__ void to_digits(const u32 (&w)[N], u32 (&d)[digit_count<N>()])
{
#pragma unroll
    for (int k = 0; k < digit_count<N>(); ++k) {
        const int bit = DIGIT_BITS * k, wi = bit / 32, s = bit % 32;
        const u32 lo = w[wi] >> s;
        const u32 hi = (s && wi + 1 < N) ? (w[wi + 1] << (32 - s)) : 0u;
        d[k] = (lo | hi) & DIGIT_MASK;
    }
}

// Short products (the truncated "high half" product of MPFR's mpfr_mulhigh_n, after Krandick & Johnson,
// "Efficient multiprecision floating point multiplication with optimal directional rounding", 1993):
// the digit columns below C0 = short_cut<N>() are left out.  Column k sums at most k+1 products below
// 2^58, so the omitted part is delta < C0 * 2^(29*C0 + 29) / (1 - 2^-29), and C0 is the largest count
// with C0 < 2^(32N - 61 - 29*C0), which keeps delta < 2^(32N-32), the weight of word N-1 of the product.
template <int N> constexpr int short_cut()
{
    int best = 0;
    for (int c = 1; c <= digit_count<N>(); ++c) {
        const int e = 32 * N - 61 - DIGIT_BITS * c;
        if (e >= 0 && (e >= 30 || c < (1 << e))) best = c;
    }
    return best;
}

// p (2N words) = a * b from the digits of a and b (each < 2^(32N)) when full; otherwise the short
// product P' = a * b - delta, 0 <= delta < 2^(32N-32), from the columns from short_cut<N>() up.  With
// HIGH, only what a rounding to N words reads is formed: words N-1 .. 2N-1 exactly, word 0 nonzero
// exactly when a bit below word N-1 is set, words 1 .. N-2 zero (no repacking of the low half).
// Carry pass and repacking of digit column sums: col[k] (k < 2D) holds column k of a product of 29-bit
// digits, columns below short_cut<N>() only when full (zero otherwise); p as digit_columns describes.
template <int N, bool HIGH = false> __device__ __forceinline__ void pack_columns(u64 (&col)[2 * digit_count<N>()], u32 (&p)[2 * N], bool full)
{
    constexpr int D = digit_count<N>(), C0 = short_cut<N>();
    u32 dg[2 * D];
    if (full) {
#pragma unroll
        for (int k = 0; k < C0; ++k) {
            dg[k] = (u32)col[k] & DIGIT_MASK;
            col[k + 1] += col[k] >> DIGIT_BITS;
        }
    } else {
#pragma unroll
        for (int k = 0; k < C0; ++k) dg[k] = 0;
    }
#pragma unroll
    for (int k = C0; k < 2 * D - 1; ++k) {
        dg[k] = (u32)col[k] & DIGIT_MASK;
        col[k + 1] += col[k] >> DIGIT_BITS;
    }
    dg[2 * D - 1] = (u32)col[2 * D - 1]; // the product is < 2^(64N) <= 2^(58D): the top digit fits in 29 bits
    if constexpr (HIGH) {
        // bits below word N-1: digits 0 .. KF-1 whole, the low LB bits of digit KF
        constexpr int LOWBITS = 32 * (N - 1), KF = LOWBITS / DIGIT_BITS, LB = LOWBITS - DIGIT_BITS * KF;
        u32 any = LB ? (dg[KF] & ((1u << LB) - 1u)) : 0u;
#pragma unroll
        for (int k = 0; k < KF; ++k) any |= dg[k];
#pragma unroll
        for (int w = 1; w < N - 1; ++w) p[w] = 0;
        p[0] = any;
    }
#pragma unroll
    for (int w = HIGH ? N - 1 : 0; w < 2 * N; ++w) {
        const int bit = 32 * w, k = bit / DIGIT_BITS, s = bit % DIGIT_BITS;
        u32 v = dg[k] >> s;
        if (k + 1 < 2 * D) v |= dg[k + 1] << (DIGIT_BITS - s);
        if (k + 2 < 2 * D && 2 * DIGIT_BITS - s < 32) v |= dg[k + 2] << (2 * DIGIT_BITS - s);
        p[w] = v;
    }
}

// p (2N words) = a * b from the digits of a and b (each < 2^(32N)) when full; otherwise the short
// product P' = a * b - delta, 0 <= delta < 2^(32N-32), from the columns from short_cut<N>() up.  With
// HIGH, only what a rounding to N words reads is formed: words N-1 .. 2N-1 exactly, word 0 nonzero
// exactly when a bit below word N-1 is set, words 1 .. N-2 zero (no repacking of the low half).
template <int N, bool HIGH = false>
__device__ __forceinline__ void digit_columns(const u32 (&da)[digit_count<N>()], const u32 (&db)[digit_count<N>()], u32 (&p)[2 * N],
                                              bool full)
{
    constexpr int D = digit_count<N>(), C0 = short_cut<N>();
    static_assert(D <= 63, "column sums of 29-bit digit products must fit in 64 bits");
    u64 col[2 * D];
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) col[k] = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            if (i + j >= C0) col[i + j] += (u64)da[i] * db[j];
    if (full) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j)
                if (i + j < C0) col[i + j] += (u64)da[i] * db[j];
    }
    pack_columns<N, HIGH>(col, p, full);
}

// p (2N words) = a * b from the digits of a and b (each < 2^(32N)).
template <int N>
__device__ __forceinline__ void digit_product(const u32 (&da)[digit_count<N>()], const u32 (&db)[digit_count<N>()], u32 (&p)[2 * N])
{
    digit_columns<N>(da, db, p, true);
}

// Whether round_product of a short product P' of normalized operands is RN(a*b).  With A = P' div
// 2^(32N-32) and B = P' mod 2^(32N-32), a*b = A * 2^(32N-32) + (B + delta), B + delta < 2^(32N-31).
// The round bit r is bit 32N-1 (top bit of P set) or 32N-2, so the low r-(32N-32) bits of A (31 or 30)
// lie between it and the omitted part: unless they are all ones, no carry reaches r and mantissa and
// round bit are those of P'; the sticky bit is too, unless those bits and B are all zero.
template <int N> __device__ __forceinline__ bool short_decides(const u32 (&p)[2 * N])
{
    const u32 low = (p[2 * N - 1] >> 31) ? 0x7fffffffu : 0x3fffffffu;
    const u32 gl = p[N - 1] & low;
    u32 any = gl;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) any |= p[i];
    return gl != low && any != 0;
}

// p = the product a * b as far as round_product needs it: the short product when it decides the
// rounding, else (or with full set) the exact product.  With discard (the caller throws p away: a zero
// or status operand) the short product stands, so such lanes never take the exact path.
// With HIGH, p holds only what round_product reads (see digit_columns).
template <int N, bool HIGH = false>
__device__ __forceinline__ void rounding_product(const u32 (&da)[digit_count<N>()], const u32 (&db)[digit_count<N>()], u32 (&p)[2 * N],
                                                 bool full, bool discard = false)
{
    // two calls rather than a retry loop: in a loop the compiler hoists the low columns out of the
    // loop as invariant and forms them on every call
    digit_columns<N, HIGH>(da, db, p, full);
    if (!full && !discard && !short_decides<N>(p)) digit_columns<N, HIGH>(da, db, p, true);
}

// p (2N words) = a * b: digit columns from 8 limbs on, schoolbook rows of v_mad_u64_u32 carry chains below.
template <int N> __device__ __forceinline__ void mantissa_product(const u32 (&a)[N], const u32 (&b)[N], u32 (&p)[2 * N])
{
    if constexpr (use_digits<N>()) {
        u32 da[digit_count<N>()], db[digit_count<N>()];
        to_digits<N>(a, da);
        to_digits<N>(b, db);
        digit_product<N>(da, db, p);
        return;
    }
#pragma unroll
    for (int i = 0; i < 2 * N; ++i) p[i] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u32 carry = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            u64 v = (u64)a[i] * b[j] + p[i + j] + carry;
            p[i + j] = (u32)v;
            carry = (u32)(v >> 32);
        }
        p[i + N] = carry;
    }
}

// Round a normalized-input product: the msb of p is bit 64N-1 or 64N-2 (inputs have their top bits set).
// With kill (a zero or status operand) the result is zero with status st whatever p holds.
// With nudge = +1 or -1 the value rounded is P + nudge * delta for some 0 < delta < half a unit of P's
// last bit: a positive nudge sets the sticky bit (a tail of exactly one half rounds up), a negative one
// puts a tail of exactly one half below the midpoint (it rounds down); any other tail rounds as P's.
template <int N>
__device__ __forceinline__ Num<N> round_product(const u32 (&p)[2 * N], i64 ea_plus_eb, int sign, u32 st, bool kill = false,
                                                int nudge = 0)
{
    const bool top = (p[2 * N - 1] >> 31) != 0;
    i64 msb = ea_plus_eb + (top ? 1 : 0);
    Num<N> r;
    // mantissa = p[N..2N) shifted left by (top ? 0 : 1)
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u32 hi = p[N + i], lo = p[N + i - 1];
        r.w[i] = top ? hi : __builtin_amdgcn_alignbit(hi, lo, 31u);
    }
    const u32 g = p[N - 1];
    const u32 round = top ? (g >> 31) : ((g >> 30) & 1u);
    bool sticky = (top ? (g << 1) : (g << 2)) != 0;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) sticky = sticky || p[i] != 0;
    const bool inc = round && (sticky || nudge > 0 || (nudge == 0 && (r.w[0] & 1u)));
    const u32 c = increment_mantissa<N>(r.w, inc);
    r.w[N - 1] |= c << 31; // a carry out leaves all words zero: the mantissa becomes 2^(32N-1)
    msb += c;
    finish_rounded<N>(r, msb, sign, st, kill);
    return r;
}

// RN(a * b) (contract of core.hpp:137-145)
template <int N> __device__ __forceinline__ Num<N> mul(const Num<N>& a, const Num<N>& b)
{
    const u32 st = a.st | b.st;
    const bool kill = st || !a.s || !b.s;
    u32 p[2 * N];
    if constexpr (use_digits<N>()) {
        u32 da[digit_count<N>()], db[digit_count<N>()];
        to_digits<N>(a.w, da);
        to_digits<N>(b.w, db);
        rounding_product<N, true>(da, db, p, false, kill);
    } else
        mantissa_product<N>(a.w, b.w, p);
    return round_product<N>(p, (i64)a.e + b.e, a.s * b.s, st, kill);
}

// RN(a * b) with the digits of b's mantissa given (a multiplier reused over a chain); needs use_digits<N>().
template <int N> __device__ __forceinline__ Num<N> mul_by_digits(const Num<N>& a, const Num<N>& b, const u32 (&bd)[digit_count<N>()])
{
    const u32 st = a.st | b.st;
    const bool kill = st || !a.s || !b.s;
    u32 ad[digit_count<N>()], p[2 * N];
    to_digits<N>(a.w, ad);
    rounding_product<N, true>(ad, bd, p, false, kill);
    return round_product<N>(p, (i64)a.e + b.e, a.s * b.s, st, kill);
}

// -------------------------------------------------------------------- div
// Two-word by one-word quotient with a reused reciprocal: the 2/1 step of Moller & Granlund,
// "Improved division by invariant integers" (2011), as in core.hpp:176-184.
// Preconditions: d >= 2^31, high < d, inv = floor((2^64-1)/d) - 2^32.
__device__ __forceinline__ void divide_word(u32 high, u32 low, u32 d, u32 inv, u32& q, u32& r)
{
    u64 product = (u64)high * inv;
    u64 sum = (u64)(u32)product + low;
    u32 qq = (u32)((product >> 32) + high + 1 + (sum >> 32));
    u32 rr = low - qq * d;
    if (rr > (u32)sum) {
        --qq;
        rr += d;
    }
    if (rr >= d) {
        ++qq;
        rr -= d;
    }
    q = qq;
    r = rr;
}

// One quotient word of Knuth's algorithm D (TAOCP vol. 2, 4.3.1): the partial remainder R (N+1 words,
// R < B before the step) takes the next dividend word, the quotient word is estimated from the top two
// remainder words and B's top word (with the reused reciprocal inv of divide_word), corrected with B's
// second word, multiplied out and subtracted (added back once when the estimate was still one too high).
template <int N> __device__ __forceinline__ u32 knuth_step(u32 (&R)[N + 1], const u32 (&b)[N], u32 next, u32 inv)
{
    const u32 d = b[N - 1], dn = b[N - 2];
#pragma unroll
    for (int i = N; i > 0; --i) R[i] = R[i - 1];
    R[0] = next;
    u64 q, rem;
    if (R[N] >= d) {
        q = 0xffffffffull;
        rem = ((((u64)R[N]) << 32) | R[N - 1]) - q * d;
    } else {
        u32 qq, rr;
        divide_word(R[N], R[N - 1], d, inv, qq, rr);
        q = qq;
        rem = rr;
    }
    while (rem < ((u64)1 << 32) && q * dn > ((rem << 32) | R[N - 2])) {
        --q;
        rem += d;
    }
    // R -= q B: the product words from a multiply-add chain (each word's high half carried in the 64-bit
    // accumulator), subtracted by a borrow chain (one v_mad_u64_u32 and one subtract pair per word)
    u64 p = 0;
    u32 bw = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        p = (u64)(u32)q * b[i] + (p >> 32);
        R[i] = sub_borrow(R[i], (u32)p, bw);
    }
    R[N] = sub_borrow(R[N], (u32)(p >> 32), bw);
    const bool under = bw != 0;
    if (under) {
        --q;
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            u64 t = (u64)R[i] + b[i] + c;
            R[i] = (u32)t;
            c = (u32)(t >> 32);
        }
        R[N] += c;
    }
    return (u32)q;
}

// The rounding of a quotient: Q (N words, top bit set) with remainder R < B (R[N] = 0); 2R against B
// decides, ties to even; e is the exponent of Q's top bit.
template <int N>
__device__ __forceinline__ Num<N> round_quotient(const u32 (&Q)[N], const u32 (&R)[N + 1], const u32 (&b)[N], i64 e, int sign)
{
    int half;
    if (R[N - 1] >> 31) half = 1;
    else {
        u32 tw[N];
#pragma unroll
        for (int i = 0; i < N; ++i) tw[i] = __builtin_amdgcn_alignbit(R[i], i ? R[i - 1] : 0u, 31u);
        half = compare_words<N>(tw, b);
    }
    const bool inc = half > 0 || (half == 0 && (Q[0] & 1u));
    Num<N> r;
    r.s = sign;
    r.st = ST_OK;
    u32 c = inc ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u64 t = (u64)Q[i] + c;
        r.w[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    if (c) {
        r.w[N - 1] = 0x80000000u;
        e += 1;
    }
    if (e > EMAX || e < -(i64)EMAX) return zero<N>(ST_OVERFLOW);
    r.e = (int)e;
    return r;
}

// RN(a / b): normalized base-2^32 long division (Knuth D), then the remainder decides the rounding
// (contract of core.hpp:185-215).  The partial remainder R is an (N+1)-word shift register: each step
// brings in the next dividend word, so no array is indexed at run time.
template <int N> __device__ __forceinline__ Num<N> div(const Num<N>& a, const Num<N>& b)
{
    const u32 st = a.st | b.st;
    if (st) return zero<N>(st);
    if (!b.s) return zero<N>(ST_DIV0);
    if (!a.s) return zero<N>(ST_OK);
    const bool lt = compare_words<N>(a.w, b.w) < 0;
    i64 e = (i64)a.e - b.e - (lt ? 1 : 0);
    if (e > EMAX || e < -(i64)EMAX - 1) return zero<N>(ST_OVERFLOW);
    // dividend U = A * 2^(32N-1+lt): its top N words are A (lt) or A >> 1, then one word A[0] << 31
    // (only when !lt), then zeros.
    u32 R[N + 1], Q[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u32 hi = i + 1 < N ? a.w[i + 1] : 0u;
        R[i] = lt ? a.w[i] : __builtin_amdgcn_alignbit(hi, a.w[i], 1u);
        Q[i] = 0;
    }
    R[N] = 0;
    u32 next = lt ? 0u : (a.w[0] << 31);
    const u32 d = b.w[N - 1];
    const u32 inv = (u32)(~(u64)0 / d - ((u64)1 << 32));
    // fully unrolled: the remainder's word shifts and the quotient words' places are register choices
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        Q[j] = knuth_step<N>(R, b.w, next, inv);
        next = 0;
    }
    return round_quotient<N>(Q, R, b.w, e, a.s * b.s);
}

// ---------------------------------------------------- division by a reused divisor
// V = floor((2^(64N) - 1) / B) for a normalized mantissa B (N+1 words, V[N] = 1): Knuth D on the
// all-ones 2N-word dividend, whose top N words exceed B by less than B (first quotient word 1).
template <int N> __device__ __forceinline__ void reciprocal(const u32 (&b)[N], u32 (&v)[N + 1])
{
    u32 R[N + 1];
#pragma unroll
    for (int i = 0; i < N; ++i) R[i] = ~b[i]; // 2^(32N) - 1 - B
    R[N] = 0;
    v[N] = 1;
    const u32 inv = (u32)(~(u64)0 / b[N - 1] - ((u64)1 << 32));
#pragma unroll
    for (int j = N - 1; j >= 0; --j) v[j] = knuth_step<N>(R, b, 0xffffffffu, inv);
}

// The low P words of the product of two digit strings (29-bit digits, DA and DB of them): columns below
// ceil(32P / 29) only, each an in-place v_mad_u64_u32 sum (at most 36 products below 2^58), one carry pass.
template <int DA, int DB, int P>
__device__ __forceinline__ void digit_product_low(const u32 (&da)[DA], const u32 (&db)[DB], u32 (&p)[P])
{
    constexpr int K = (32 * P + DIGIT_BITS - 1) / DIGIT_BITS;
    static_assert(DA <= 37 && DB <= 37, "column sums of 29-bit digit products must fit in 64 bits"); // 37 * 2^58 + carry < 2^64
    u64 col[K];
#pragma unroll
    for (int k = 0; k < K; ++k) col[k] = 0;
#pragma unroll
    for (int i = 0; i < DA; ++i)
#pragma unroll
        for (int j = 0; j < DB; ++j)
            if (i + j < K) col[i + j] += (u64)da[i] * db[j];
    u32 dg[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        dg[k] = (u32)col[k] & DIGIT_MASK;
        if (k + 1 < K) col[k + 1] += col[k] >> DIGIT_BITS;
    }
#pragma unroll
    for (int w = 0; w < P; ++w) {
        const int bit = 32 * w, k = bit / DIGIT_BITS, s = bit % DIGIT_BITS;
        u32 v = dg[k] >> s;
        if (k + 1 < K) v |= dg[k + 1] << (DIGIT_BITS - s);
        if (k + 2 < K && 2 * DIGIT_BITS - s < 32) v |= dg[k + 2] << (2 * DIGIT_BITS - s);
        p[w] = v;
    }
}

// 29-bit digits of an M-word string (M need not be a limb count).
template <int M> constexpr int digits_of_words() { return (32 * M + DIGIT_BITS - 1) / DIGIT_BITS; }
template <int M> __device__ __forceinline__ void words_to_digits(const u32 (&w)[M], u32 (&d)[digits_of_words<M>()])
{
#pragma unroll
    for (int k = 0; k < digits_of_words<M>(); ++k) {
        const int bit = DIGIT_BITS * k, wi = bit / 32, s = bit % 32;
        const u32 lo = w[wi] >> s;
        const u32 hi = (s && wi + 1 < M) ? (w[wi + 1] << (32 - s)) : 0u;
        d[k] = (lo | hi) & DIGIT_MASK;
    }
}

// RN(a / b), bit for bit the result of div, given v = reciprocal of b's mantissa.  With U = A *
// 2^(32N-1+lt) and Q = floor(U / B) as in div: V is within one unit below 2^(64N) / B and U < 2^(64N),
// so the estimate floor(U * V / 2^(64N)) = floor(A * V / 2^(32N+1-lt)) is Q or Q - 1; the remainder
// U - estimate * B lies in [0, 2B), below 2^(32N+1), so its low N+1 words (computed modulo 2^(32(N+1)))
// decide which, and 2R against B rounds as in div.  Two products with independent multiply-adds in
// place of N dependent quotient steps: the short-latency form for divisors used many times.
template <int N> __device__ __forceinline__ Num<N> div_by_reciprocal(const Num<N>& a, const Num<N>& b, const u32 (&v)[N + 1])
{
    const u32 st = a.st | b.st;
    if (st) return zero<N>(st);
    if (!b.s) return zero<N>(ST_DIV0);
    if (!a.s) return zero<N>(ST_OK);
    const bool lt = compare_words<N>(a.w, b.w) < 0;
    i64 e = (i64)a.e - b.e - (lt ? 1 : 0);
    if (e > EMAX || e < -(i64)EMAX - 1) return zero<N>(ST_OVERFLOW);
    constexpr int DA = digits_of_words<N>(), DV = digits_of_words<N + 1>();
    u32 da[DA], dv[DV], p[2 * N + 1];
    words_to_digits<N>(a.w, da);
    words_to_digits<N + 1>(v, dv);
    digit_product_low<DA, DV, 2 * N + 1>(da, dv, p);
    u32 Q[N];
#pragma unroll
    for (int k = 0; k < N; ++k) Q[k] = lt ? p[N + k] : __builtin_amdgcn_alignbit(p[N + k + 1], p[N + k], 1u);
    // low N+1 words of Q * B and of U
    u32 dq[DA], db[DA], qb[N + 1], R[N + 1];
    words_to_digits<N>(Q, dq);
    words_to_digits<N>(b.w, db);
    digit_product_low<DA, DA, N + 1>(dq, db, qb);
#pragma unroll
    for (int k = 0; k <= N; ++k) {
        u32 uk = 0;
        if (k == N) uk = lt ? a.w[0] : __builtin_amdgcn_alignbit(a.w[1], a.w[0], 1u);
        else if (k == N - 1) uk = lt ? 0u : (a.w[0] << 31);
        R[k] = uk;
    }
    u32 bw = 0;
#pragma unroll
    for (int k = 0; k <= N; ++k) {
        R[k] = sub_borrow(R[k], qb[k], bw);
    }
    // R in [0, 2B): one more unit of the quotient when R >= B
    bool ge = R[N] != 0;
    if (!ge) ge = compare_words<N>(*reinterpret_cast<const u32(*)[N]>(R), b.w) >= 0;
    if (ge) {
        u32 c = 0;
#pragma unroll
        for (int k = 0; k <= N; ++k) {
            R[k] = sub_borrow(R[k], (k < N ? b.w[k] : 0u), c);
        }
        c = 1;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const u64 t = (u64)Q[k] + c;
            Q[k] = (u32)t;
            c = (u32)(t >> 32);
        }
    }
    return round_quotient<N>(Q, R, b.w, e, a.s * b.s);
}

// ------------------------------------------------------------------- sqrt
// RN(sqrt(a)), bit-serial: restoring integer square root of Z = mantissa * 2^(32N-1+odd), rounded up
// exactly when Z - Q^2 > Q (contract of core.hpp:216-235; docs/numerics.md "Square and square root").
// Only the top N+1 words of Z can be nonzero.  The radicand is consumed two bits at a time from a shift
// register.  The fallback of sqrt below.
template <int N> __device__ __noinline__ Num<N> sqrt_restoring(const Num<N>& a)
{
    if (a.st || !a.s) return zero<N>(a.st);
    if (a.s < 0) return zero<N>(ST_INVALID);
    int e = a.e / 2;
    if (a.e < 0 && (a.e % 2)) --e;
    const int odd = a.e - 2 * e;
    // Z's top N+1 words (the rest are zero): mantissa << (31 + odd) within N+1 words
    u32 z[N + 1];
#pragma unroll
    for (int i = 0; i <= N; ++i) {
        u32 hi = i < N ? a.w[i] : 0u, lo = i ? a.w[i - 1] : 0u;
        // word i of mantissa << (31+odd) within N+1 words
        z[i] = __builtin_amdgcn_alignbit(hi, lo, odd ? 0u : 1u);
    }
    u32 rem[N + 1], root[N];
#pragma unroll
    for (int i = 0; i <= N; ++i) rem[i] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) root[i] = 0;
    // 32N result bits; bits 2*bit, 2*bit+1 of Z for bit = 32N-1 .. 0. Words of Z from the top:
    // z[N], z[N-1], ..., z[0], then N-1 zero words.
    for (int word = 0; word < 2 * N; ++word) {
        u32 cur = z[N];
#pragma unroll
        for (int i = N; i > 0; --i) z[i] = z[i - 1];
        z[0] = 0;
        for (int pair = 0; pair < 16; ++pair) {
            const u32 two = cur >> 30;
            cur <<= 2;
            // rem = 4 rem + two
#pragma unroll
            for (int i = N; i >= 0; --i) rem[i] = (rem[i] << 2) | (i ? rem[i - 1] >> 30 : two);
            // trial = 4 root_old + 1 = (root << 2) | 1 with root the value before this bit
            u32 trial[N + 1];
#pragma unroll
            for (int i = N; i >= 0; --i) {
                u32 hi = i < N ? root[i] : 0u, lo = i ? root[i - 1] : 0u;
                trial[i] = __builtin_amdgcn_alignbit(hi, lo, 30u);
            }
            trial[0] |= 1u;
            u32 bw = 0;
            u32 diff[N + 1];
#pragma unroll
            for (int i = 0; i <= N; ++i) {
                diff[i] = sub_borrow(rem[i], trial[i], bw);
            }
            const bool ge = bw == 0;
#pragma unroll
            for (int i = 0; i <= N; ++i) rem[i] = ge ? diff[i] : rem[i];
            // root = 2 root + ge
#pragma unroll
            for (int i = N - 1; i >= 0; --i) root[i] = (root[i] << 1) | (i ? root[i - 1] >> 31 : (ge ? 1u : 0u));
        }
    }
    // round up iff rem > root
    bool gt;
    if (rem[N]) gt = true;
    else {
        u32 rl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) rl[i] = rem[i];
        gt = compare_words<N>(rl, root) > 0;
    }
    Num<N> r;
    r.s = 1;
    r.st = ST_OK;
    u32 c = gt ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u64 t = (u64)root[i] + c;
        r.w[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    if (c) {
        r.w[N - 1] = 0x80000000u;
        e += 1;
    }
    r.e = e;
    return r;
}

// Bits [P, P + 32K) of a W-word string, P a compile-time bit position (words past the end read as 0).
template <int K, int P, int W> __device__ __forceinline__ void bits_at(const u32 (&a)[W], u32 (&o)[K])
{
    constexpr int q = P / 32, b = P % 32;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const u32 lo = q + k < W ? a[q + k] : 0u, hi = q + k + 1 < W ? a[q + k + 1] : 0u;
        o[k] = b ? __builtin_amdgcn_alignbit(hi, lo, (u32)b) : lo;
    }
}

// Z ~ 2^F / sqrt(x), F = 32M - 3, for x = (top M words of the W-word X) / 2^(32M) in [1/4, 1): Newton's
// iteration for the reciprocal square root, z <- z (3 - x z^2) / 2, in M-word fixed point.  Every term
// stays non-negative (x z^2 is near 1, so 3 - x z^2 is near 2), so each step is three integer products
// and a subtraction from a constant, with no normalization shifts; the step at M words refines the result
// at M/2 + 1 words (each step doubles the accurate bits), down to two words seeded by rsqrt_seed from X's
// top 64 bits.  Z <= 2^(F+1) * (1 + tiny) fits in M words.
//
// rsqrt_seed: z ~ 2^61 / sqrt(x) for x = top / 2^64 in [1/4, 1), in integer arithmetic: a two-piece
// constant start (relative error below 0.2) refined by five Newton steps z <- z (3 - x z^2) / 2 in 64-bit
// fixed point with 2^61 scaling (products through 128-bit intermediates; each step doubles the accurate
// bits: about 2.5, 5, 10, 20, 40 and then the 60 the format holds, far above the 31 that the two-word step
// of rsqrt_fixed needs to continue the doubling).
__device__ __forceinline__ u64 mul_hi64(u64 a, u64 b)
{
    const u64 al = (u32)a, ah = a >> 32, bl = (u32)b, bh = b >> 32;
    const u64 ll = al * bl, lh = al * bh, hl = ah * bl, hh = ah * bh;
    const u64 mid = (ll >> 32) + (u32)lh + (u32)hl;
    return hh + (lh >> 32) + (hl >> 32) + (mid >> 32);
}
__device__ __forceinline__ u64 rsqrt_seed(u64 top)
{
    const u64 three = (u64)3 << 61;
    u64 z = top >= ((u64)1 << 63) ? (u64)0x2666666666666666ull /* 1.2 * 2^61 */ : (u64)0x3666666666666666ull; // 1.7 * 2^61
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const u64 z2 = mul_hi64(z, z) << 3;        // z^2 * 2^61 (z^2 <= 4: below 2^63)
        const u64 xz2 = mul_hi64(top, z2);         // x z^2 * 2^61
        const u64 t = three - xz2;                 // (3 - x z^2) * 2^61, near 2 * 2^61
        z = mul_hi64(z, t) << 2;                   // z (3 - x z^2) / 2 * 2^61
    }
    return z;
}

template <int M, int W> __device__ void rsqrt_fixed(const u32 (&X)[W], u32 (&Z)[M])
{
    constexpr int F = 32 * M - 3, DM = digits_of_words<M>();
    if constexpr (M <= 2) {
        static_assert(M == 2, "the seed fills two words");
        const u64 z = rsqrt_seed(((u64)X[W - 1] << 32) | X[W - 2]);
        Z[0] = (u32)z;
        Z[1] = (u32)(z >> 32);
    } else {
        constexpr int MP = M / 2 + 1;
        u32 zp[MP];
        rsqrt_fixed<MP, W>(X, zp);
#pragma unroll
        for (int i = 0; i < M; ++i) Z[i] = i >= M - MP ? zp[i - (M - MP)] : 0u; // 2^(F - F') = 2^(32(M - MP))
    }
    u32 dz[DM], dx[DM], da[DM], dd[DM], A[2 * M], Ap[M], Xm[M], B[2 * M], D[M], P[2 * M];
    words_to_digits<M>(Z, dz);
    digit_product_low<DM, DM, 2 * M>(dz, dz, A);
    bits_at<M, F>(A, Ap); // z^2 * 2^F
#pragma unroll
    for (int i = 0; i < M; ++i) Xm[i] = X[W - M + i];
    words_to_digits<M>(Xm, dx);
    words_to_digits<M>(Ap, da);
    digit_product_low<DM, DM, 2 * M>(dx, da, B); // x z^2 * 2^(F + 32M)
    u32 bw = 0;
#pragma unroll
    for (int i = 0; i < M; ++i) { // D = 3 * 2^F - x z^2 * 2^F
        D[i] = sub_borrow(i == M - 1 ? 0x60000000u : 0u, B[M + i], bw);
    }
    words_to_digits<M>(D, dd);
    digit_product_low<DM, DM, 2 * M>(dz, dd, P);
    bits_at<M, F + 1>(P, Z); // z (3 - x z^2) / 2 * 2^F
}

// RN(sqrt(a)).  The mantissa m (N words, top bit set) and the exponent parity define the integer
// X = m * 2^(32N-1+odd) < 2^(64N), whose integer square root r has exactly 32N bits; the result is r, plus
// one ulp exactly when X - r^2 > r (the contract of core.hpp:216-235: no ties are possible).  r starts as
// x * z from the fixed-point reciprocal square root z of x = X / 2^(64N) at N+1 words (rsqrt_fixed:
// products only), which lands within a unit or two; one squaring gives X - r^2 and unit steps on r and the
// remainder (X - (r+1)^2 = X - r^2 - (2r+1)) make r^2 <= X < (r+1)^2 exactly, so the result does not depend
// on how close the estimate is.  Should that take more than a few steps, the bit-serial restoring method
// decides.
template <int N> __device__ __forceinline__ Num<N> sqrt(const Num<N>& a)
{
    if (a.st || !a.s) return zero<N>(a.st);
    if (a.s < 0) return zero<N>(ST_INVALID);
    int e = a.e / 2;
    if (a.e < 0 && (a.e % 2)) --e;
    const int odd = a.e - 2 * e;
    // X: words 0..N-2 zero, words N-1..2N-1 = m << (31 + odd) in N+1 words
    u32 X[2 * N];
#pragma unroll
    for (int i = 0; i < N - 1; ++i) X[i] = 0;
#pragma unroll
    for (int i = 0; i <= N; ++i) {
        const u32 hi = i < N ? a.w[i] : 0u, lo = i ? a.w[i - 1] : 0u;
        X[N - 1 + i] = __builtin_amdgcn_alignbit(hi, lo, odd ? 0u : 1u);
    }
    // r = x z 2^(32N) = Xt Z / 2^(32 + F), Xt = top N+1 words of X, F = 32(N+1) - 3
    constexpr int DT = digits_of_words<N + 1>();
    u32 Z[N + 1], Xt[N + 1], dz[DT], dt[DT], P[2 * N + 2], r[N], over[1];
    rsqrt_fixed<N + 1, 2 * N>(X, Z);
#pragma unroll
    for (int i = 0; i <= N; ++i) Xt[i] = X[N - 1 + i];
    words_to_digits<N + 1>(Z, dz);
    words_to_digits<N + 1>(Xt, dt);
    digit_product_low<DT, DT, 2 * N + 2>(dt, dz, P);
    bits_at<N, 32 * N + 61>(P, r);
    bits_at<1, 64 * N + 61>(P, over);
    if (over[0]) // estimate at or above 2^(32N): start from the largest 32N-bit value
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = 0xffffffffu;
    else if (!(r[N - 1] >> 31)) // below 2^(32N-1): start from the smallest
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = i == N - 1 ? 0x80000000u : 0u;
    // rem = X - r^2 as a (2N+1)-word two's-complement value
    u32 rem[2 * N + 1];
    {
        u32 sq[2 * N];
        mantissa_product<N>(r, r, sq);
        u32 bw = 0;
#pragma unroll
        for (int i = 0; i < 2 * N; ++i) {
            rem[i] = sub_borrow(X[i], sq[i], bw);
        }
        rem[2 * N] = bw ? 0xffffffffu : 0u;
    }
    bool done = false;
#pragma unroll 1
    for (int step = 0; step < 6 && !done; ++step) {
        // t2 = 2r + 1 (N+1 words)
        u32 t2[N + 1];
#pragma unroll
        for (int i = 0; i <= N; ++i) {
            const u32 hi = i < N ? r[i] : 0u, lo = i ? r[i - 1] : 0u;
            t2[i] = __builtin_amdgcn_alignbit(hi, lo, 31u);
        }
        t2[0] |= 1u;
        if (rem[2 * N] >> 31) { // r^2 > X: r <- r - 1, rem <- rem + 2r - 1 (= rem + t2 - 2)
            u32 c = 0;
#pragma unroll
            for (int i = 0; i <= 2 * N; ++i) {
                const u64 t = (u64)rem[i] + (i <= N ? t2[i] : 0u) + c;
                rem[i] = (u32)t;
                c = (u32)(t >> 32);
            }
            u32 bw = 2;
#pragma unroll
            for (int i = 0; i <= 2 * N; ++i) {
                rem[i] = sub_borrow(rem[i], 0u, bw);
            }
            u32 b = 1;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                r[i] = sub_borrow(r[i], 0u, b);
            }
            continue;
        }
        // (r+1)^2 <= X  <=>  rem >= 2r + 1: r <- r + 1, rem <- rem - (2r + 1)
        bool high = false;
#pragma unroll
        for (int i = N + 1; i <= 2 * N; ++i) high = high || rem[i] != 0;
        u32 rl[N + 1];
#pragma unroll
        for (int i = 0; i <= N; ++i) rl[i] = rem[i];
        if (high || compare_words<N + 1>(rl, t2) >= 0) {
            u32 bw = 0;
#pragma unroll
            for (int i = 0; i <= 2 * N; ++i) {
                rem[i] = sub_borrow(rem[i], (i <= N ? t2[i] : 0u), bw);
            }
            u32 c = 1;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const u64 t = (u64)r[i] + c;
                r[i] = (u32)t;
                c = (u32)(t >> 32);
            }
            continue;
        }
        done = true;
    }
    if (!done) return sqrt_restoring<N>(a);
    // round up iff X - r^2 > r (the remainder is below 2r + 1 < 2^(32N+1))
    u32 rr[N + 1], rl[N + 1];
#pragma unroll
    for (int i = 0; i <= N; ++i) {
        rr[i] = i < N ? r[i] : 0u;
        rl[i] = rem[i];
    }
    const bool gt = compare_words<N + 1>(rl, rr) > 0;
    Num<N> res;
    res.s = 1;
    res.st = ST_OK;
    u32 c = gt ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const u64 t = (u64)r[i] + c;
        res.w[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    if (c) {
        res.w[N - 1] = 0x80000000u;
        e += 1;
    }
    res.e = e;
    return res;
}

// -------------------------------------------------------------------- fma
// acc (two's complement) += sign * x
template <int W> __device__ __forceinline__ void accumulate_signed(u32 (&acc)[W], const u32 (&x)[W], int sign)
{
    if (sign > 0) {
        u32 cy = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            u64 t = (u64)acc[i] + x[i] + cy;
            acc[i] = (u32)t;
            cy = (u32)(t >> 32);
        }
    } else {
        u32 bw = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            acc[i] = sub_borrow(acc[i], x[i], bw);
        }
    }
}

// acc (two's complement) += sign (one unit at bit 0)
template <int W> __device__ __forceinline__ void accumulate_unit(u32 (&acc)[W], int sign)
{
    const u32 add = sign > 0 ? 1u : 0xffffffffu; // +1 or -1 (sign-extended)
    u32 cy = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        u64 t = (u64)acc[i] + (i ? (sign > 0 ? 0u : 0xffffffffu) : add) + cy;
        acc[i] = (u32)t;
        cy = (u32)(t >> 32);
    }
}

// RN(a*b + c) with one rounding (the mpfr_fma contract of core.hpp:317-326).  With P = a*b exact
// (2N words) and C = c (N words), three cases, each giving the one correctly rounded result:
//   c below half an ulp of P's last bit   only c's sign matters: P*2^32 +/- 1 (a sticky unit) is rounded;
//   P below a quarter ulp of c           RN(c + p) = c (the nearest neighbours of c are at least half an
//                                         ulp of c away, a quarter below a power of two);
//   otherwise                            a (2N+3)-word window under the larger term's top bit holds the sum
//                                         exactly, or up to a sticky unit when the smaller term reaches
//                                         below it (fma_tail), rounded once.
// The cases after P is formed: p = the 2N-word product of two nonzero status-free operands whose
// exponents sum to eab, ps = the product's sign.
// Whether c is zero or |c| < 2^(p_lsb-1), half a unit of the last bit of the product P (2N words, exponent
// sum eab): the exact sum is then P + t * delta, t the sign of c relative to P (0 for a zero c),
// 0 < delta < that half unit, which is P's rounding with the tail nudged by t.
template <int N> __device__ __forceinline__ bool fma_addend_below(i64 eab, const Num<N>& c)
{
    return !c.s || (i64)c.e < eab - 2 * (32 * N - 1) - 1;
}

template <int N> __device__ __forceinline__ Num<N> fma_tail(const u32 (&p)[2 * N], i64 eab, int ps, const Num<N>& c)
{
    // one branch for a zero and a far addend, so a wavefront whose lanes mix plain products (the
    // multiplication steps of a chain) and far addends runs one rounding of P, not two
    if (fma_addend_below<N>(eab, c)) return round_product<N>(p, eab, ps, ST_OK, false, ps * c.s);
    const bool top = (p[2 * N - 1] >> 31) != 0;
    const i64 p_msb = eab + (top ? 1 : 0);
    const i64 c_msb = c.e, c_lsb = (i64)c.e - (32 * N - 1);
    if (p_msb < c_lsb - 2) return c; // |p| < 2^(c_lsb-2)
    // Near case: a window of W = 2N+3 words with the larger term's top bit (top bit of P or of c, the one
    // with the higher exponent) at window bit 32W-2; the smaller term, aligned to it, may reach below bit 0
    // only when it lies 96 or more bits below the larger one's top (the larger term's lowest bit is at
    // window bit 95 or higher), and then its bits below the window only decide the sticky flag.  The sum is
    // formed once for both signs: T = big + (small or its complement) + carry-in, which is floor of the
    // exact window value (a subtraction with a sticky remainder is one unit lower), and a sticky remainder
    // sets T's last bit, so T lies in the same rounding cell as the exact value and is no tie (no rounding
    // boundary of the N-word result lies within one window unit, and no cancellation beyond one bit is
    // possible when a sticky remainder is).
    constexpr int W = 2 * N + 3;
    const bool p_big = p_msb >= c_msb;
    const i64 H = p_big ? p_msb : c_msb;
    // the exponent gap: at most 64N with P the larger (c_msb >= p_lsb - 1) and 32N + 1 with c the larger
    // (p_msb >= c_lsb - 2), so the small term moves by at most 2N words
    const u32 d = (u32)(p_big ? p_msb - c_msb : c_msb - p_msb);
    // top-aligned images: P shifted left by 64 + pb bits (its top bit, 64N-1 or 64N-2, to 64N+94); c shifted
    // left by 32(N+2) + 31 bits (its top bit 32N-1 to 64N+94)
    const u32 pb = top ? 31u : 32u;
    u32 big[W], sm[W];
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const u32 phi = (i - 2 >= 0 && i - 2 < 2 * N) ? p[i - 2] : 0u, plo = (i - 3 >= 0 && i - 3 < 2 * N) ? p[i - 3] : 0u;
        const u32 pa = __builtin_amdgcn_alignbit(phi, plo, 32u - pb);
        const u32 chi = (i - (N + 2) >= 0 && i - (N + 2) < N) ? c.w[i - (N + 2)] : 0u;
        const u32 clo = (i - (N + 3) >= 0 && i - (N + 3) < N) ? c.w[i - (N + 3)] : 0u;
        const u32 ca = __builtin_amdgcn_alignbit(chi, clo, 1u);
        big[i] = p_big ? pa : ca;
        sm[i] = p_big ? ca : pa;
    }
    const bool sticky = shr_sticky<W, 2 * N>(sm, d);
    const int sb = p_big ? ps : c.s, ss = p_big ? c.s : ps;
    const bool sub = sb != ss;
    const u32 flip = sub ? 0xffffffffu : 0u;
    u32 T[W], cy = (sub && !sticky) ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < W; ++i) T[i] = add_carry(big[i], sm[i] ^ flip, cy);
    T[0] |= sticky ? 1u : 0u;
    int sign = sb;
    if (sub && (T[W - 1] >> 31)) { // the smaller-exponent term had the larger mantissa (d <= 1, no sticky): negate
        sign = ss;
        u32 c1 = 1;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const u64 v = (u64)(u32)~T[i] + c1;
            T[i] = (u32)v;
            c1 = (u32)(v >> 32);
        }
    }
    return pack<N, W>(T, H - (64 * N + 94), sign, ST_OK);
}

template <int N> __device__ __forceinline__ Num<N> fma(const Num<N>& a, const Num<N>& b, const Num<N>& c)
{
    const u32 st = a.st | b.st | c.st;
    if (st) return zero<N>(st);
    const int ps = a.s * b.s;
    if (!ps) return c.s ? c : zero<N>(ST_OK); // RN(c) = c
    u32 p[2 * N];
    mantissa_product<N>(a.w, b.w, p);
    return fma_tail<N>(p, (i64)a.e + b.e, ps, c);
}

// fma with the digits of b's mantissa given (a multiplier reused over a chain); needs use_digits<N>().
// With c = 0 this is RN(a * b), the result of mul_by_digits: a chain whose steps mix mul and fma runs
// every step through here with c = 0 on its mul steps, so the lanes of a wave share one product per step.
template <int N>
__device__ __forceinline__ Num<N> fma_by_digits(const Num<N>& a, const Num<N>& b, const u32 (&bd)[digit_count<N>()], const Num<N>& c)
{
    const u32 st = a.st | b.st | c.st;
    const int ps = a.s * b.s;
    const bool kill = st || !ps;
    u32 ad[digit_count<N>()], p[2 * N];
    to_digits<N>(a.w, ad);
    rounding_product<N>(ad, bd, p, c.s != 0, kill); // with c = 0 only the rounding of P is needed
    // c = 0 (a chain's mul step) and status operands: the product's rounding, cleared when kill; an addend
    // below half a unit of P's last bit: the same rounding with the tail nudged (fma_tail's first case),
    // taken here so that the two share one copy of the rounding
    const i64 eab = (i64)a.e + b.e;
    const bool plain = !c.s || st;
    if (plain || (ps && fma_addend_below<N>(eab, c))) return round_product<N>(p, eab, ps, st, kill, plain ? 0 : ps * c.s);
    if (!ps) return c; // RN(c) = c
    return fma_tail<N>(p, eab, ps, c);
}

// Position of the highest set bit of x, or -1 (a scan over compile-time indices).
template <int W> __device__ __forceinline__ int highest_bit(const u32 (&x)[W])
{
    int h = -1;
#pragma unroll
    for (int i = 0; i < W; ++i)
        if (x[i]) h = 32 * i + 31 - __clz(x[i]);
    return h;
}

// RN(s1 * int(t1) * 2^e1 + s2 * int(t2) * 2^e2) with one rounding, for exact terms of W1 and W2 words
// (either may be zero).  The window holds the larger term entirely and reaches far enough below it that
// a smaller term losing bits below the window lies more than 32N+60 bits below the larger msb, where it
// only contributes the sign of one sticky unit (the argument of core.hpp:251-253).  Used for the update
// form RN(c - sum) of the dense products (linalg.metal:13-16,100-105 round c against the exact sum).
template <int N, int W1, int W2>
__device__ __forceinline__ Num<N> round_sum2(const u32 (&t1)[W1], i64 e1, int s1, const u32 (&t2)[W2], i64 e2, int s2)
{
    const int h1 = highest_bit<W1>(t1), h2 = highest_bit<W2>(t2);
    if (h1 < 0) s1 = 0;
    if (h2 < 0) s2 = 0;
    if (!s1 && !s2) return zero<N>(ST_OK);
    const i64 m1 = e1 + h1, m2 = e2 + h2;
    const i64 mx = !s2 ? m1 : !s1 ? m2 : (m1 > m2 ? m1 : m2);
    constexpr int FW = W1 + W2 + N + 2;
    const i64 L = mx + 3 - 32 * (i64)FW; // the larger msb sits at window bit 32FW-3
    u32 acc[FW], w[FW];
#pragma unroll
    for (int i = 0; i < FW; ++i) acc[i] = 0;
    bool d1 = false, d2 = false;
    if (s1) {
        d1 = place<FW, W1>(w, t1, e1 - L);
        accumulate_signed<FW>(acc, w, s1);
    }
    if (s2) {
        d2 = place<FW, W2>(w, t2, e2 - L);
        accumulate_signed<FW>(acc, w, s2);
    }
    i64 scale = L;
    if (d1 || d2) {
        shl_bits(acc, 1);
        accumulate_unit<FW>(acc, d1 ? s1 : s2);
        scale = L - 1;
    }
    int sign = 1;
    if (acc[FW - 1] >> 31) {
        sign = -1;
        u32 cy = 1;
#pragma unroll
        for (int i = 0; i < FW; ++i) {
            u64 v = (u64)(u32)~acc[i] + cy;
            acc[i] = (u32)v;
            cy = (u32)(v >> 32);
        }
    }
    return pack<N, FW>(acc, scale, sign, ST_OK);
}

template <int N> __device__ __forceinline__ Num<N> fms(const Num<N>& a, const Num<N>& b, const Num<N>& c)
{
    return fma<N>(a, b, negate<N>(c));
}

// ---------------------------------------------------------------- complex
// Composed complex operations, each real primitive rounded (core.hpp:414-419).
template <int N> struct Cpx {
    Num<N> re, im;
};

template <int N> __device__ __forceinline__ Cpx<N> cadd(const Cpx<N>& a, const Cpx<N>& b)
{
    return {add<N>(a.re, b.re), add<N>(a.im, b.im)};
}

template <int N> __device__ __forceinline__ Cpx<N> cmul(const Cpx<N>& a, const Cpx<N>& b)
{
    Num<N> re = sub<N>(mul<N>(a.re, b.re), mul<N>(a.im, b.im));
    Num<N> im = add<N>(mul<N>(a.re, b.im), mul<N>(a.im, b.re));
    return {re, im};
}

template <int N> __device__ __forceinline__ Cpx<N> cdiv(const Cpx<N>& a, const Cpx<N>& b)
{
    Num<N> d = add<N>(mul<N>(b.re, b.re), mul<N>(b.im, b.im));
    const Num<N> nre = add<N>(mul<N>(a.re, b.re), mul<N>(a.im, b.im));
    const Num<N> nim = sub<N>(mul<N>(a.im, b.re), mul<N>(a.re, b.im));
    // both parts divide by d: its reciprocal once, then two divisions by it (each RN(n / d), as div)
    u32 v[N + 1];
    if (d.s && !d.st) reciprocal<N>(d.w, v);
    else
#pragma unroll
        for (int k = 0; k <= N; ++k) v[k] = 0; // not read: a zero or status divisor is decided first
    return {div_by_reciprocal<N>(nre, d, v), div_by_reciprocal<N>(nim, d, v)};
}

// ---------------------------------------------------------------- casts
// Exact widening, RN narrowing from M words to N words (kernels.metal:251-258).
template <int N, int M> __device__ __forceinline__ Num<N> cast(const Num<M>& x)
{
    if (x.st || !x.s) return zero<N>(x.st);
    u32 t[M];
#pragma unroll
    for (int i = 0; i < M; ++i) t[i] = x.w[i];
    // x is normalized: msb = x.e
    return round_normalized<N, M>(t, (i64)x.e, x.s, ST_OK);
}

} // namespace lf

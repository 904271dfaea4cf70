// coop_core.hpp — cooperative multiplications: G consecutive lanes of a wavefront (a group, G = 2 or 4,
// inside one quad) form the digit columns of one value's product together.
//
// Every lane of a group holds the whole value and the whole factor (one value's state is replicated);
// what is split is the product's digit columns, the bulk of a multiplication's work.  Lane k of the group
// forms the columns c_lo + k + G m (m = 0, 1, ...) of the 29-bit digit product (limb_core.hpp digit
// columns), so each lane does about 1/G of the multiply-adds.  For the index arithmetic to be the same in
// every lane (registers are indexed at compile time), lane k holds the factor's digits shifted up by
// G - 1 - k places, bs[j'] = bd[j' - (G - 1 - k)] (zero outside [0, D)): then column c_lo + k + G m is
// sum over i + j' = c_lo + G m + G - 1 of ad[i] bs[j'], with no k in the indices.  The factor is fixed for
// a chain, so bs is formed once per chain.  The column sums are then exchanged inside the quad by DPP
// (quad_perm reads of each group lane), after which every lane of the group holds every column and runs
// the same carry pass and rounding (limb_core.hpp pack_columns, round_product): the result is the
// single-lane result bit for bit, in all G lanes.
#pragma once
#include "limb_core.hpp"

namespace lf {

template <int N, int G, bool FULL> constexpr int coop_lo() { return FULL ? 0 : short_cut<N>(); }
// columns per lane: ceil((2D - 1 - c_lo) / G) (column 2D-1 is never formed: it is the top carry)
template <int N, int G, bool FULL> constexpr int coop_cols() { return (2 * digit_count<N>() - 1 - coop_lo<N, G, FULL>() + G - 1) / G; }
template <int N, int G> constexpr int coop_factor_digits() { return digit_count<N>() + G - 1; }

// bs[j'] = bd[j' - (G - 1 - k)] for lane k of a group, zero outside the factor's digits
template <int N, int G>
__device__ __forceinline__ void coop_shift_factor(const u32 (&bd)[digit_count<N>()], u32 k, u32 (&bs)[coop_factor_digits<N, G>()])
{
    constexpr int D = digit_count<N>();
#pragma unroll
    for (int jp = 0; jp < D + G - 1; ++jp) {
        u32 v = 0;
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            const int j = jp - (G - 1 - kk);
            if (j >= 0 && j < D) v = (k == (u32)kk) ? bd[j] : v;
        }
        bs[jp] = v;
    }
}

// S[m] = column c_lo + k + G m of the digit product ad * bd, for the lane whose shifted factor is bs
template <int N, int G, bool FULL>
__device__ __forceinline__ void coop_partial(const u32 (&ad)[digit_count<N>()], const u32 (&bs)[coop_factor_digits<N, G>()],
                                             u64 (&S)[coop_cols<N, G, FULL>()])
{
    constexpr int D = digit_count<N>(), LO = coop_lo<N, G, FULL>(), M = coop_cols<N, G, FULL>();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int c = LO + G * m + G - 1; // i + j' = c
        u64 s = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const int jp = c - i;
            if (jp >= 0 && jp < D + G - 1) s += (u64)ad[i] * bs[jp];
        }
        S[m] = s;
    }
}

#ifndef LF_HOST_BUILD
// Value of lane `src` (0 <= src < G) of this lane's group, through DPP inside the quad.
template <int G, int SRC> __device__ __forceinline__ u32 coop_read(u32 v)
{
    static_assert(G == 2 || G == 4, "groups of 2 or 4 lanes");
    constexpr int ctrl = G == 4 ? (SRC | SRC << 2 | SRC << 4 | SRC << 6) : (SRC | SRC << 2 | (2 + SRC) << 4 | (2 + SRC) << 6);
    return (u32)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
}

template <int G, int SRC> __device__ __forceinline__ u64 coop_read64(u64 v)
{
    return (u64)coop_read<G, SRC>((u32)v) | ((u64)coop_read<G, SRC>((u32)(v >> 32)) << 32);
}

// col[c] = column c of the product for c >= c_lo (the columns below are zero), in every lane of the group
template <int N, int G, bool FULL>
__device__ __forceinline__ void coop_gather(const u64 (&S)[coop_cols<N, G, FULL>()], u64 (&col)[2 * digit_count<N>()])
{
    constexpr int D = digit_count<N>(), LO = coop_lo<N, G, FULL>(), M = coop_cols<N, G, FULL>();
#pragma unroll
    for (int c = 0; c < 2 * D; ++c) col[c] = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        if (LO + G * m + 0 < 2 * D - 1) col[LO + G * m + 0] = coop_read64<G, 0>(S[m]);
        if (LO + G * m + 1 < 2 * D - 1) col[LO + G * m + 1] = coop_read64<G, 1>(S[m]);
        if constexpr (G == 4) {
            if (LO + G * m + 2 < 2 * D - 1) col[LO + G * m + 2] = coop_read64<G, 2>(S[m]);
            if (LO + G * m + 3 < 2 * D - 1) col[LO + G * m + 3] = coop_read64<G, 3>(S[m]);
        }
    }
}

// RN(a * b) with the factor's digits given twice: bd (for the rare exact fallback) and bs (shifted for
// lane k of the group); the mul_by_digits contract, the same result in every lane of the group.
template <int N, int G>
__device__ __forceinline__ Num<N> coop_mul_by_digits(const Num<N>& a, const Num<N>& b, const u32 (&bd)[digit_count<N>()],
                                                     const u32 (&bs)[coop_factor_digits<N, G>()])
{
    const u32 st = a.st | b.st;
    const bool kill = st || !a.s || !b.s;
    u32 ad[digit_count<N>()], p[2 * N];
    to_digits<N>(a.w, ad);
    u64 S[coop_cols<N, G, false>()], col[2 * digit_count<N>()];
    coop_partial<N, G, false>(ad, bs, S);
    coop_gather<N, G, false>(S, col);
    pack_columns<N, true>(col, p, false);
    // the group's lanes hold the same p, so they take this branch together
    if (!kill && !short_decides<N>(p)) digit_columns<N, true>(ad, bd, p, true);
    return round_product<N>(p, (i64)a.e + b.e, a.s * b.s, st, kill);
}

// The digit-column product of ad and bd, split over the group's lanes: the exact product p (2N words)
// when full, else the short product (the digit_columns contract, without HIGH).
template <int N, int G, bool FULL>
__device__ __forceinline__ void coop_columns(const u32 (&ad)[digit_count<N>()], const u32 (&bs)[coop_factor_digits<N, G>()], u32 (&p)[2 * N])
{
    u64 S[coop_cols<N, G, FULL>()], col[2 * digit_count<N>()];
    coop_partial<N, G, FULL>(ad, bs, S);
    coop_gather<N, G, FULL>(S, col);
    pack_columns<N, false>(col, p, FULL);
}

// RN(a * b + c) with one rounding (the fma_by_digits contract), the product's columns split over the
// group's lanes: the exact product when c is nonzero, else the short product, made exact when it does
// not decide the rounding.  Every lane of the group returns the same value.
template <int N, int G>
__device__ __forceinline__ Num<N> coop_fma_by_digits(const Num<N>& a, const Num<N>& b, const u32 (&bs)[coop_factor_digits<N, G>()],
                                                     const Num<N>& c)
{
    const u32 st = a.st | b.st | c.st;
    const int ps = a.s * b.s;
    const bool kill = st || !ps;
    u32 ad[digit_count<N>()], p[2 * N];
    to_digits<N>(a.w, ad);
    // the group's lanes hold the same operands and columns, so they take these branches together
    bool full = c.s != 0;
    if (!full) {
        coop_columns<N, G, false>(ad, bs, p);
        full = !kill && !short_decides<N>(p);
    }
    if (full) coop_columns<N, G, true>(ad, bs, p);
    if (!c.s || st) return round_product<N>(p, (i64)a.e + b.e, ps, st, kill);
    if (!ps) return c; // RN(c) = c
    return fma_tail<N>(p, (i64)a.e + b.e, ps, c);
}
#endif

} // namespace lf

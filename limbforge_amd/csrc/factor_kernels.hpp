// factor_kernels.hpp — kernels of the blocked Cholesky factorization and the triangular solves on a
// device-resident work matrix in limb planes (row-major element i*n + j).
//
// The rounding sequence is the reference's (docs/numerics.md:674-748, src/linalg.mm:411-486): with
// D(c; x, y; K) = RN(c - sum_{k<K} x_k y_k) rounded once,
//   panel [k0, k1): for every column j in order, s_jj = D(a_jj; l_j, l_j; [k0, j)), l_jj = sqrt(s_jj),
//                   then l_ij = div(D(a_ij; l_i, l_j; [k0, j)), l_jj) for the rows i > j;
//   the trailing updates by a block are products of the update GEMM (gemm.hip), not kernels here;
//   forward solve in block [r0, r1): x_i = div(D(b_i; l_i, x; [r0, i)), l_ii), i ascending;
//   backward solve in block [r0, r1): x_i = div(D(b_i; L[., i], x; (i, r1)), l_ii), i descending.
// Each entry depends only on final entries, so computing a panel column by column (all rows of the
// column in parallel) gives the bits of the reference's row-by-row host order.
//
// D is computed in one of two exact ways, chosen per call from the terms' exponents (same result):
//   window    every exact 2N-word product and c added into one W-word two's-complement register window
//   clusters  exponent clusters with a sticky sign (the reference's linalg.hpp:298-317), each cluster
//             summed exactly in the thread's own global scratch (positive and negative accumulators)
#pragma once
#include "factor_launch.hpp"
#include "gemm_kernels.hpp"

namespace lf {

// A vector of plane elements: element k at plane index base + k*stride.
struct VecView {
    const u32* planes;
    u64 pitch;
    i64 base, stride;
};

__device__ __forceinline__ u64 vec_at(const VecView& v, u32 k) { return (u64)(v.base + (i64)k * v.stride); }

template <int N> __device__ __forceinline__ void vec_meta(const VecView& v, u32 k, int& e, int& s, u32& st)
{
    const u64 a = vec_at(v, k);
    e = (int)v.planes[(u64)N * v.pitch + a];
    s = (int)v.planes[(u64)(N + 1) * v.pitch + a];
    st = v.planes[(u64)(N + 2) * v.pitch + a];
}

template <int N> __device__ __forceinline__ Num<N> vec_load(const VecView& v, u32 k) { return load_out<N>(v.planes, v.pitch, vec_at(v, k)); }

// Per-thread scratch of the cluster path: two accumulators of `words` words each.
struct DotScratch {
    u32* base;
    u64 words;
    int32_t* error; // set (and the result zeroed) if a cluster would not fit: the host sizes scratch so that none does
};

template <int N>
__device__ __noinline__ Num<N> dot_sub_clusters(const Num<N>& c, const VecView& x, const VecView& y, u32 K, const DotScratch& sc)
{
    constexpr i64 NONE = -((i64)1 << 62);
    int logk = 0;
    while (((u64)1 << logk) < (u64)K + 1) ++logk;
    const i64 width = 64 * N, gap = 32 * N + logk + 4;
    // term k < K: -x_k y_k; term K: c
    auto scale_of = [&](u32 k, int& sg) -> i64 {
        if (k == K) {
            sg = c.s;
            return (i64)c.e - (32 * N - 1);
        }
        int ex, sx, ey, sy;
        u32 tx, ty;
        vec_meta<N>(x, k, ex, sx, tx);
        vec_meta<N>(y, k, ey, sy, ty);
        sg = -sx * sy;
        return sg ? (i64)ex + ey - 2 * (32 * N - 1) : NONE;
    };
    // next cluster below `cutoff`: its highest and lowest term scales; false when none is left
    auto next = [&](i64 cutoff, i64& hi, i64& low) -> bool {
        hi = NONE;
        for (u32 k = 0; k <= K; ++k) {
            int sg;
            const i64 s = scale_of(k, sg);
            if (sg && s < cutoff && s > hi) hi = s;
        }
        if (hi == NONE) return false;
        low = hi;
        for (bool changed = true; changed;) {
            changed = false;
            for (u32 k = 0; k <= K; ++k) {
                int sg;
                const i64 s = scale_of(k, sg);
                if (sg && s < low && s + width - 1 + gap >= low) {
                    low = s;
                    changed = true;
                }
            }
        }
        return true;
    };
    constexpr int KW = N + 4; // kept words of the first nonzero cluster sum: a sticky word, then the top N+3
    u32 S[KW];
#pragma unroll
    for (int w = 0; w < KW; ++w) S[w] = 0;
    i64 scale1 = 0, cutoff = ((i64)1 << 62), hi, low;
    int nsums = 0, sign1 = 0, sign2 = 0;
    while (nsums < 2 && next(cutoff, hi, low)) {
        const u64 W = (u64)((hi + width - low) / 32 + 3);
        if (W > sc.words) {
            *sc.error = 1;
            return zero<N>(ST_OK);
        }
        u32* pos = sc.base;
        u32* neg = sc.base + W;
        for (u64 w = 0; w < 2 * W; ++w) pos[w] = 0;
        for (u32 k = 0; k <= K; ++k) {
            int sg;
            const i64 s = scale_of(k, sg);
            if (!sg || s < low || s >= cutoff) continue;
            u32* dst = sg > 0 ? pos : neg;
            if (k == K) global_add(dst, W, c.w, N, s - low);
            else {
                const Num<N> a = vec_load<N>(x, k), b = vec_load<N>(y, k);
                u32 p[2 * N];
                mantissa_product<N>(a.w, b.w, p);
                global_add(dst, W, p, 2 * N, s - low);
            }
        }
        int cmp = 0;
        for (u64 w = W; w-- > 0;)
            if (pos[w] != neg[w]) {
                cmp = pos[w] > neg[w] ? 1 : -1;
                break;
            }
        if (cmp) {
            if (!nsums) {
                u32* big = cmp > 0 ? pos : neg;
                const u32* small = cmp > 0 ? neg : pos;
                u32 bw = 0;
                u64 top = 0;
                for (u64 w = 0; w < W; ++w) {
                    const u64 v = (u64)big[w] - small[w] - bw;
                    big[w] = (u32)v;
                    bw = (u32)(v >> 63);
                    if (big[w]) top = w;
                }
                const i64 first = (i64)top - (N + 2);
                bool sticky = false;
                for (i64 w = 0; w < first; ++w) sticky = sticky || big[w] != 0;
                S[0] = sticky ? 1u : 0u;
#pragma unroll
                for (int t = 0; t < N + 3; ++t) {
                    const i64 w = first + t;
                    S[t + 1] = (w >= 0 && w < (i64)W) ? big[w] : 0u;
                }
                scale1 = low + 32 * (first - 1);
                sign1 = cmp;
            } else
                sign2 = cmp;
            ++nsums;
        }
        cutoff = low;
    }
    if (!nsums) return zero<N>(ST_OK);
    if (nsums == 1) return pack<N, KW>(S, scale1, sign1, ST_OK);
    // S * 2^D -/+ 1 puts the sign of the remaining clusters below the rounding position of S
    constexpr int DW = N + 2, EW = KW + DW;
    u32 ext[EW];
#pragma unroll
    for (int w = 0; w < EW; ++w) ext[w] = w >= DW ? S[w - DW] : 0u;
    if (sign2 == sign1) ext[0] = 1;
    else {
        u32 bw = 1;
#pragma unroll
        for (int w = 0; w < EW; ++w) {
            const u64 v = (u64)ext[w] - bw;
            ext[w] = (u32)v;
            bw = (u32)(v >> 63);
        }
    }
    return pack<N, EW>(ext, scale1 - 32 * DW, sign1, ST_OK);
}

// ---------------------------------------------------------------- one exact dot per wavefront
// D(c; x, y; K) = RN(c - sum_{k<K} x_k y_k) with one rounding (the contract of the reference's
// exact_dot_add with subtract, linalg.hpp:268-318): zero with the OR of the statuses of c and of every
// x_k, y_k; c itself when no product is nonzero.
//
// One wavefront computes one D.  In chunks of 32 terms, lane l < 32 forms the exact product of term
// k = chunk + l and writes it into LDS row l as a dense W-word two's-complement window (the product shifted
// to its offset from the smallest term, negated for a subtracted product); lane w then sums window column
// w over the rows as an unsigned 64-bit column sum, independent loads with no carries between columns;
// the lanes then resolve the carries together (mod 2^(32W), which holds the signed total), negate a
// negative window and round once (wave_round).  A window of W = 2N + 24 words holds terms spread over up
// to ~700 bits; wider spreads take the cluster path on lane 0.  Every lane of the wave calls it with the
// same arguments; the result is valid in lane 0 (in every lane on the window path).
constexpr int WAVE = 64;
constexpr int DOT_ROWS = 32; // terms per chunk

// Row stride of the exact-dot rows: the W-word window plus one word, odd (W is even), so that the lanes
// storing one column of their rows hit distinct banks.
template <int N> constexpr int dot_row_stride() { return factor_window(N) + 1; }
template <int N> struct alignas(16) WaveLds {
    u32 row[DOT_ROWS + 1][dot_row_stride<N>()]; // row DOT_ROWS holds c
    u32 neg[factor_window(N) + 1];              // per word: negative terms whose sign fill starts there
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ long long wave_max(long long v)
{
#pragma unroll
    for (int m = 32; m; m >>= 1) {
        const long long o = __shfl_xor(v, m);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ long long wave_min(long long v)
{
#pragma unroll
    for (int m = 32; m; m >>= 1) {
        const long long o = __shfl_xor(v, m);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int m = 32; m; m >>= 1) v += __shfl_xor(v, m);
    return v;
}
__device__ __forceinline__ int wave_or(int v)
{
#pragma unroll
    for (int m = 32; m; m >>= 1) v |= __shfl_xor(v, m);
    return v;
}

// The value of lane l - 1 in lane l (0 in lane 0): a DPP wave shift right by one lane (wave_shr:1), at ALU
// latency where a shuffle goes through LDS.  Every lane of the wave must be active.
__device__ __forceinline__ u32 wave_from_previous_lane(u32 v) { return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false); }

// Maximum (MAX) or minimum of a 32-bit value over the 64 lanes of the wave, in every lane: DPP row shifts by
// 1, 2, 4 and 8 leave each row's reduction in its lane 15, the row broadcasts of lanes 15 and 31 carry it to
// lane 63, which is then read (ALU latency per step, where a cross-lane shuffle waits on LDS).  Every lane
// of the wave must be active.
template <bool MAX> __device__ __forceinline__ int wave_extreme32(int v)
{
    constexpr int ID = MAX ? (int)0x80000000 : 0x7fffffff;
    auto pick = [](int a, int b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x111, 0xF, 0xF, false)); // row_shr:1
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x112, 0xF, 0xF, false)); // row_shr:2
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x114, 0xF, 0xF, false)); // row_shr:4
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x118, 0xF, 0xF, false)); // row_shr:8
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x142, 0xA, 0xF, false)); // row_bcast:15 into rows 1, 3
    v = pick(v, __builtin_amdgcn_update_dpp(ID, v, 0x143, 0xC, 0xF, false)); // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// Row r = sg * src (S words) * 2^rel in the (cleared) rows of a wave, as the S+1 words of the shifted term
// at word offset o = rel / 32, two's complement within those words for a negative term; the rest of a
// negative term's W-word two's complement (ones from word o+S+1 up) is left out and counted in neg[o+S+1]
// instead, for the column sums to take off (sg in {-1, 0, 1}; rel >= 0 and the term fits: rel + 32S <= 32W - 1).
template <int N, int S> __device__ __forceinline__ void put_row(WaveLds<N>& L, int r, const u32 (&src)[S], i64 rel, int sg)
{
    constexpr int W = factor_window(N);
    static_assert(S + 1 <= W, "a term fits the window");
    // a term that fits has o + S <= W - 1 (its top word, q[S], is then inside the row); the clamp only
    // keeps the stores inside the row should that ever not hold
    const int o = min((int)(rel >> 5), W - 1 - S);
    const u32 b = (u32)(rel & 31);
    u32 q[S + 1];
#pragma unroll
    for (int t = 0; t <= S; ++t) {
        const u32 hi = t < S ? src[t] : 0u, lo = t > 0 ? src[t - 1] : 0u;
        q[t] = sg ? (b ? ((hi << b) | (lo >> (32u - b))) : hi) : 0u;
    }
    if (sg < 0) {
        bool nonzero = false;
        u32 c = 1;
#pragma unroll
        for (int t = 0; t <= S; ++t) {
            nonzero = nonzero || q[t] != 0;
            const u64 v = (u64)(u32)~q[t] + c;
            q[t] = (u32)v;
            c = (u32)(v >> 32);
        }
        if (nonzero && o + S + 1 < W) atomicAdd(&L.neg[o + S + 1], 1u);
    }
    u32* row = &L.row[r][o];
#pragma unroll
    for (int t = 0; t <= S; ++t) row[t] = q[t];
}

// Clear the rows and the sign-fill counts of a wave's exact-dot LDS (all lanes, 16-byte stores).
template <int N> __device__ __forceinline__ void clear_rows(WaveLds<N>& L, u32 lane)
{
    static_assert(sizeof(WaveLds<N>) % 4 == 0, "whole words");
    constexpr u32 words = sizeof(WaveLds<N>) / 4, quads = words / 4;
    uint4* q = reinterpret_cast<uint4*>(&L);
    for (u32 k = lane; k < quads; k += WAVE) q[k] = make_uint4(0u, 0u, 0u, 0u);
    u32* w = reinterpret_cast<u32*>(&L);
    for (u32 k = 4 * quads + lane; k < words; k += WAVE) w[k] = 0u;
}

// Carries of one 64-word group of a multiword sum, all lanes at once: lane l's word has a carry-out g
// (generate) or is all ones p (propagate); with A = G | P and B = G as bit vectors over the lanes
// (G and P disjoint), A + B + cin ripples exactly those carries, so bit l of (A + B + cin) ^ A ^ B is the
// carry into lane l's word.  Returns that bit; cin becomes the group's carry-out.
__device__ __forceinline__ u32 wave_carry_in(bool g, bool p, u32 lane, u64& cin)
{
    const u64 G = __ballot(g), A = G | __ballot(p);
    const u64 s1 = A + G, s2 = s1 + cin;
    const u64 c = s2 ^ A ^ G;
    cin = (s1 < A || s2 < s1) ? 1u : 0u;
    return (u32)(c >> lane) & 1u;
}

// The rounded value of a W-word two's-complement window whose column w (w = lane + 64t, t < CPL) holds
// the 64-bit sum S[t] of that word over the rows, less H[t] * 2^(32w) for the sign fills left out of the
// rows (ones from word w up, mod 2^(32W): -2^(32w) each): RN-even of the window's value * 2^scale, in every
// lane.  Lane-parallel: the column sums are split into words plus a carry word into the next column,
// single-bit carries are resolved across lanes by wave_carry_in, a negative total is negated the same
// way, and the top word, the N mantissa words, the round bit and the sticky bit are found with ballots
// and lane reads (the serial form is pack() of the resolved window).
template <int N>
__device__ __forceinline__ Num<N> wave_round(const u64 (&S)[(factor_window(N) + WAVE - 1) / WAVE], const u32 (&H)[(factor_window(N) + WAVE - 1) / WAVE],
                                          i64 scale, u32 lane)
{
    constexpr int W = factor_window(N);
    constexpr int CPL = (W + WAVE - 1) / WAVE;
    u32 word[CPL];
    {
        u64 cin = 0;
        u32 hlast = 0; // carry word of column 64t - 1 (lane 63 of the previous group)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const bool in = (int)lane + WAVE * t < W;
            const u32 h = (u32)(S[t] >> 32); // at most the row count: a single carry bit into the next word
            u32 hprev = wave_from_previous_lane(h);
            hprev = lane == 0 ? hlast : hprev;
            hlast = __builtin_amdgcn_readlane(h, WAVE - 1);
            const u64 ts = (u64)(u32)S[t] + hprev;
            const u32 u = in ? (u32)ts : 0u;
            const u32 c = wave_carry_in(in && (ts >> 32) != 0, in && u == 0xffffffffu, lane, cin);
            word[t] = in ? u + c : 0u;
        }
    }
    {   // minus the sign fills: H[t] ones-words start at column w (a borrow lookahead like the carries)
        u64 bin = 0;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const bool in = (int)lane + WAVE * t < W;
            const u32 d = word[t] - H[t];
            const u32 bb = wave_carry_in(in && word[t] < H[t], in && d == 0u, lane, bin);
            word[t] = in ? d - bb : 0u;
        }
    }
    constexpr int TL = (W - 1) % WAVE, TG = (W - 1) / WAVE;
    const int sign = (__builtin_amdgcn_readlane(word[TG], TL) >> 31) ? -1 : 1;
    if (sign < 0) {
        u64 cin = 1;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const bool in = (int)lane + WAVE * t < W;
            const u32 v = in ? ~word[t] : 0u;
            const u32 c = wave_carry_in(false, in && v == 0xffffffffu, lane, cin);
            word[t] = in ? v + c : 0u;
        }
    }
    // the top nonzero word
    int top = -1;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const u64 nz = __ballot(word[t] != 0);
        if (nz) top = WAVE * t + 63 - __clzll(nz);
    }
    if (top < 0) return zero<N>(ST_OK);
    u32 topword = 0;
#pragma unroll
    for (int t = 0; t < CPL; ++t)
        if (top / WAVE == t) topword = __builtin_amdgcn_readlane(word[t], top % WAVE);
    const int lz = __clz(topword);
    // normalized words: the window shifted left by lz bits
    u32 nw[CPL];
    {
        u32 last = 0;
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            u32 prev = wave_from_previous_lane(word[t]);
            prev = lane == 0 ? last : prev;
            last = __builtin_amdgcn_readlane(word[t], WAVE - 1);
            nw[t] = lz ? (word[t] << lz) | (prev >> (32 - lz)) : word[t];
        }
    }
    const int rj = top - N; // the word holding the round bit
    bool sticky = false;
#pragma unroll
    for (int t = 0; t < CPL; ++t) sticky = sticky || __ballot(nw[t] != 0 && (int)lane + WAVE * t < rj) != 0;
    auto read = [&](int j) -> u32 {
        u32 v = 0;
#pragma unroll
        for (int t = 0; t < CPL; ++t)
            if (j >= 0 && j / WAVE == t) v = __builtin_amdgcn_readlane(nw[t], j % WAVE);
        return v;
    };
    const u32 rw = read(rj);
    sticky = sticky || (rw << 1) != 0;
    u32 m[N];
#pragma unroll
    for (int k = 0; k < N; ++k) m[k] = read(rj + 1 + k);
    return round_words<N>(m, rw >> 31, sticky, scale + 32 * (i64)top + 31 - lz, sign, ST_OK);
}

template <int N>
__device__ Num<N> wave_dot_sub(const Num<N>& c, const VecView& x, const VecView& y, u32 K, u32 lane, WaveLds<N>& L, const DotScratch& sc)
{
    constexpr int W = factor_window(N);
    constexpr int CPL = (W + WAVE - 1) / WAVE; // window columns per lane
    constexpr i64 NONE = -((i64)1 << 62);
    // pass 1: statuses, scale range and count of the nonzero products.  With one chunk of terms (K <= 32)
    // each lane loads its term once, here, and forms its exact product at once (it depends only on the
    // loads, so it overlaps the cross-lane reductions that follow); the status OR and the count come from
    // ballots and the scale range from DPP reductions over 32-bit scales (valid exponents keep them in
    // range; a lane whose scale is not takes the 64-bit shuffle reductions).
    const bool one_chunk = K <= (u32)DOT_ROWS;
    Num<N> ta = zero<N>(ST_OK), tb = zero<N>(ST_OK);
    u32 pt[2 * N];
    u32 st = 0;
    i64 hi = NONE, lo = -NONE;
    int count = 0;
    if (one_chunk) {
        if (lane < K) {
            ta = vec_load<N>(x, lane);
            tb = vec_load<N>(y, lane);
        }
        mantissa_product<N>(ta.w, tb.w, pt);
        const u32 stl = ta.st | tb.st;
        const bool nz = ta.s && tb.s;
        const i64 s64 = nz ? (i64)ta.e + tb.e - 2 * (32 * N - 1) : 0;
        st = __ballot(stl != 0) ? (u32)wave_or((int)stl) : 0u;
        count = __popcll(__ballot(nz));
        if (__ballot(nz && (s64 < -((i64)1 << 30) || s64 > ((i64)1 << 30)))) {
            hi = wave_max(nz ? s64 : NONE);
            lo = wave_min(nz ? s64 : -NONE);
        } else if (count) {
            hi = wave_extreme32<true>(nz ? (int)s64 : (int)0x80000000);
            lo = wave_extreme32<false>(nz ? (int)s64 : 0x7fffffff);
        }
    } else {
        for (u32 k = lane; k < K; k += WAVE) {
            int ex, sx, ey, sy;
            u32 tx, ty;
            vec_meta<N>(x, k, ex, sx, tx);
            vec_meta<N>(y, k, ey, sy, ty);
            st |= tx | ty;
            if (sx && sy) {
                const i64 s = (i64)ex + ey - 2 * (32 * N - 1);
                hi = s > hi ? s : hi;
                lo = s < lo ? s : lo;
                ++count;
            }
        }
        st = (u32)wave_or((int)st);
        count = wave_sum(count);
        hi = wave_max(hi);
        lo = wave_min(lo);
    }
    st |= c.st;
    if (st) return zero<N>(st);
    if (!count) return c;
    // the window spans from the lowest term's last bit to the highest top bit: a product's 64N bits above
    // its scale, c's 32N above its own (c, a 32N-bit value, sits about 32N bits above products of its
    // magnitude, so charging it a product's width would send ordinary dots past about 21 limbs to the
    // cluster path)
    i64 top = hi + 64 * N;
    if (c.s) {
        const i64 s = (i64)c.e - (32 * N - 1);
        top = s + 32 * N > top ? s + 32 * N : top;
        lo = s < lo ? s : lo;
        ++count;
    }
    int logc = 0;
    while ((1 << logc) < count) ++logc;
    if (top - lo + logc + 1 > 32 * W) {
        Num<N> r = zero<N>(ST_OK);
        if (lane == 0) r = dot_sub_clusters<N>(c, x, y, K, sc);
        return r;
    }
    // pass 2: rows of exact products, column sums
    u64 S[CPL];
    u32 H[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        S[t] = 0;
        H[t] = 0;
    }
    for (u32 kb = 0; kb < K; kb += DOT_ROWS) {
        const u32 rows = K - kb < (u32)DOT_ROWS ? K - kb : (u32)DOT_ROWS;
        clear_rows<N>(L, lane);
        wave_sync();
        if (lane < rows) {
            if (one_chunk) put_row<N, 2 * N>(L, (int)lane, pt, (ta.s && tb.s) ? (i64)ta.e + tb.e - 2 * (32 * N - 1) - lo : 0, -ta.s * tb.s);
            else {
                const u32 k = kb + lane;
                const Num<N> a = vec_load<N>(x, k), b = vec_load<N>(y, k);
                u32 p[2 * N];
                mantissa_product<N>(a.w, b.w, p);
                put_row<N, 2 * N>(L, (int)lane, p, (a.s && b.s) ? (i64)a.e + b.e - 2 * (32 * N - 1) - lo : 0, -a.s * b.s);
            }
        }
        const bool with_c = kb == 0 && c.s != 0;
        if (with_c && lane == WAVE - 1) put_row<N, N>(L, DOT_ROWS, c.w, (i64)c.e - (32 * N - 1) - lo, c.s);
        wave_sync();
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int w = (int)lane + WAVE * t;
            if (w < W) {
                u64 sum = with_c ? (u64)L.row[DOT_ROWS][w] : 0u;
#pragma unroll 8
                for (u32 r = 0; r < rows; ++r) sum += L.row[r][w];
                S[t] += sum;
                H[t] += L.neg[w];
            }
        }
        wave_sync();
    }
    const Num<N> r = wave_round<N>(S, H, lo, lane);
    wave_sync(); // the rows and columns are reused by the wave's next dot
    return r;
}

// ---------------------------------------------------------------- Cholesky
template <int N> __device__ __forceinline__ Num<N> element(const FactorArgs& f, u32 i, u32 j)
{
    return load_out<N>(f.w, f.pitch, (u64)i * f.n + j);
}

// The reciprocal (limb_core.hpp reciprocal) of the pivot or diagonal entry i, N+1 planes.
template <int N> __device__ __forceinline__ void load_recip(const u32* r, u64 rpitch, u32 i, u32 (&v)[N + 1])
{
#pragma unroll
    for (int k = 0; k <= N; ++k) v[k] = r[(u64)k * rpitch + i];
}
template <int N> __device__ __forceinline__ void store_recip(u32* r, u64 rpitch, u32 i, const Num<N>& d)
{
    u32 v[N + 1];
    if (d.s && !d.st) reciprocal<N>(d.w, v);
    else
#pragma unroll
        for (int k = 0; k <= N; ++k) v[k] = 0; // never used: division by zero or a status is decided first
#pragma unroll
    for (int k = 0; k <= N; ++k) r[(u64)k * rpitch + i] = v[k];
}

// Waves of the diagonal-block workgroup: 12 up to 4 limbs (three waves per SIMD leave 168 VGPRs a lane,
// enough there), 8 above (at 168 VGPRs the kernel spilled from 5 limbs on, 350 bytes a lane at 8 limbs;
// at 256 bits 8 waves without those spills ran the factorization's blocks 5% faster than 12, with four
// numerator waves fewer); fewer when their LDS rows would pass 144 KB.
template <int N> constexpr int diag_waves()
{
    constexpr int by_regs = N <= 4 ? 12 : 8;
    constexpr int by_lds = (int)(147456 / sizeof(WaveLds<N>));
    return by_lds < 1 ? 1 : (by_lds < by_regs ? by_lds : by_regs);
}

// The diagonal block [k0, k1) of a panel, column by column, in one workgroup of diag_waves<N>()
// wavefronts.  Column j in two phases: (A) wave 0 forms the pivot's exact dot and its square root
// (lane 0) while the other waves form the numerators D(a_ij; l_i, l_j; [k0, j)) of the rows below, one
// row per wave at a time, written in place (they read only columns before j); (B) one thread per row
// divides its numerator by l_jj.  Stops at a failing pivot, before any entry of its column is written
// (numerators of the rows below included: phase B is never reached).
template <int N> __global__ void __launch_bounds__(diag_waves<N>() * WAVE) chol_diag_kernel(FactorArgs f)
{
    constexpr int DW = diag_waves<N>();
    __shared__ WaveLds<N> lds[DW];
    __shared__ int stop;
    const u32 wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    WaveLds<N>& L = lds[wave];
    const DotScratch sc{f.scratch + (u64)wave * 2 * f.swords, f.swords, &f.info->error};
    for (u32 j = f.k0; j < f.k1; ++j) {
        const VecView rj{f.w, f.pitch, (i64)j * f.n + f.k0, 1};
        if (wave == 0) {
            const Num<N> s = wave_dot_sub<N>(element<N>(f, j, j), rj, rj, j - f.k0, lane, L, sc);
            if (lane == 0) {
                if (s.st || s.s <= 0) {
                    f.info->pivot = (int32_t)j;
                    f.info->sign = s.st ? 0 : s.s;
                    f.info->status = s.st;
                    stop = 1;
                } else {
                    store_out<N>(f.w, f.pitch, f.n, j, j, false, sqrt<N>(s));
                    stop = 0;
                }
            }
        }
        if (DW == 1 || wave > 0) {
            const u32 step = DW > 1 ? DW - 1 : 1;
            for (u32 i = j + 1 + (DW > 1 ? wave - 1 : 0); i < f.k1; i += step) {
                const VecView ri{f.w, f.pitch, (i64)i * f.n + f.k0, 1};
                const Num<N> v = wave_dot_sub<N>(element<N>(f, i, j), ri, rj, j - f.k0, lane, L, sc);
                if (lane == 0) store_out<N>(f.w, f.pitch, f.n, i, j, false, v);
            }
        }
        __syncthreads();
        if (stop) return;
        const Num<N> piv = element<N>(f, j, j);
        if (threadIdx.x == blockDim.x - 1) store_recip<N>(f.recip, f.rpitch, j, piv); // for the rows below the block
        for (u32 i = j + 1 + threadIdx.x; i < f.k1; i += blockDim.x - 1)
            store_out<N>(f.w, f.pitch, f.n, i, j, false, div<N>(element<N>(f, i, j), piv));
        __syncthreads();
    }
}

// The rows below the diagonal block, one wavefront per row i >= k1: columns [k0, p) in order, p the
// failing pivot of the block or k1.
template <int N> __global__ void __launch_bounds__(64) chol_panel_kernel(FactorArgs f)
{
    __shared__ WaveLds<N> L;
    const u32 r = blockIdx.x, lane = threadIdx.x;
    const u32 i = f.k1 + r;
    const int32_t piv = f.info->pivot;
    const u32 pe = piv >= 0 ? (u32)piv : f.k1;
    const DotScratch sc{f.scratch + (u64)r * 2 * f.swords, f.swords, &f.info->error};
    const VecView ri{f.w, f.pitch, (i64)i * f.n + f.k0, 1};
    for (u32 j = f.k0; j < pe; ++j) {
        const VecView rj{f.w, f.pitch, (i64)j * f.n + f.k0, 1};
        const Num<N> v = wave_dot_sub<N>(element<N>(f, i, j), ri, rj, j - f.k0, lane, L, sc);
        if (lane == 0) {
            u32 rv[N + 1];
            load_recip<N>(f.recip, f.rpitch, j, rv);
            store_out<N>(f.w, f.pitch, f.n, i, j, false, div_by_reciprocal<N>(v, element<N>(f, j, j), rv));
        }
        __syncthreads(); // entry (i, j) is a term of the row's next dot
    }
}

// Look-ahead part of a block's trailing update: after the panel [k0, k1), the entries (i, j), j <= i, of the
// next panel's columns j in [k1, k2) and rows i in [k1, n) become D(a_ij; l_i, l_j; [k0, k1)) in place -- the
// same one rounding of the same exact value as the update product gives them -- so that the next panel can
// start while the rest of the update (rows and columns from k2) runs on another stream.  One wavefront per
// entry, grid-stride over the (n - k1) x (k2 - k1) rectangle; wavefront b uses scratch slot b.
template <int N> __global__ void __launch_bounds__(64) chol_lookahead_kernel(FactorArgs f, u32 k2)
{
    __shared__ WaveLds<N> L;
    const u32 lane = threadIdx.x, cols = k2 - f.k1;
    const u64 entries = (u64)(f.n - f.k1) * cols;
    const DotScratch sc{f.scratch + (u64)blockIdx.x * 2 * f.swords, f.swords, &f.info->error};
    for (u64 e = blockIdx.x; e < entries; e += gridDim.x) {
        const u32 i = f.k1 + (u32)(e / cols), j = f.k1 + (u32)(e % cols);
        if (j > i) continue; // the whole wavefront skips the strict upper part
        const VecView ri{f.w, f.pitch, (i64)i * f.n + f.k0, 1};
        const VecView rj{f.w, f.pitch, (i64)j * f.n + f.k0, 1};
        const Num<N> v = wave_dot_sub<N>(element<N>(f, i, j), ri, rj, f.k1 - f.k0, lane, L, sc);
        if (lane == 0) store_out<N>(f.w, f.pitch, f.n, i, j, false, v);
    }
}

// Entries (i, j) with i >= j >= p after a failing pivot p: zero with status invalid.  Thread per entry of
// the trailing (n-p) x (n-p) square.
template <int N> __global__ void __launch_bounds__(256) chol_invalid_kernel(u32* w, u64 pitch, u32 n, u32 p)
{
    const u64 m = n - p, idx = (u64)blockIdx.x * 256u + threadIdx.x;
    if (idx >= m * m) return;
    const u32 i = p + (u32)(idx / m), j = p + (u32)(idx % m);
    if (j <= i) store_out<N>(w, pitch, n, i, j, false, zero<N>(ST_INVALID));
}

// The strict upper triangle of the work matrix set to zero (the factor's upper part).
template <int N> __global__ void __launch_bounds__(256) zero_upper_kernel(u32* w, u64 pitch, u32 n)
{
    const u64 idx = (u64)blockIdx.x * 256u + threadIdx.x;
    if (idx >= (u64)n * n) return;
    const u32 i = (u32)(idx / n), j = (u32)(idx % n);
    if (j > i) store_out<N>(w, pitch, n, i, j, false, zero<N>(ST_OK));
}

// ---------------------------------------------------------------- triangular solves
// Forward block [r0, r1) of L X = B, one wavefront per right-hand side.
template <int N> __global__ void __launch_bounds__(64) solve_forward_kernel(SolveArgs a)
{
    __shared__ WaveLds<N> L;
    const u32 c = blockIdx.x, lane = threadIdx.x;
    const DotScratch sc{a.scratch + (u64)c * 2 * a.swords, a.swords, a.error};
    const VecView xc{a.x, a.xpitch, (i64)a.r0 * a.nrhs + c, (i64)a.nrhs};
    for (u32 i = a.r0; i < a.r1; ++i) {
        const VecView li{a.f, a.fpitch, (i64)i * a.n + a.r0, 1};
        const Num<N> v = wave_dot_sub<N>(load_out<N>(a.x, a.xpitch, (u64)i * a.nrhs + c), li, xc, i - a.r0, lane, L, sc);
        if (lane == 0) {
            u32 rv[N + 1];
            load_recip<N>(a.recip, a.rpitch, i, rv);
            store_out<N>(a.x, a.xpitch, a.nrhs, i, c, false, div_by_reciprocal<N>(v, load_out<N>(a.f, a.fpitch, (u64)i * a.n + i), rv));
        }
        __syncthreads(); // x_i is a term of the next row's dot
    }
}

// Backward block [r0, r1) of L^T X = B (column i of L as the vector), one wavefront per right-hand side.
template <int N> __global__ void __launch_bounds__(64) solve_backward_kernel(SolveArgs a)
{
    __shared__ WaveLds<N> L;
    const u32 c = blockIdx.x, lane = threadIdx.x;
    const DotScratch sc{a.scratch + (u64)c * 2 * a.swords, a.swords, a.error};
    for (u32 i = a.r1; i-- > a.r0;) {
        Num<N> v = load_out<N>(a.x, a.xpitch, (u64)i * a.nrhs + c);
        if (i + 1 < a.r1) {
            const VecView col{a.f, a.fpitch, (i64)(i + 1) * a.n + i, (i64)a.n};
            const VecView xc{a.x, a.xpitch, (i64)(i + 1) * a.nrhs + c, (i64)a.nrhs};
            v = wave_dot_sub<N>(v, col, xc, a.r1 - i - 1, lane, L, sc);
        }
        if (lane == 0) {
            u32 rv[N + 1];
            load_recip<N>(a.recip, a.rpitch, i, rv);
            store_out<N>(a.x, a.xpitch, a.nrhs, i, c, false, div_by_reciprocal<N>(v, load_out<N>(a.f, a.fpitch, (u64)i * a.n + i), rv));
        }
        __syncthreads();
    }
}

// ---- in-block solves with the block staged in LDS
// The in-block solves run one wavefront per right-hand side, a few waves on the whole GPU, so each row's
// global loads (its terms, b_i, the reciprocal, l_ii, and the x_k just written) stand on the chain's
// critical path.  These forms copy the block's part of L (the lower triangle, packed so that each row's
// terms are contiguous: by rows for the forward pass, by columns for the backward pass), its x column and
// its reciprocals into LDS once, then read only LDS in the recurrence; x_i goes to LDS for the later rows
// and to global memory as the result.  Same arithmetic and order as the kernels above.
constexpr u32 STAGED_ROWS = 32; // largest block the staged forms take

// LDS words of the staged block of B rows: the packed triangle (N+3 planes of B(B+1)/2 entries), the x
// column (N+3 planes of B entries) and the reciprocals (N+1 planes of B entries).
template <int N> __host__ __device__ constexpr u32 staged_tri_pitch(u32 B) { return B * (B + 1) / 2; }
template <int N> __host__ __device__ constexpr u64 staged_words(u32 B) { return (u64)(N + 3) * staged_tri_pitch<N>(B) + (u64)(N + 3) * B + (u64)(N + 1) * B; }

// first index of column i of a column-packed lower triangle of B rows: sum over c < i of (B - c)
__device__ __forceinline__ u32 tri_col_start(u32 B, u32 i) { return i * B - i * (i - 1) / 2; }

template <int N> __device__ __forceinline__ void stage_copy(u32* dst, u64 dpitch, u64 di, const u32* src, u64 spitch, u64 si)
{
#pragma unroll
    for (int w = 0; w < N + 3; ++w) dst[(u64)w * dpitch + di] = src[(u64)w * spitch + si];
}

// Copies the block's L triangle (by rows or by columns), x column and reciprocals into LDS.
template <int N>
__device__ __forceinline__ void stage_solve_block(const SolveArgs& a, u32 c, bool by_columns, u32* tri, u32* xs, u32* rs, u32 lane)
{
    const u32 B = a.r1 - a.r0, TP = staged_tri_pitch<N>(B);
    for (u32 i = 0; i < B; ++i) {
        if (by_columns) // column i: rows k = i .. B-1
            for (u32 k = i + lane; k < B; k += WAVE)
                stage_copy<N>(tri, TP, tri_col_start(B, i) + (k - i), a.f, a.fpitch, (u64)(a.r0 + k) * a.n + (a.r0 + i));
        else // row i: columns k = 0 .. i
            for (u32 k = lane; k <= i; k += WAVE)
                stage_copy<N>(tri, TP, i * (i + 1) / 2 + k, a.f, a.fpitch, (u64)(a.r0 + i) * a.n + (a.r0 + k));
    }
    for (u32 i = lane; i < B; i += WAVE) {
        stage_copy<N>(xs, B, i, a.x, a.xpitch, (u64)(a.r0 + i) * a.nrhs + c);
#pragma unroll
        for (int w = 0; w <= N; ++w) rs[(u64)w * B + i] = a.recip[(u64)w * a.rpitch + a.r0 + i];
    }
}

template <int N> __device__ __forceinline__ void load_recip_lds(const u32* rs, u32 B, u32 i, u32 (&v)[N + 1])
{
#pragma unroll
    for (int k = 0; k <= N; ++k) v[k] = rs[(u64)k * B + i];
}

// Forward block [r0, r1) of L X = B, r1 - r0 <= STAGED_ROWS, block staged in LDS (dynamic shared memory of
// staged_words<N>(r1 - r0) words); one wavefront per right-hand side.
template <int N> __global__ void __launch_bounds__(64) solve_forward_staged_kernel(SolveArgs a)
{
    __shared__ WaveLds<N> L;
    extern __shared__ u32 stage[];
    const u32 c = blockIdx.x, lane = threadIdx.x, B = a.r1 - a.r0, TP = staged_tri_pitch<N>(B);
    u32* tri = stage;
    u32* xs = tri + (u64)(N + 3) * TP;
    u32* rs = xs + (u64)(N + 3) * B;
    stage_solve_block<N>(a, c, false, tri, xs, rs, lane);
    __syncthreads();
    const DotScratch sc{a.scratch + (u64)c * 2 * a.swords, a.swords, a.error};
    const VecView xv{xs, B, 0, 1};
    for (u32 i = 0; i < B; ++i) {
        const VecView li{tri, TP, (i64)(i * (i + 1) / 2), 1};
        const Num<N> v = wave_dot_sub<N>(load_out<N>(xs, B, i), li, xv, i, lane, L, sc);
        if (lane == 0) {
            u32 rv[N + 1];
            load_recip_lds<N>(rs, B, i, rv);
            const Num<N> x = div_by_reciprocal<N>(v, load_out<N>(tri, TP, i * (i + 1) / 2 + i), rv);
            store_out<N>(xs, B, 1, i, 0, false, x);
            store_out<N>(a.x, a.xpitch, a.nrhs, a.r0 + i, c, false, x);
        }
        __syncthreads(); // x_i is a term of the next row's dot
    }
}

// Backward block [r0, r1) of L^T X = B, staged as the forward form (the triangle packed by columns).
template <int N> __global__ void __launch_bounds__(64) solve_backward_staged_kernel(SolveArgs a)
{
    __shared__ WaveLds<N> L;
    extern __shared__ u32 stage[];
    const u32 c = blockIdx.x, lane = threadIdx.x, B = a.r1 - a.r0, TP = staged_tri_pitch<N>(B);
    u32* tri = stage;
    u32* xs = tri + (u64)(N + 3) * TP;
    u32* rs = xs + (u64)(N + 3) * B;
    stage_solve_block<N>(a, c, true, tri, xs, rs, lane);
    __syncthreads();
    const DotScratch sc{a.scratch + (u64)c * 2 * a.swords, a.swords, a.error};
    for (u32 i = B; i-- > 0;) {
        Num<N> v = load_out<N>(xs, B, i);
        if (i + 1 < B) {
            const VecView col{tri, TP, (i64)tri_col_start(B, i) + 1, 1};
            const VecView xv{xs, B, (i64)i + 1, 1};
            v = wave_dot_sub<N>(v, col, xv, B - i - 1, lane, L, sc);
        }
        if (lane == 0) {
            u32 rv[N + 1];
            load_recip_lds<N>(rs, B, i, rv);
            const Num<N> x = div_by_reciprocal<N>(v, load_out<N>(tri, TP, tri_col_start(B, i)), rv);
            store_out<N>(xs, B, 1, i, 0, false, x);
            store_out<N>(a.x, a.xpitch, a.nrhs, a.r0 + i, c, false, x);
        }
        __syncthreads();
    }
}

// Off-block update of a solve by the block [r0, r1), for updates too small to pay for the update GEMM's
// planning: x_ic = D(x_ic; l, x_c; [r0, r1)) with l = row i of L (forward, rows i in [r1, n)) or column i of L
// (backward, rows i in [0, r0)), one wavefront per entry, grid-stride over the rows x nrhs entries; wavefront
// b uses scratch slot b.  The same one rounding of the same exact value as the update product.
template <int N> __global__ void __launch_bounds__(64) solve_update_kernel(SolveArgs a, u32 backward)
{
    __shared__ WaveLds<N> L;
    const u32 lane = threadIdx.x;
    const u32 row0 = backward ? 0u : a.r1, rows = backward ? a.r0 : a.n - a.r1;
    const u64 entries = (u64)rows * a.nrhs;
    const DotScratch sc{a.scratch + (u64)blockIdx.x * 2 * a.swords, a.swords, a.error};
    for (u64 e = blockIdx.x; e < entries; e += gridDim.x) {
        const u32 i = row0 + (u32)(e / a.nrhs), c = (u32)(e % a.nrhs);
        const VecView l = backward ? VecView{a.f, a.fpitch, (i64)a.r0 * a.n + i, (i64)a.n} : VecView{a.f, a.fpitch, (i64)i * a.n + a.r0, 1};
        const VecView xc{a.x, a.xpitch, (i64)a.r0 * a.nrhs + c, (i64)a.nrhs};
        const Num<N> v = wave_dot_sub<N>(load_out<N>(a.x, a.xpitch, (u64)i * a.nrhs + c), l, xc, a.r1 - a.r0, lane, L, sc);
        if (lane == 0) store_out<N>(a.x, a.xpitch, a.nrhs, i, c, false, v);
    }
}

// Measurement support: cycles (s_memtime) of one of the factorization's per-row steps on one wavefront
// alone, averaged over `reps` dependent repetitions, one launch per step so that each kernel holds only
// the code it times: STEP 0 exact dots of 16 then 31 terms (out[0], out[1]), 1 a division by a reciprocal
// (out[2]), 2 a Knuth division (out[3]), 3 a square root (out[4]), 4 a reciprocal (out[5]).  v holds 67
// normalized values in N+3 planes of pitch `pitch` (terms 0-30 and 32-62, addend 64, divisor 65, dividend
// 66); sink keeps the results live.
template <int N, int STEP>
__global__ void __launch_bounds__(64) factor_steps_kernel(const u32* v, u64 pitch, u32* scratch, u64 swords, int32_t* err, int reps, u64* out,
                                                         u32* sink)
{
    __shared__ WaveLds<N> L;
    const u32 lane = threadIdx.x;
    u32 chk = 0;
    if constexpr (STEP == 0) {
        const DotScratch sc{scratch, swords, err};
        const VecView x{v, pitch, 0, 1}, y{v, pitch, 32, 1};
        Num<N> acc = load_out<N>(v, pitch, 64);
        for (int t = 0; t < 2; ++t) {
            const u32 K = t ? 31u : 16u;
            const u64 t0 = __builtin_amdgcn_s_memtime();
            for (int r = 0; r < reps; ++r) {
                const Num<N> d = wave_dot_sub<N>(acc, x, y, K, lane, L, sc);
                acc.w[0] ^= d.w[0] & 1u; // each dot depends on the one before
                chk += d.w[N - 1];
            }
            const u64 t1 = __builtin_amdgcn_s_memtime();
            if (lane == 0) out[t] = (t1 - t0) / (u64)reps;
        }
    } else if (lane == 0) {
        const Num<N> d = load_out<N>(v, pitch, 65);
        Num<N> a = load_out<N>(v, pitch, 66);
        a.s = 1;
        u32 rv[N + 1];
        if constexpr (STEP == 1) reciprocal<N>(d.w, rv);
        const u64 t0 = __builtin_amdgcn_s_memtime();
        for (int r = 0; r < reps; ++r) {
            if constexpr (STEP == 4) {
                u32 w[N + 1];
                reciprocal<N>(a.w, w);
#pragma unroll
                for (int k = 0; k < N; ++k) a.w[k] ^= w[k] & 1u; // every word of the reciprocal feeds the next
                chk += w[0];
            } else {
                const Num<N> q = STEP == 1 ? div_by_reciprocal<N>(a, d, rv) : STEP == 2 ? div<N>(a, d) : sqrt<N>(a);
#pragma unroll
                for (int k = 0; k < N - 1; ++k) a.w[k] ^= q.w[k] & 1u;
                chk += q.w[N - 1];
            }
        }
        const u64 t1 = __builtin_amdgcn_s_memtime();
        out[STEP + 1] = (t1 - t0) / (u64)reps;
    }
    sink[lane] = chk;
}

// Reciprocals of the n diagonal entries of the factor, one thread each (for the solves' divisions).
template <int N> __global__ void __launch_bounds__(64) reciprocals_kernel(const u32* w, u64 pitch, u32 n, u32* recip, u64 rpitch)
{
    const u32 i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    store_recip<N>(recip, rpitch, i, load_out<N>(w, pitch, (u64)i * n + i));
}

} // namespace lf

// Host build of the device arithmetic core (limbforge_amd/csrc/limb_core.hpp) checked against the
// oracle on the reference tests' case classes, for every width.  Test infrastructure: compiled by
// tests/test_device_core_cpu.py with LF_HOST_BUILD defined.
#include "dot2_core.hpp"
#include "limb_core.hpp"
#include "residue_reduce.hpp"
#include "coop_core.hpp"
#include "lf_oracle.h"
#include "lfo_cases.h"
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <gmp.h>
#include <vector>

static long failures = 0, checks = 0;

template <int N> lf::Num<N> to_dev(const lfo_num& x)
{
    lf::Num<N> r;
    for (int i = 0; i < N; ++i) r.w[i] = x.limb[i];
    r.e = x.exponent;
    r.s = x.sign;
    r.st = x.status;
    return r;
}

template <int N> bool same(const lf::Num<N>& d, const lfo_num& o)
{
    if (d.s != o.sign || d.e != o.exponent || d.st != o.status) return false;
    for (int i = 0; i < N; ++i)
        if (d.w[i] != o.limb[i]) return false;
    return true;
}

template <int N> void report(const char* op, int i, const lf::Num<N>& d, const lfo_num& o)
{
    ++checks;
    if (same<N>(d, o)) return;
    if (++failures <= 30)
        std::printf("MISMATCH %s bits=%d case=%d dev(s=%d e=%d st=%u top=%08x lo=%08x) oracle(s=%d e=%d st=%u top=%08x lo=%08x)\n", op,
                    32 * N, i, d.s, d.e, d.st, d.w[N - 1], d.w[0], o.sign, o.exponent, o.status, o.limb[N - 1], o.limb[0]);
}

template <int N> void run_width(int cases)
{
    lfo_mt64 g;
    lfo_mt_seed(&g, 4242 + N);
    for (int i = 0; i < cases; ++i) {
        lfo_num a, b, c, root;
        lfo_binary_case(&g, N, i, &a, &b);
        auto da = to_dev<N>(a), db = to_dev<N>(b);
        report<N>("add", i, lf::add<N>(da, db), lfo_add(N, a, b));
        report<N>("sub", i, lf::sub<N>(da, db), lfo_sub(N, a, b));
        report<N>("mul", i, lf::mul<N>(da, db), lfo_mul(N, a, b));
        if (a.sign && b.sign) { // structured mantissas: the short product's undecided cases (runs of zeros
                                // or ones next to the round bit) must fall back to the exact product
            lfo_num s = a, t = b;
            const int za = (int)(lfo_mt_next(&g) % N), zb = (int)(lfo_mt_next(&g) % N);
            const uint32_t fill[4] = {0u, 0xffffffffu, 0u, 0x80000000u};
            for (int w = 0; w < za; ++w) s.limb[w] = fill[i % 4];
            for (int w = 0; w < zb; ++w) t.limb[w] = fill[(i / 4) % 4];
            if (i % 5 == 1) for (int w = 0; w < N; ++w) s.limb[w] = 0xffffffffu;
            if (i % 7 == 2) { for (int w = 0; w < N; ++w) t.limb[w] = 0; t.limb[N - 1] = 0x80000000u; }
            report<N>("mul structured", i, lf::mul<N>(to_dev<N>(s), to_dev<N>(t)), lfo_mul(N, s, t));
            if constexpr (lf::use_digits<N>()) {
                uint32_t td[lf::digit_count<N>()];
                lf::to_digits<N>(to_dev<N>(t).w, td);
                report<N>("mul structured digits", i, lf::mul_by_digits<N>(to_dev<N>(s), to_dev<N>(t), td), lfo_mul(N, s, t));
            }
        }
        if constexpr (lf::use_digits<N>()) { // the chain's multiply with the factor's digits formed once
            uint32_t bd[lf::digit_count<N>()];
            lf::to_digits<N>(db.w, bd);
            auto x = da;
            lfo_num y = a;
            for (int s = 0; s < 3; ++s) {
                x = lf::mul_by_digits<N>(x, db, bd);
                y = lfo_mul(N, y, b);
            }
            report<N>("mul chain", i, x, y);
        }
        {   // a mixed chain as the batching kernels run it: mul steps are fma steps with c = 0
            lfo_num g0 = lfo_random_number(&g, N, 64);
            if (i % 7 == 3) g0 = lfo_zero(0);
            if (i % 11 == 5) g0.status = 4;
            const uint64_t mask = lfo_mt_next(&g);
            auto x = da;
            const auto dg = to_dev<N>(g0);
            lfo_num y = a;
            for (int s = 0; s < 6; ++s) {
                const bool f_step = (mask >> s) & 1u;
                if constexpr (lf::use_digits<N>()) {
                    uint32_t bd[lf::digit_count<N>()];
                    lf::to_digits<N>(db.w, bd);
                    x = lf::fma_by_digits<N>(x, db, bd, f_step ? dg : lf::zero<N>(0));
                } else
                    x = lf::fma<N>(x, db, f_step ? dg : lf::zero<N>(0));
                y = f_step ? lfo_fma(N, y, b, g0) : lfo_mul(N, y, b);
            }
            report<N>("mixed chain", i, x, y);
        }
        report<N>("div", i, lf::div<N>(da, db), lfo_div(N, a, b));
        {   // division by a reused divisor's reciprocal: random, equal, power-of-two and all-ones mantissas
            lfo_num t = b, u = a;
            if (i % 6 == 1) for (int w = 0; w < N; ++w) u.limb[w] = t.limb[w];
            if (i % 6 == 2) { for (int w = 0; w < N; ++w) t.limb[w] = 0; t.limb[N - 1] = 0x80000000u; }
            if (i % 6 == 3) for (int w = 0; w < N; ++w) t.limb[w] = 0xffffffffu;
            if (i % 6 == 4) { for (int w = 0; w < N; ++w) u.limb[w] = 0xffffffffu; }
            if (i % 6 == 5) { for (int w = 0; w < N; ++w) u.limb[w] = 0; u.limb[N - 1] = 0x80000000u; }
            const auto dt = to_dev<N>(t);
            uint32_t v[N + 1];
            if (dt.s) lf::reciprocal<N>(dt.w, v);
            else for (int w = 0; w <= N; ++w) v[w] = 0;
            report<N>("div by reciprocal", i, lf::div_by_reciprocal<N>(to_dev<N>(u), dt, v), lfo_div(N, u, t));
        }
        root = a;
        lfo_root_case(&g, N, i, &root);
        if (i % 5 == 0 && root.sign) root.sign = -root.sign;
        report<N>("sqrt", i, lf::sqrt<N>(to_dev<N>(root)), lfo_sqrt(N, root));
        lfo_fma_case(&g, N, i, &a, &b, &c);
        da = to_dev<N>(a), db = to_dev<N>(b);
        auto dc = to_dev<N>(c);
        report<N>("fma", i, lf::fma<N>(da, db, dc), lfo_fma(N, a, b, c));
        report<N>("fms", i, lf::fms<N>(da, db, dc), lfo_fms(N, a, b, c));
        // fma at the borders of its near and far cases: c's top bit around P's last bit, P's top bit around
        // c's last bit, c a power of two, both signs
        if (i % 4 == 0 && a.sign && b.sign) {
            const long pe = (long)a.exponent + b.exponent, plsb = pe - 2 * (32L * N - 1);
            const int shifts[] = {-3, -2, -1, 0, 1};
            for (int k = 0; k < 5; ++k)
                for (int side = 0; side < 2; ++side) {
                    lfo_num cc = c;
                    if (!cc.sign) cc = lfo_random_number(&g, N, 10);
                    // side 0: c_msb = p_lsb + shift; side 1: c_lsb = p_msb + 2 - shift (p_msb ~ pe or pe+1)
                    cc.exponent = side == 0 ? (int)(plsb + shifts[k]) : (int)(pe + 2 + shifts[k] + 32L * N - 1);
                    if (k % 2) {
                        for (int w = 0; w < N; ++w) cc.limb[w] = 0;
                        cc.limb[N - 1] = 0x80000000u;
                    }
                    if (cc.exponent > 1000000000 || cc.exponent < -1000000000) continue;
                    cc.sign = (k + side) % 2 ? 1 : -1;
                    report<N>("fma border", i, lf::fma<N>(da, db, to_dev<N>(cc)), lfo_fma(N, a, b, cc));
                }
        }
        if (i % 3 == 0) { // fma with c below half a unit of P's last bit on a product whose tail is exactly one
                          // half: a odd in [2^(32N-1), 2/3 * 2^(32N)) times b = 1.5 puts P's lowest set bit on
                          // the round bit (P = 3a * 2^(32N-2) below 2^(64N-1)); c's sign decides the tie
            lfo_num ta = lfo_random_number(&g, N, 50), tb = lfo_pow2(N, (int)(lfo_mt_next(&g) % 41) - 20);
            ta.limb[N - 1] = 0x80000000u | (ta.limb[N - 1] & 0x1fffffffu); // below 2/3 of 2^(32N): top bits 100
            ta.limb[0] |= 1u;
            tb.limb[N - 1] = 0xc0000000u;
            if (ta.sign && tb.sign) {
                for (int side = 0; side < 4; ++side) {
                    lfo_num cc = lfo_random_number(&g, N, 10);
                    const long plsb = (long)ta.exponent + tb.exponent - 2 * (32L * N - 1);
                    cc.exponent = (int)(plsb - 2 - (long)(lfo_mt_next(&g) % 40));
                    cc.sign = side % 2 ? 1 : -1;
                    if (side >= 2) ta.sign = -ta.sign;
                    if (cc.exponent > 1000000000 || cc.exponent < -1000000000) continue;
                    report<N>("fma far tie", i, lf::fma<N>(to_dev<N>(ta), to_dev<N>(tb), to_dev<N>(cc)), lfo_fma(N, ta, tb, cc));
                    if constexpr (lf::use_digits<N>()) {
                        uint32_t bd[lf::digit_count<N>()];
                        lf::to_digits<N>(to_dev<N>(tb).w, bd);
                        report<N>("fma far tie digits", i, lf::fma_by_digits<N>(to_dev<N>(ta), to_dev<N>(tb), bd, to_dev<N>(cc)),
                                  lfo_fma(N, ta, tb, cc));
                    }
                }
            }
        }
        if (a.sign && b.sign) { // fma over the whole near range: c's top bit from 96N+8 below to 96N+8 above
                                // P's, and c equal to RN(a*b) within one binade (cancellation)
            const long pe = (long)a.exponent + b.exponent, span = 96L * N + 8;
            for (int rep = 0; rep < 4; ++rep) {
                lfo_num cc = lfo_random_number(&g, N, 10);
                cc.exponent = (int)(pe + (long)(lfo_mt_next(&g) % (2 * span + 1)) - span);
                if (rep % 2 == 0) {
                    const lfo_num pr = lfo_mul(N, a, b);
                    if (pr.sign) {
                        for (int w = 0; w < N; ++w) cc.limb[w] = pr.limb[w];
                        if (rep == 2) cc.limb[0] ^= 1u << (lfo_mt_next(&g) % 32);
                        cc.exponent = pr.exponent + (int)(lfo_mt_next(&g) % 3) - 1;
                    }
                }
                if (cc.exponent > 1000000000 || cc.exponent < -1000000000) continue;
                cc.sign = (lfo_mt_next(&g) & 1) ? 1 : -1;
                report<N>("fma sweep", i, lf::fma<N>(da, db, to_dev<N>(cc)), lfo_fma(N, a, b, cc));
            }
        }
        if (i % 4 == 0) {
            lfo_num x[2], y[2], r[2];
            for (int p = 0; p < 2; ++p) { x[p] = lfo_random_number(&g, N, 1000); y[p] = lfo_random_number(&g, N, 1000); }
            if (i % 12 == 4) { x[1] = x[0]; x[1].sign = -x[1].sign; y[1] = y[0]; }
            lf::Cpx<N> cx{to_dev<N>(x[0]), to_dev<N>(x[1])}, cy{to_dev<N>(y[0]), to_dev<N>(y[1])};
            lfo_cdiv(N, x, y, r);
            auto q = lf::cdiv<N>(cx, cy);
            report<N>("cdiv.re", i, q.re, r[0]);
            report<N>("cdiv.im", i, q.im, r[1]);
            lfo_cmul(N, x, y, r);
            auto m = lf::cmul<N>(cx, cy);
            report<N>("cmul.re", i, m.re, r[0]);
            report<N>("cmul.im", i, m.im, r[1]);
        }
        {   // complex fma / fms: cancellation between the products, addends far above and below, huge spreads
            lfo_num x[2], y[2], z[2], r[2];
            for (int p = 0; p < 2; ++p) {
                x[p] = lfo_random_number(&g, N, 1000);
                y[p] = lfo_random_number(&g, N, 1000);
                z[p] = lfo_random_number(&g, N, 3000);
            }
            const long pe = (long)x[0].exponent + y[0].exponent;
            switch (i % 10) {
            case 0: x[1] = x[0]; y[1] = y[0]; z[0].exponent = (int)(pe - (long)(lfo_mt_next(&g) % (uint64_t)(4 * 32 * N))); break;
            case 1: y[1].exponent = y[0].exponent - 5 * 32 * N; break;
            case 2: z[0].exponent = (int)(pe + 3 * 32 * N); break;
            case 3: z[0].exponent = (int)(pe - 4 * 32 * N); z[1] = lfo_zero(0); break;
            case 4: x[0].exponent = 300000000; z[0].exponent = -300000000; y[1].exponent = -200000000; break;
            case 5: x[1] = lfo_zero(0); break;
            case 6: x[0] = lfo_zero(0); x[1] = lfo_zero(0); break;
            case 7: z[1].status = 2; break;
            case 8: x[0].exponent = 500000000; y[0].exponent = 499999999; break;
            default: break;
            }
            lf::Cpx<N> cx{to_dev<N>(x[0]), to_dev<N>(x[1])}, cy{to_dev<N>(y[0]), to_dev<N>(y[1])}, cz{to_dev<N>(z[0]), to_dev<N>(z[1])};
            lfo_cfma(N, x, y, z, r);
            auto f = lf::cfma<N>(cx, cy, cz);
            report<N>("cfma.re", i, f.re, r[0]);
            report<N>("cfma.im", i, f.im, r[1]);
            lfo_cfms(N, x, y, z, r);
            f = lf::cfms<N>(cx, cy, cz);
            report<N>("cfms.re", i, f.re, r[0]);
            report<N>("cfms.im", i, f.im, r[1]);
        }
    }
}

template <int N> void run_all(int lo, int hi, int cases)
{
    if (N >= lo && N <= hi) run_width<N>(cases);
    if constexpr (N < 32) run_all<N + 1>(lo, hi, cases);
}

// The residue GEMM's epilogue, (65536 G11 + 256 Gmid + G00) mod m (residue_combine) and the same times a
// CRT factor y mod m through balanced digit weights (residue_fold, with y = 1, 2, (m+1)/2, m-1 and a random
// value per modulus), against exact 64-bit remainders over every modulus the planner can use (the 1200
// largest primes <= 65279), at the accumulators' extremes (|G11|, |G00| <= 65472 * 2^14,
// |Gmid| <= 2 * 65472 * 2^14), near multiples of m and random values.
static void check_residue_combine()
{
    std::vector<int> primes;
    for (int q = 65279; primes.size() < 1200; --q) {
        bool p = true;
        for (int d = 2; d * d <= q; ++d)
            if (q % d == 0) {
                p = false;
                break;
            }
        if (p) primes.push_back(q);
    }
    uint64_t state = 0x9E3779B97F4A7C15ull;
    auto next = [&] {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return state;
    };
    const int64_t B = 65472LL * 16384, BM = 2 * B;
    const int64_t ends[] = {0, 1, -1, B, -B, B - 1, -B + 1, 1 << 29, -(1 << 29)};
    const int64_t mids[] = {0, 1, -1, BM, -BM, BM - 1, -BM + 1, 1 << 30, -(1 << 30)};
    int64_t ys[5] = {};
    auto check = [&](int64_t a, int64_t b, int64_t c, int m) {
        ++checks;
        const int c32 = int((uint64_t(1) << 32) % uint64_t(m)), cb = c32 > m / 2 ? c32 - m : c32;
        const uint32_t r = lf::residue_combine((int)a, (int)b, (int)c, m, 1.0f / (float)m, cb);
        const int64_t t = a * 65536 + b * 256 + c, want = ((t % m) + m) % m;
        if ((int64_t)r != want && ++failures <= 30)
            std::printf("MISMATCH residue_combine g11=%lld gmid=%lld g00=%lld m=%d: %u, want %lld\n", (long long)a, (long long)b, (long long)c, m, r,
                        (long long)want);
        for (int64_t y : ys) {
            ++checks;
            auto bal = [&](int64_t w) {
                const int64_t k = w * y % m;
                return int(k > m / 2 ? k - m : k);
            };
            const uint32_t f = lf::residue_fold((int)a, (int)b, (int)c, bal(65536), bal(256), bal(1), m, 1.0f / (float)m, cb);
            const int64_t wf = want * y % m;
            if ((int64_t)f != wf && ++failures <= 30)
                std::printf("MISMATCH residue_fold g11=%lld gmid=%lld g00=%lld m=%d y=%lld: %u, want %lld\n", (long long)a, (long long)b, (long long)c, m,
                            (long long)y, f, (long long)wf);
        }
    };
    for (int m : primes) {
        const int64_t fixed[4] = {1, 2, (m + 1) / 2, m - 1};
        for (int k = 0; k < 4; ++k) ys[k] = fixed[k];
        ys[4] = 1 + int64_t(next() % uint64_t(m - 1));
        for (int64_t a : ends)
            for (int64_t b : mids)
                for (int64_t c : ends) check(a, b, c, m);
        for (int k = 0; k < 200; ++k) {
            check(int64_t(next() % uint64_t(2 * B + 1)) - B, int64_t(next() % uint64_t(2 * BM + 1)) - BM, int64_t(next() % uint64_t(2 * B + 1)) - B, m);
            const int64_t a = int64_t(next() % 32768) * m + int64_t(next() % 3) - 1;
            check(a <= B ? a : -m, int64_t(next() % 1024) * m + m / 2, int64_t(next() % 4096) * m + (m - 1) / 2, m);
        }
    }
}

// Cooperative digit columns (coop_core.hpp): for every lane k of a group of G, the lane's share of the
// columns from its shifted factor digits, assembled over the group, equals the product's columns; packed,
// they give digit_columns' words (short and exact forms), on random mantissas of every digit width.
template <int N, int G, bool FULL> static void check_coop_once(uint64_t& state)
{
    constexpr int D = lf::digit_count<N>(), LO = lf::coop_lo<N, G, FULL>(), M = lf::coop_cols<N, G, FULL>();
    auto next = [&] {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return state;
    };
    uint32_t a[N], b[N];
    for (int i = 0; i < N; ++i) {
        a[i] = (uint32_t)next();
        b[i] = (uint32_t)next();
    }
    a[N - 1] |= 0x80000000u;
    b[N - 1] |= 0x80000000u;
    if (next() % 4 == 0) // all-ones and sparse mantissas exercise the carries
        for (int i = 0; i < N - 1; ++i) a[i] = (next() & 1) ? 0xffffffffu : 0u;
    uint32_t ad[D], bd[D];
    lf::to_digits<N>(a, ad);
    lf::to_digits<N>(b, bd);
    uint64_t col[2 * D] = {};
    for (uint32_t k = 0; k < (uint32_t)G; ++k) {
        uint32_t bs[lf::coop_factor_digits<N, G>()];
        lf::coop_shift_factor<N, G>(bd, k, bs);
        uint64_t S[M];
        lf::coop_partial<N, G, FULL>(ad, bs, S);
        for (int m = 0; m < M; ++m)
            if (LO + G * m + (int)k < 2 * D - 1) col[LO + G * m + k] = S[m];
    }
    uint64_t col2[2 * D];
    for (int c = 0; c < 2 * D; ++c) col2[c] = col[c];
    uint32_t got[2 * N], want[2 * N];
    lf::pack_columns<N, true>(col, got, FULL);
    lf::digit_columns<N, true>(ad, bd, want, FULL);
    ++checks;
    bool same = true;
    for (int w = 0; w < 2 * N; ++w) same = same && got[w] == want[w];
    if (!same && ++failures <= 30) std::printf("MISMATCH coop columns N=%d G=%d full=%d\n", N, G, FULL ? 1 : 0);
    // every word of the product (the form the fused multiply-add rounds against its addend)
    lf::pack_columns<N, false>(col2, got, FULL);
    lf::digit_columns<N, false>(ad, bd, want, FULL);
    ++checks;
    same = true;
    for (int w = 0; w < 2 * N; ++w) same = same && got[w] == want[w];
    if (!same && ++failures <= 30) std::printf("MISMATCH coop columns (all words) N=%d G=%d full=%d\n", N, G, FULL ? 1 : 0);
}

template <int N> static void check_coop(uint64_t& state, int cases)
{
    if constexpr (lf::use_digits<N>()) {
        for (int c = 0; c < cases; ++c) {
            check_coop_once<N, 2, false>(state);
            check_coop_once<N, 4, false>(state);
            check_coop_once<N, 2, true>(state);
            check_coop_once<N, 4, true>(state);
        }
    }
    if constexpr (N < 32) check_coop<N + 1>(state, cases);
}

// Divisions with words drawn from {0, 1, 2, 2^31 - 1, 2^31, 2^31 + 1, 2^32 - 2, 2^32 - 1} (three in four)
// and random words: these take Knuth's add-back branch (the quotient word estimate one too large after
// the two-word correction) tens of thousands of times per million cases; against the oracle.
template <int N> static void check_div_structured(uint64_t& state, long cases)
{
    const uint32_t pool[] = {0u, 1u, 2u, 0x7fffffffu, 0x80000000u, 0x80000001u, 0xfffffffeu, 0xffffffffu};
    auto next = [&] {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return state;
    };
    for (long c = 0; c < cases; ++c) {
        lfo_num a{}, b{};
        for (int i = 0; i < N; ++i) {
            a.limb[i] = (next() & 3) ? pool[next() & 7] : (uint32_t)next();
            b.limb[i] = (next() & 3) ? pool[next() & 7] : (uint32_t)next();
        }
        a.limb[N - 1] |= 0x80000000u;
        b.limb[N - 1] |= 0x80000000u;
        a.sign = (next() & 1) ? 1 : -1;
        b.sign = 1;
        a.exponent = (int)(next() % 7) - 3;
        b.exponent = (int)(next() % 7) - 3;
        report<N>("div structured", (int)c, lf::div<N>(to_dev<N>(a), to_dev<N>(b)), lfo_div(N, a, b));
        if (c % 8 == 0) { // the same quotients through the reciprocal
            uint32_t v[N + 1];
            lf::reciprocal<N>(to_dev<N>(b).w, v);
            report<N>("div_by_reciprocal structured", (int)c, lf::div_by_reciprocal<N>(to_dev<N>(a), to_dev<N>(b), v), lfo_div(N, a, b));
        }
    }
}

// RN to N words (ties to even) of ps * P * 2^pl + c, exactly, with GMP: the reference for fma_tail on
// arbitrary 2N-word P (P need not be a product of two N-word mantissas).
template <int N> static lf::Num<N> exact_round_sum(const uint32_t (&p)[2 * N], long pl, int ps, const lf::Num<N>& c)
{
    const long cl = (long)c.e - (32L * N - 1), m = pl < cl ? pl : cl;
    mpz_t P, C, V, q, r, half;
    mpz_inits(P, C, V, q, r, half, NULL);
    mpz_import(P, 2 * N, -1, 4, 0, 0, p);
    mpz_import(C, N, -1, 4, 0, 0, c.w);
    mpz_mul_2exp(P, P, (unsigned long)(pl - m));
    mpz_mul_2exp(C, C, (unsigned long)(cl - m));
    if (ps < 0) mpz_neg(P, P);
    if (c.s < 0) mpz_neg(C, C);
    mpz_add(V, P, C);
    lf::Num<N> res = lf::zero<N>(0);
    const int sg = mpz_sgn(V);
    if (sg) {
        mpz_abs(V, V);
        long nb = (long)mpz_sizeinbase(V, 2);
        const long sh = nb - 32L * N;
        if (sh <= 0) mpz_mul_2exp(q, V, (unsigned long)-sh);
        else {
            mpz_fdiv_q_2exp(q, V, (unsigned long)sh);
            mpz_fdiv_r_2exp(r, V, (unsigned long)sh);
            mpz_set_ui(half, 1);
            mpz_mul_2exp(half, half, (unsigned long)(sh - 1));
            const int cmp = mpz_cmp(r, half);
            if (cmp > 0 || (cmp == 0 && mpz_odd_p(q))) mpz_add_ui(q, q, 1);
            if ((long)mpz_sizeinbase(q, 2) > 32L * N) {
                mpz_fdiv_q_2exp(q, q, 1);
                ++nb;
            }
        }
        uint32_t w[N] = {};
        size_t cnt = 0;
        mpz_export(w, &cnt, -1, 4, 0, 0, q);
        for (int i = 0; i < N; ++i) res.w[i] = w[i];
        res.e = (int)(m + nb - 1);
        res.s = sg;
    }
    mpz_clears(P, C, V, q, r, half, NULL);
    return res;
}

// fma_tail's window path where the result sits next to a rounding tie and only the sticky remainder of
// the smaller term (its bits below the window, with a zero gap above them inside the window) decides it:
// P the larger term with P + C_hi or P - C_hi an exact tie (c's bits just below P's last bit zero, its
// lowest bit set), and c the larger term with C + P_hi carrying to an odd 32N+1-bit value or C - P_hi
// leaving exactly half a unit less the remainder; plus c anywhere in the window range.  Against GMP.
template <int N> static void check_fma_tail_sticky(uint64_t& state, int cases)
{
    if constexpr (N >= 4) {
        auto next = [&] {
            state ^= state << 13;
            state ^= state >> 7;
            state ^= state << 17;
            return state;
        };
        mpz_t P, C, X, t, u;
        mpz_inits(P, C, X, t, u, NULL);
        for (int it = 0; it < cases; ++it)
            for (int scen = 0; scen < 5; ++scen) {
                uint32_t p[2 * N];
                for (int i = 0; i < 2 * N; ++i) p[i] = (uint32_t)next();
                const int top = (int)(next() & 1);
                p[2 * N - 1] = top ? (p[2 * N - 1] | 0x80000000u) : ((p[2 * N - 1] & 0x7fffffffu) | 0x40000000u);
                const long eab = (long)(next() % 401) - 200, pl = eab - 2 * (32L * N - 1), pmsb = eab + top;
                lf::Num<N> c;
                for (int i = 0; i < N; ++i) c.w[i] = (uint32_t)next();
                c.w[N - 1] |= 0xc0000000u;
                c.st = 0;
                const bool add = scen % 2 == 0;
                const int ps = (next() & 1) ? 1 : -1;
                c.s = add ? ps : -ps;
                mpz_import(P, 2 * N, -1, 4, 0, 0, p);
                if (scen < 2) { // P the larger term
                    const long d = 32L * N + 96 + (long)(next() % (uint64_t)(32 * N - 97));
                    c.e = (int)(pmsb - d);
                    const long cl = (long)c.e - (32L * N - 1);
                    mpz_import(C, N, -1, 4, 0, 0, c.w);
                    for (long j = pl - 96 - cl; j < pl - cl; ++j)
                        if (j >= 0 && j < 32L * N) mpz_clrbit(C, (unsigned long)j);
                    mpz_setbit(C, 0);
                    mpz_fdiv_q_2exp(t, C, (unsigned long)(pl - cl)); // C_hi at P's scale
                    if (add) mpz_add(X, P, t);
                    else mpz_sub(X, P, t);
                    const long sh = (long)mpz_sizeinbase(X, 2) - 32L * N;
                    mpz_fdiv_r_2exp(u, X, (unsigned long)sh);
                    mpz_sub(P, P, u);
                    mpz_setbit(t, 0); // any value: reuse t for 2^(sh-1)
                    mpz_set_ui(t, 1);
                    mpz_mul_2exp(t, t, (unsigned long)(sh - 1));
                    mpz_add(P, P, t);
                    uint32_t w[N] = {};
                    size_t cnt = 0;
                    mpz_export(w, &cnt, -1, 4, 0, 0, C);
                    for (int i = 0; i < N; ++i) c.w[i] = w[i];
                } else if (scen < 4) { // c the larger term
                    const long d = 97 + (long)(next() % (uint64_t)(32 * N - 95));
                    c.e = (int)(pmsb + d);
                    const long cl = (long)c.e - (32L * N - 1), w0 = (long)c.e - (64L * N + 94);
                    for (long j = w0 - pl; j < cl - pl; ++j)
                        if (j >= 0 && j < 64L * N) mpz_clrbit(P, (unsigned long)j);
                    mpz_setbit(P, 0);
                    mpz_fdiv_q_2exp(t, P, (unsigned long)(cl - pl)); // P_hi at c's scale
                    if (add) { // C = 2^(32N) - P_hi + j, j odd: C + P_hi = 2^(32N) + j
                        mpz_set_ui(C, 1);
                        mpz_mul_2exp(C, C, 32UL * N);
                        mpz_sub(C, C, t);
                        mpz_add_ui(C, C, 1 + 2 * (next() % 1000));
                    } else { // P's window bits below c: only the half unit, so C - P = (C - P_hi - 1) + half - rem
                        mpz_import(C, N, -1, 4, 0, 0, c.w);
                        if (cl - 1 - pl >= 0) mpz_setbit(P, (unsigned long)(cl - 1 - pl));
                    }
                    uint32_t w[N] = {};
                    size_t cnt = 0;
                    mpz_export(w, &cnt, -1, 4, 0, 0, C);
                    for (int i = 0; i < N; ++i) c.w[i] = w[i];
                } else { // c anywhere in the window range
                    c.e = (int)(pl - 3 + (long)(next() % (uint64_t)(96 * N + 6)));
                }
                uint32_t q[2 * N] = {};
                size_t cnt = 0;
                mpz_export(q, &cnt, -1, 4, 0, 0, P);
                if (cnt > (size_t)(2 * N) || (q[2 * N - 1] >> 30) != (top ? 2u + (q[2 * N - 1] >> 30 & 1u) : 1u)) continue;
                if (!(c.w[N - 1] >> 31)) continue;
                for (int i = 0; i < 2 * N; ++i) p[i] = q[i];
                const lf::Num<N> got = lf::fma_tail<N>(p, eab, ps, c), want = exact_round_sum<N>(p, pl, ps, c);
                ++checks;
                bool ok = got.s == want.s && got.st == want.st && (!got.s || got.e == want.e);
                for (int i = 0; i < N && ok && got.s; ++i) ok = got.w[i] == want.w[i];
                if (!ok && ++failures <= 30)
                    std::printf("MISMATCH fma_tail sticky N=%d scenario=%d case=%d got(s=%d e=%d top=%08x lo=%08x) want(s=%d e=%d top=%08x lo=%08x)\n",
                                N, scen, it, got.s, got.e, got.w[N - 1], got.w[0], want.s, want.e, want.w[N - 1], want.w[0]);
            }
        mpz_clears(P, C, X, t, u, NULL);
    }
    if constexpr (N < 32) check_fma_tail_sticky<N + 1>(state, cases);
}

int main(int argc, char** argv)
{
    int lo = argc > 1 ? std::atoi(argv[1]) : 2, hi = argc > 2 ? std::atoi(argv[2]) : 32, cases = argc > 3 ? std::atoi(argv[3]) : 300;
    run_all<2>(lo, hi, cases);
    check_residue_combine();
    uint64_t coop_state = 0x2545F4914F6CDD1Dull;
    check_coop<8>(coop_state, 200);
    uint64_t tail_state = 0x9E3779B97F4A7C15ull;
    check_fma_tail_sticky<4>(tail_state, cases);
    uint64_t div_state = 0xD1B54A32D192ED03ull;
    check_div_structured<2>(div_state, 400L * cases);
    check_div_structured<3>(div_state, 200L * cases);
    check_div_structured<4>(div_state, 100L * cases);
    check_div_structured<8>(div_state, 20L * cases);
    check_div_structured<32>(div_state, 2L * cases);
    std::printf("device core (host build): %ld checks, %ld failures\n", checks, failures);
    return failures ? 1 : 0;
}

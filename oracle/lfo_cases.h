/*
 * lfo_cases.h — seeded input generation for the oracle's pinning checks and the golden fixtures.
 * TEST INFRASTRUCTURE ONLY (header-only, C).
 *
 * The generator is std::mt19937_64 (Matsumoto & Nishimura's 64-bit Mersenne Twister with the
 * parameters fixed by the C++11 standard), so seeded streams match the reference tests'
 * std::mt19937_64 streams.  random_number restates tests/reference.hpp:21-26; the case classes
 * restate the fixtures of tests/test_arithmetic.cpp:16-62 and tests/test_fused.cpp:10-28.
 * The binade-boundary and root-boundary fixtures use GMP (mpz) as the reference tests do.
 */
#ifndef LFO_CASES_H
#define LFO_CASES_H
#include <gmp.h>
#include <stdint.h>
#include <string.h>
#include "lf_oracle.h"

typedef struct { uint64_t mt[312]; int idx; } lfo_mt64;

static inline void lfo_mt_seed(lfo_mt64* g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i) g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

static inline uint64_t lfo_mt_next(lfo_mt64* g)
{
    static const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
            uint64_t xa = x >> 1;
            if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* tests/reference.hpp:21-26: uniform limbs, top bit set, random sign, exponent in [-span, span] */
static inline lfo_num lfo_random_number(lfo_mt64* g, int n, int span)
{
    lfo_num x = lfo_zero(0);
    for (int i = 0; i < n; ++i) x.limb[i] = (uint32_t)lfo_mt_next(g);
    x.limb[n - 1] |= 0x80000000u;
    x.sign = (lfo_mt_next(g) & 1) ? 1 : -1;
    x.exponent = (int)(lfo_mt_next(g) % (uint64_t)(2 * span + 1)) - span;
    return x;
}

static inline lfo_num lfo_power_of_two(int e) /* 2^e */
{
    lfo_num x = lfo_zero(0);
    x.limb[LFO_MAX_WORDS - 1] = 0; /* filled by caller's width */
    x.sign = 1;
    x.exponent = e;
    return x;
}

static inline lfo_num lfo_pow2(int n, int e)
{
    lfo_num x = lfo_power_of_two(e);
    x.limb[n - 1] = 0x80000000u;
    return x;
}

static inline lfo_num lfo_neg(lfo_num x) { x.sign = -x.sign; return x; }

/* Binary-op cases of test_arithmetic.cpp:20-49 for index i (b may be modified to the fixture). */
static inline void lfo_binary_case(lfo_mt64* g, int n, int i, lfo_num* a, lfo_num* b)
{
    const int bits = 32 * n;
    *a = lfo_random_number(g, n, 1000);
    *b = lfo_random_number(g, n, 1000);
    if (i % 3 == 0) b->exponent = a->exponent - (int)(lfo_mt_next(g) % (uint64_t)(bits + 80)); /* alignment / cancellation */
    if (i % 11 == 0) { *b = lfo_neg(*a); b->limb[0] ^= (uint32_t)(lfo_mt_next(g) & 255); }
    if (i % 17 == 0) *a = lfo_zero(0);
    if (i % 23 == 0) *b = lfo_zero(0);
    if (i % 101 == 0) a->status = LFO_INVALID;
    lfo_num one = lfo_pow2(n, 0), half_ulp = lfo_pow2(n, -bits);
    switch (i) {
    case 1: *a = one; *b = half_ulp; break;
    case 2: *a = one; a->limb[0] |= 1; *b = half_ulp; break;
    case 3: *a = one; *b = lfo_neg(half_ulp); break;                 /* rounding at a binade boundary */
    case 4: *a = one; a->limb[0] = 1; *b = lfo_neg(one); break;        /* exact deep cancellation */
    case 5: *a = one; a->exponent = 1000000000; *b = lfo_pow2(n, 1); break;
    case 6: *a = one; a->exponent = -1000000000; *b = lfo_pow2(n, -1); break;
    case 7: *a = one; *b = one; for (int k = 0; k < n; ++k) a->limb[k] = b->limb[k] = 0xffffffffu; break;
    default: break;
    }
    if (i >= 8 && i < 40) {
        /* exact products just below a binade boundary: rounding can rescue emin-1 (test_arithmetic.cpp:37-49) */
        mpz_t za, zb, target, rem, half;
        mpz_inits(za, zb, target, rem, half, NULL);
        mpz_set_ui(target, 1); mpz_mul_2exp(target, target, 2 * bits - 1);
        mpz_set_ui(half, 1); mpz_mul_2exp(half, half, bits - 2);
        do {
            *a = lfo_random_number(g, n, 0);
            mpz_import(za, n, -1, 4, 0, 0, a->limb);
            mpz_fdiv_qr(zb, rem, target, za);
        } while (mpz_sgn(rem) == 0 || mpz_cmp(rem, half) >= 0);
        *b = lfo_zero(0);
        size_t cnt;
        mpz_export(b->limb, &cnt, -1, 4, 0, 0, zb);
        a->sign = b->sign = i % 2 ? -1 : 1;
        a->exponent = i % 3 ? -1000000000 : 1000000000;
        b->exponent = i % 3 ? -1 : 0;
        mpz_clears(za, zb, target, rem, half, NULL);
    }
}

/* sqrt inputs: the binary a operand with fixtures near root rounding boundaries (test_arithmetic.cpp:50-62) */
static inline void lfo_root_case(lfo_mt64* g, int n, int i, lfo_num* x)
{
    const int bits = 32 * n;
    if (i == 0) { *x = lfo_zero(0); return; }
    if (i == 1) { *x = lfo_pow2(n, 0); return; }
    if (i == 2) { *x = lfo_pow2(n, 1); return; }
    if (i == 3) { *x = lfo_pow2(n, 2); return; }
    if (i >= 40 && i < 104) {
        lfo_num q = lfo_random_number(g, n, 0);
        mpz_t za, zb;
        mpz_inits(za, zb, NULL);
        mpz_import(za, n, -1, 4, 0, 0, q.limb);
        mpz_mul(zb, za, za);
        mpz_add(zb, zb, za); /* floor((q+1/2)^2) */
        int parity = (int)(mpz_sizeinbase(zb, 2) == (size_t)(2 * bits));
        int shift = bits - 1 + parity;
        mpz_fdiv_q_2exp(zb, zb, shift);
        if (i % 2) mpz_add_ui(zb, zb, 1);
        if (mpz_sizeinbase(zb, 2) > (size_t)bits) mpz_sub_ui(zb, zb, 1);
        *x = lfo_zero(0);
        size_t cnt;
        mpz_export(x->limb, &cnt, -1, 4, 0, 0, zb);
        x->sign = 1;
        x->exponent = parity - (i % 3 ? 0 : 2);
        mpz_clears(za, zb, NULL);
    }
}

/* exact mantissa product helper used by the fma cases (the rounded product, as mul) */
/* fma cases of test_fused.cpp:10-28 */
static inline void lfo_fma_case(lfo_mt64* g, int n, int i, lfo_num* a, lfo_num* b, lfo_num* c)
{
    const int bits = 32 * n;
    *a = lfo_random_number(g, n, 40);
    *b = lfo_random_number(g, n, 40);
    *c = lfo_random_number(g, n, 80);
    lfo_num half;
    switch (i % 23) {
    case 0: *c = lfo_neg(lfo_mul(n, *a, *b)); break;
    case 1: *c = lfo_neg(lfo_mul(n, *a, *b)); c->limb[0] ^= (uint32_t)(lfo_mt_next(g) & 15); break;
    case 2: c->exponent = a->exponent + b->exponent - (int)(lfo_mt_next(g) % (uint64_t)(3 * bits + 80)); break;
    case 3: c->exponent = a->exponent + b->exponent + (int)(lfo_mt_next(g) % (uint64_t)(2 * bits + 80)); break;
    case 4:
        *a = lfo_pow2(n, 0);
        half = lfo_zero(0); half.limb[n - 1] = 0x80000000u; half.sign = (lfo_mt_next(g) & 1) ? 1 : -1; half.exponent = b->exponent - bits;
        *c = half; break;
    case 5:
        *a = lfo_pow2(n, 0);
        half = lfo_zero(0); half.limb[n - 1] = 0x80000000u; half.sign = (lfo_mt_next(g) & 1) ? 1 : -1; half.exponent = b->exponent - bits;
        *c = half; c->limb[0] |= 1; break;
    case 6: *a = lfo_zero(0); break;
    case 7: *c = lfo_zero(0); break;
    case 8: *b = lfo_zero(0); *c = lfo_zero(0); break;
    case 9: c->status = LFO_INVALID; break;
    case 10: a->exponent = 600000000; b->exponent = 600000000; *c = lfo_neg(lfo_mul(n, *a, *b)); break;
    case 11: a->exponent = 600000000; b->exponent = 600000000; c->exponent = 999999990; break;
    case 12: a->exponent = -600000000; b->exponent = -500000000; c->exponent = -999999999; break;
    case 13: a->exponent = b->exponent = 0; *c = lfo_neg(*a); c->exponent = b->exponent + a->exponent; break;
    default: break;
    }
}

#endif

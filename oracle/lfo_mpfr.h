/*
 * lfo_mpfr.h — MPFR reference operations for pinning the oracle and writing golden fixtures.
 * TEST INFRASTRUCTURE ONLY (header-only, C; needs MPFR and GMP).
 *
 * Restates the semantics of the reference's independent MPFR oracle, tests/reference.hpp:
 *   - a widened MPFR exponent range so results outside +-1e9 map to exponent_overflow (:11-15);
 *   - real ops with RN-even conversion back to `bits` (:33-44), statuses and division by zero;
 *   - composed complex ops from rounded real ops (:45-55);
 *   - mpfr_fma / mpfr_fms (:57-62);
 *   - exact products at 2*bits summed by one mpfr_sum for dot products (:64-94);
 * and the bridge conversions of include/limbforge/mpfr_bridge.hpp:22-42 (from_mpfr: RN-even to bits,
 * WolfNum exponent = MPFR exponent - 1, out of range -> exponent_overflow).
 */
#ifndef LFO_MPFR_H
#define LFO_MPFR_H
#include <mpfr.h>
#include <stdlib.h>
#include "lf_oracle.h"

static inline void lfo_mpfr_range(void)
{
    mpfr_set_emin(-3000000000L);
    mpfr_set_emax(3000000000L);
}

/* exact conversion (out has at least 32n bits of precision) */
static inline void lfo_to_mpfr(int n, mpfr_ptr out, const lfo_num* x)
{
    if (!x->sign) { mpfr_set_zero(out, 1); return; }
    mpz_t z;
    mpz_init(z);
    mpz_import(z, n, -1, 4, 0, 0, x->limb);
    mpfr_set_z(out, z, MPFR_RNDN);
    mpfr_mul_2si(out, out, (long)x->exponent - (32L * n - 1), MPFR_RNDN);
    if (x->sign < 0) mpfr_neg(out, out, MPFR_RNDN);
    mpz_clear(z);
}

static inline lfo_num lfo_from_mpfr(int n, mpfr_srcptr in)
{
    const int bits = 32 * n;
    if (!mpfr_number_p(in)) return lfo_zero(LFO_INVALID);
    if (mpfr_zero_p(in)) return lfo_zero(LFO_OK);
    mpfr_t x;
    mpfr_init2(x, bits);
    mpfr_set(x, in, MPFR_RNDN);
    long e = (long)mpfr_get_exp(x) - 1;
    if (e > LFO_EMAX || e < -LFO_EMAX) { mpfr_clear(x); return lfo_zero(LFO_EXPONENT_OVERFLOW); }
    lfo_num r = lfo_zero(LFO_OK);
    r.sign = mpfr_sgn(x);
    r.exponent = (int32_t)e;
    mpfr_abs(x, x, MPFR_RNDN);
    mpfr_mul_2si(x, x, bits - 1 - e, MPFR_RNDN);
    mpz_t z;
    mpz_init(z);
    mpfr_get_z(z, x, MPFR_RNDN);
    size_t cnt = 0;
    mpz_export(r.limb, &cnt, -1, 4, 0, 0, z);
    mpz_clear(z);
    mpfr_clear(x);
    return r;
}

/* real op (add, sub, mul, div, square, sqrt) with the reference oracle's status rules */
static inline lfo_num lfo_mpfr_real(int n, int op, const lfo_num* a, const lfo_num* b)
{
    const int bits = 32 * n;
    int unary = op == LFO_SQUARE || op == LFO_SQRT;
    uint32_t st = a->status | (unary ? 0 : b->status);
    if (st) return lfo_zero(st);
    if (op == LFO_DIV && !b->sign) return lfo_zero(LFO_DIVISION_BY_ZERO);
    mpfr_t x, y, z;
    mpfr_inits2(bits, x, y, z, (mpfr_ptr)0);
    lfo_to_mpfr(n, x, a);
    if (!unary) lfo_to_mpfr(n, y, b);
    switch (op) {
    case LFO_ADD: mpfr_add(z, x, y, MPFR_RNDN); break;
    case LFO_SUB: mpfr_sub(z, x, y, MPFR_RNDN); break;
    case LFO_MUL: mpfr_mul(z, x, y, MPFR_RNDN); break;
    case LFO_SQUARE: mpfr_sqr(z, x, MPFR_RNDN); break;
    case LFO_SQRT: mpfr_sqrt(z, x, MPFR_RNDN); break;
    default: mpfr_div(z, x, y, MPFR_RNDN); break;
    }
    lfo_num r = lfo_from_mpfr(n, z);
    mpfr_clears(x, y, z, (mpfr_ptr)0);
    return r;
}

static inline void lfo_mpfr_complex(int n, int op, const lfo_num a[2], const lfo_num b[2], lfo_num r[2])
{
    if (op == LFO_CADD) {
        r[0] = lfo_mpfr_real(n, LFO_ADD, &a[0], &b[0]);
        r[1] = lfo_mpfr_real(n, LFO_ADD, &a[1], &b[1]);
        return;
    }
    lfo_num p0 = lfo_mpfr_real(n, LFO_MUL, &a[0], &b[0]);
    lfo_num p1 = lfo_mpfr_real(n, LFO_MUL, &a[1], &b[1]);
    lfo_num p2 = lfo_mpfr_real(n, LFO_MUL, &a[0], &b[1]);
    lfo_num p3 = lfo_mpfr_real(n, LFO_MUL, &a[1], &b[0]);
    if (op == LFO_CMUL) {
        r[0] = lfo_mpfr_real(n, LFO_SUB, &p0, &p1);
        r[1] = lfo_mpfr_real(n, LFO_ADD, &p2, &p3);
        return;
    }
    lfo_num s0 = lfo_mpfr_real(n, LFO_MUL, &b[0], &b[0]);
    lfo_num s1 = lfo_mpfr_real(n, LFO_MUL, &b[1], &b[1]);
    lfo_num d = lfo_mpfr_real(n, LFO_ADD, &s0, &s1);
    lfo_num re = lfo_mpfr_real(n, LFO_ADD, &p0, &p1);
    lfo_num im = lfo_mpfr_real(n, LFO_SUB, &p3, &p2);
    r[0] = lfo_mpfr_real(n, LFO_DIV, &re, &d);
    r[1] = lfo_mpfr_real(n, LFO_DIV, &im, &d);
}

static inline lfo_num lfo_mpfr_fused(int n, const lfo_num* a, const lfo_num* b, const lfo_num* c, int subtract)
{
    const int bits = 32 * n;
    uint32_t st = a->status | b->status | c->status;
    if (st) return lfo_zero(st);
    mpfr_t x, y, z, r;
    mpfr_inits2(bits, x, y, z, r, (mpfr_ptr)0);
    lfo_to_mpfr(n, x, a);
    lfo_to_mpfr(n, y, b);
    lfo_to_mpfr(n, z, c);
    if (subtract) mpfr_fms(r, x, y, z, MPFR_RNDN);
    else mpfr_fma(r, x, y, z, MPFR_RNDN);
    lfo_num out = lfo_from_mpfr(n, r);
    mpfr_clears(x, y, z, r, (mpfr_ptr)0);
    return out;
}

/* RN(c -/+ sum_k a_k b_k): exact products at 2*bits and c, one mpfr_sum (reference.hpp:74-94).
 * c may be NULL.  a and b are arrays of working numbers with strides. */
static inline lfo_num lfo_mpfr_dot(int n, const lfo_num* c, int subtract, const lfo_num* a, ptrdiff_t sa,
                                   const lfo_num* b, ptrdiff_t sb, size_t K)
{
    const int bits = 32 * n;
    uint32_t st = c ? c->status : 0;
    for (size_t k = 0; k < K; ++k) st |= a[(ptrdiff_t)k * sa].status | b[(ptrdiff_t)k * sb].status;
    if (st) return lfo_zero(st);
    mpfr_t x, y, r;
    mpfr_inits2(bits, x, y, r, (mpfr_ptr)0);
    mpfr_t* p = (mpfr_t*)malloc((K + 1) * sizeof(mpfr_t));
    mpfr_ptr* terms = (mpfr_ptr*)malloc((K + 1) * sizeof(mpfr_ptr));
    for (size_t k = 0; k <= K; ++k) { mpfr_init2(p[k], 2 * bits); terms[k] = p[k]; }
    for (size_t k = 0; k < K; ++k) {
        lfo_to_mpfr(n, x, &a[(ptrdiff_t)k * sa]);
        lfo_to_mpfr(n, y, &b[(ptrdiff_t)k * sb]);
        if (mpfr_mul(p[k], x, y, MPFR_RNDN)) abort(); /* exact at 2*bits */
        if (subtract) mpfr_neg(p[k], p[k], MPFR_RNDN);
    }
    size_t count = K;
    if (c) { lfo_to_mpfr(n, p[K], c); count = K + 1; }
    mpfr_sum(r, terms, count, MPFR_RNDN);
    lfo_num out = lfo_from_mpfr(n, r);
    for (size_t k = 0; k <= K; ++k) mpfr_clear(p[k]);
    free(p);
    free(terms);
    mpfr_clears(x, y, r, (mpfr_ptr)0);
    return out;
}

static inline int lfo_equal(int n, const lfo_num* a, const lfo_num* b)
{
    return a->sign == b->sign && a->exponent == b->exponent && a->status == b->status &&
           memcmp(a->limb, b->limb, 4 * (size_t)n) == 0;
}

#endif

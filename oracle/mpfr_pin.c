/*
 * mpfr_pin — pins the oracle (lf_oracle.c) bit for bit against MPFR, and against the reference's own
 * core.hpp / exact_dot_add compiled into oracle/_ref/libref.so when that library is given.
 * TEST INFRASTRUCTURE ONLY.  Run by tests/test_oracle_pin.py.
 *
 * usage: mpfr_pin <bits|all> <cases> [path/to/libref.so]
 *
 * Coverage per width: add/sub/mul/div/square on the fixture classes of tests/test_arithmetic.cpp:16-49,
 * sqrt on the root-boundary fixtures (:50-62), fma/fms on tests/test_fused.cpp:10-28, composed complex
 * add/mul/div (reference.hpp:45-55), exact dot products and updates (reference.hpp:74-94) with
 * cancellation, ties, far terms, statuses and K = 0, precision casts, and (once) 10^6 divide_word
 * checks against 64-bit division (test_arithmetic.cpp:5-15).
 */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lf_oracle.h"
#include "lfo_cases.h"
#include "lfo_mpfr.h"

typedef int (*ref_elementwise_fn)(int, int, const void*, const void*, const void*, void*, size_t, size_t, size_t, size_t, size_t);
typedef void (*ref_dot_fn)(int, const void*, int, const void*, ptrdiff_t, const void*, ptrdiff_t, size_t, void*);
static ref_elementwise_fn ref_elementwise;
static ref_dot_fn ref_dot;

static long failures, checks, ref_checks;

static void report(const char* what, int bits, long i, const lfo_num* got, const lfo_num* want, int n)
{
    if (++failures > 20) return;
    fprintf(stderr, "MISMATCH %s bits=%d case=%ld: got sign=%d exp=%d st=%u top=%08x low=%08x ; want sign=%d exp=%d st=%u top=%08x low=%08x\n",
            what, bits, i, got->sign, got->exponent, got->status, got->limb[n - 1], got->limb[0], want->sign, want->exponent,
            want->status, want->limb[n - 1], want->limb[0]);
}

static void check(const char* what, int n, long i, const lfo_num* got, const lfo_num* want)
{
    ++checks;
    if (!lfo_equal(n, got, want)) report(what, 32 * n, i, got, want, n);
}

/* oracle vs libref on one element-wise op over records */
static void check_ref(int bits, int op, const lfo_num* a, const lfo_num* b, const lfo_num* c, const lfo_num* got, int cplx)
{
    if (!ref_elementwise) return;
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12, parts = cplx ? 2 : 1;
    unsigned char ra[2 * 140], rb[2 * 140], rc[2 * 140], ro[2 * 140];
    for (size_t p = 0; p < parts; ++p) {
        lfo_store(n, &a[p], ra + p * E);
        lfo_store(n, &b[p], rb + p * E);
        lfo_store(n, &c[p], rc + p * E);
    }
    ref_elementwise(bits, op, ra, rb, rc, ro, 1, 1, 0, 1, 0);
    for (size_t p = 0; p < parts; ++p) {
        lfo_num r;
        lfo_load(n, ro + p * E, &r);
        ++ref_checks;
        if (!lfo_equal(n, &r, &got[p])) report("oracle vs libref", bits, op, &got[p], &r, n);
    }
}

static void pin_width(int bits, int cases, uint64_t seed)
{
    const int n = bits / 32;
    lfo_mt64 g;
    lfo_mt_seed(&g, seed + (uint64_t)bits);
    static const int binops[] = {LFO_ADD, LFO_SUB, LFO_MUL, LFO_DIV, LFO_SQUARE};
    static const char* names[] = {"add", "sub", "mul", "div", "", "", "", "square", "sqrt", "fma", "fms"};
    for (int i = 0; i < cases; ++i) {
        lfo_num a, b, r, m, root;
        lfo_binary_case(&g, n, i, &a, &b);
        for (int q = 0; q < 5; ++q) {
            int op = binops[q];
            switch (op) {
            case LFO_ADD: r = lfo_add(n, a, b); break;
            case LFO_SUB: r = lfo_sub(n, a, b); break;
            case LFO_MUL: r = lfo_mul(n, a, b); break;
            case LFO_DIV: r = lfo_div(n, a, b); break;
            default: r = lfo_square(n, a); break;
            }
            m = lfo_mpfr_real(n, op, &a, &b);
            check(names[op], n, i, &r, &m);
            check_ref(bits, op, &a, &b, &b, &r, 0);
        }
        root = a;
        lfo_root_case(&g, n, i, &root);
        r = lfo_sqrt(n, root);
        m = lfo_mpfr_real(n, LFO_SQRT, &root, &root);
        check("sqrt", n, i, &r, &m);
        check_ref(bits, LFO_SQRT, &root, &root, &root, &r, 0);
        /* negative radicand */
        if (i % 7 == 0) {
            lfo_num neg = root;
            if (neg.sign) neg.sign = -1;
            r = lfo_sqrt(n, neg);
            m = lfo_mpfr_real(n, LFO_SQRT, &neg, &neg);
            if (neg.sign < 0 && !neg.status) m = lfo_zero(LFO_INVALID); /* MPFR gives NaN -> invalid */
            check("sqrt(neg)", n, i, &r, &m);
        }
    }
    for (int i = 0; i < cases; ++i) {
        lfo_num a, b, c, r, m;
        lfo_fma_case(&g, n, i, &a, &b, &c);
        for (int sub = 0; sub < 2; ++sub) {
            r = sub ? lfo_fms(n, a, b, c) : lfo_fma(n, a, b, c);
            m = lfo_mpfr_fused(n, &a, &b, &c, sub);
            check(sub ? "fms" : "fma", n, i, &r, &m);
            check_ref(bits, sub ? LFO_FMS : LFO_FMA, &a, &b, &c, &r, 0);
        }
    }
    for (int i = 0; i < cases / 2; ++i) {
        lfo_num a[2], b[2], r[2], m[2];
        for (int p = 0; p < 2; ++p) { a[p] = lfo_random_number(&g, n, 1000); b[p] = lfo_random_number(&g, n, 1000); }
        switch (i % 9) {
        case 0: b[1] = lfo_zero(0); break;
        case 1: b[0] = lfo_zero(0); break;
        case 2: b[0] = lfo_zero(0); b[1] = lfo_zero(0); break; /* division by zero */
        case 3: a[1] = lfo_neg(a[0]); b[1] = b[0]; break;     /* cancellation in the numerators */
        case 4: b[1].exponent = b[0].exponent - (int)(lfo_mt_next(&g) % (uint64_t)(2 * bits)); break;
        case 5: a[0].status = LFO_INVALID; break;
        default: break;
        }
        for (int op = LFO_CADD; op <= LFO_CDIV; ++op) {
            if (op == LFO_CADD) lfo_cadd(n, a, b, r);
            else if (op == LFO_CMUL) lfo_cmul(n, a, b, r);
            else lfo_cdiv(n, a, b, r);
            lfo_mpfr_complex(n, op, a, b, m);
            check(op == LFO_CADD ? "cadd.re" : op == LFO_CMUL ? "cmul.re" : "cdiv.re", n, i, &r[0], &m[0]);
            check(op == LFO_CADD ? "cadd.im" : op == LFO_CMUL ? "cmul.im" : "cdiv.im", n, i, &r[1], &m[1]);
            check_ref(bits, op, a, b, b, r, 1);
        }
    }
    /* complex fma / fms: each part RN(x y + z w + v), exact products at 2*bits and one mpfr_sum */
    for (int i = 0; i < cases / 2; ++i) {
        lfo_num a[2], b[2], c[2], r[2], m[2];
        for (int p = 0; p < 2; ++p) {
            a[p] = lfo_random_number(&g, n, 1000); b[p] = lfo_random_number(&g, n, 1000); c[p] = lfo_random_number(&g, n, 2000);
        }
        switch (i % 11) {
        case 0: c[0] = lfo_zero(0); break;
        case 1: a[1] = lfo_zero(0); b[1] = lfo_zero(0); break;
        case 2: /* the two products cancel in the real part; the addend decides */
            a[1] = a[0]; b[1] = b[0]; c[0].exponent = a[0].exponent + b[0].exponent - (int)(lfo_mt_next(&g) % (uint64_t)(3 * bits)); break;
        case 3: /* products far apart: the smaller one is a sticky unit */
            b[1].exponent = b[0].exponent - 5 * bits; break;
        case 4: /* addend far above the products */
            c[0].exponent = a[0].exponent + b[0].exponent + 2 * bits; c[1].exponent = c[0].exponent; break;
        case 5: /* addend far below, and tied to the last bit of the real part */
            c[0].exponent = a[0].exponent + b[0].exponent - 4 * bits; break;
        case 6: a[0].status = LFO_INVALID; break;
        case 7: c[1].status = LFO_DIVISION_BY_ZERO; break;
        case 8: /* exponents near the limit */
            a[0].exponent = 500000000; b[0].exponent = 499999999; break;
        case 9: /* huge spread between the three terms */
            a[0].exponent = 300000000; c[0].exponent = -300000000; b[1].exponent = -200000000; break;
        default: break;
        }
        for (int sub = 0; sub < 2; ++sub) {
            lfo_num cc[2] = {c[0], c[1]}, mi = a[1];
            if (sub) { cc[0].sign = -cc[0].sign; cc[1].sign = -cc[1].sign; }
            mi.sign = -mi.sign;
            if (sub) lfo_cfms(n, a, b, c, r); else lfo_cfma(n, a, b, c, r);
            lfo_num x2[2] = {a[0], mi}, y2[2] = {b[0], b[1]};
            m[0] = lfo_mpfr_dot(n, &cc[0], 0, x2, 1, y2, 1, 2);
            lfo_num x3[2] = {a[0], a[1]}, y3[2] = {b[1], b[0]};
            m[1] = lfo_mpfr_dot(n, &cc[1], 0, x3, 1, y3, 1, 2);
            check(sub ? "cfms.re" : "cfma.re", n, i, &r[0], &m[0]);
            check(sub ? "cfms.im" : "cfma.im", n, i, &r[1], &m[1]);
            check_ref(bits, sub ? LFO_CFMS : LFO_CFMA, a, b, c, r, 1);
        }
    }
    /* segmented dot products and tree sums against MPFR replays of their fixed step sequences */
    for (int i = 0; i < cases / 16 + 4; ++i) {
        const size_t segs = 1 + (size_t)(lfo_mt_next(&g) % 5), len = (size_t)(i % 6 == 0 ? 0 : lfo_mt_next(&g) % 40);
        const int shared = i % 3 == 1, cplx = i % 2;
        const size_t P = cplx ? 2 : 1, E = 4 * (size_t)n + 12, na = segs * len * P, nb = (shared ? len : segs * len) * P;
        lfo_num* a = (lfo_num*)calloc(na + 1, sizeof(lfo_num));
        lfo_num* b = (lfo_num*)calloc(nb + 1, sizeof(lfo_num));
        for (size_t k = 0; k < na; ++k) a[k] = lfo_random_number(&g, n, 60);
        for (size_t k = 0; k < nb; ++k) b[k] = lfo_random_number(&g, n, 60);
        if (na > 3 && i % 5 == 2) a[3] = lfo_zero(0);
        if (na > 5 && i % 7 == 3) a[5].status = LFO_INVALID;
        unsigned char* ra = (unsigned char*)malloc((na + 1) * E);
        unsigned char* rb = (unsigned char*)malloc((nb + 1) * E);
        unsigned char* ro = (unsigned char*)malloc(segs * P * E);
        for (size_t k = 0; k < na; ++k) lfo_store(n, &a[k], ra + k * E);
        for (size_t k = 0; k < nb; ++k) lfo_store(n, &b[k], rb + k * E);
        lfo_segmented_dot(bits, cplx, segs, len, shared, ra, rb, ro);
        for (size_t q = 0; q < segs; ++q) {
            lfo_num acc[2] = {lfo_zero(0), lfo_zero(0)};
            for (size_t k = 0; k < len; ++k) {
                const lfo_num* x = &a[(q * len + k) * P];
                const lfo_num* y = &b[(shared ? k : q * len + k) * P];
                if (cplx) { /* each part RN(x.re y.re - x.im y.im + acc.re), RN(x.re y.im + x.im y.re + acc.im) */
                    lfo_num mi = x[1], u[2] = {x[0], mi}, v[2] = {y[0], y[1]}, u2[2] = {x[0], x[1]}, v2[2] = {y[1], y[0]};
                    u[1].sign = -u[1].sign;
                    lfo_num re = lfo_mpfr_dot(n, &acc[0], 0, u, 1, v, 1, 2), im = lfo_mpfr_dot(n, &acc[1], 0, u2, 1, v2, 1, 2);
                    acc[0] = re;
                    acc[1] = im;
                } else
                    acc[0] = lfo_mpfr_fused(n, x, y, &acc[0], 0);
            }
            for (size_t p2 = 0; p2 < P; ++p2) {
                lfo_num r;
                lfo_load(n, ro + (q * P + p2) * E, &r);
                check(cplx ? "segmented dot (complex)" : "segmented dot", n, i, &r, &acc[p2]);
            }
        }
        /* tree sum of a (all of it, as one array) */
        const size_t cnt = na / P;
        lfo_tree_sum(bits, cplx, ra, cnt, ro);
        for (size_t p2 = 0; p2 < P; ++p2) {
            lfo_num* v = (lfo_num*)calloc(cnt + 1, sizeof(lfo_num));
            for (size_t k = 0; k < cnt; ++k) v[k] = a[k * P + p2];
            size_t c = cnt;
            while (c > 1) {
                for (size_t j = 0; j < (c + 1) / 2; ++j) v[j] = 2 * j + 1 < c ? lfo_mpfr_real(n, LFO_ADD, &v[2 * j], &v[2 * j + 1]) : v[2 * j];
                c = (c + 1) / 2;
            }
            lfo_num want = cnt ? v[0] : lfo_zero(0), r;
            lfo_load(n, ro + p2 * E, &r);
            check("tree sum", n, i, &r, &want);
            free(v);
        }
        free(a); free(b); free(ra); free(rb); free(ro);
    }
    /* exact dot products and updates */
    static const size_t Ks[] = {0, 1, 2, 3, 4, 7, 17, 64, 200};
    for (int i = 0; i < cases / 8 + 9; ++i) {
        size_t K = Ks[i % 9];
        lfo_num* a = (lfo_num*)calloc(K + 1, sizeof(lfo_num));
        lfo_num* b = (lfo_num*)calloc(K + 1, sizeof(lfo_num));
        int span = (i % 4 == 0) ? 1000 : (i % 4 == 1 ? 40 : (i % 4 == 2 ? 4 * bits : 300000000));
        for (size_t k = 0; k < K; ++k) { a[k] = lfo_random_number(&g, n, span); b[k] = lfo_random_number(&g, n, span); }
        int cls = (i / 9) % 6;
        if (K >= 2 && cls == 1) { a[1] = lfo_neg(a[0]); b[1] = b[0]; }                       /* exact cancellation */
        if (K >= 3 && cls == 2) { a[1] = lfo_neg(a[0]); b[1] = b[0]; b[2].exponent -= bits + 10; } /* tiny survivor */
        if (K >= 2 && cls == 3) { a[1] = lfo_pow2(n, a[0].exponent + b[0].exponent - bits); b[1] = lfo_pow2(n, 0); b[1].sign = (i & 1) ? 1 : -1;
                                  for (int w = 0; w < n; ++w) { a[0].limb[w] = b[0].limb[w] = 0; } a[0].limb[n - 1] = b[0].limb[n - 1] = 0x80000000u; } /* tie */
        if (K >= 2 && cls == 4) a[K / 2] = lfo_zero(0);
        if (K >= 2 && cls == 5 && i % 2) b[K - 1].status = LFO_EXPONENT_OVERFLOW;
        lfo_num c = lfo_random_number(&g, n, span);
        if (i % 5 == 0) c = lfo_zero(0);
        for (int mode = 0; mode < 3; ++mode) { /* 0: dot, 1: c + dot, 2: c - dot */
            const size_t E = 4 * (size_t)n + 12;
            unsigned char* ra = (unsigned char*)malloc((K + 1) * E);
            unsigned char* rb = (unsigned char*)malloc((K + 1) * E);
            unsigned char rc[140], ro[140], rr[140];
            for (size_t k = 0; k < K; ++k) { lfo_store(n, &a[k], ra + k * E); lfo_store(n, &b[k], rb + k * E); }
            lfo_store(n, &c, rc);
            lfo_exact_dot_add(bits, mode ? rc : NULL, mode == 2, ra, 1, rb, 1, K, ro);
            lfo_num got, want = lfo_mpfr_dot(n, mode ? &c : NULL, mode == 2, a, 1, b, 1, K);
            lfo_load(n, ro, &got);
            check(mode == 0 ? "dot" : mode == 1 ? "dot_add" : "dot_sub", n, i, &got, &want);
            if (ref_dot) {
                ref_dot(bits, mode ? rc : NULL, mode == 2, ra, 1, rb, 1, K, rr);
                lfo_num rv;
                lfo_load(n, rr, &rv);
                ++ref_checks;
                if (!lfo_equal(n, &rv, &got)) report("dot oracle vs libref", bits, i, &got, &rv, n);
            }
            free(ra);
            free(rb);
        }
        free(a);
        free(b);
    }
    /* casts to every other width (exact widening, RN narrowing) */
    for (int i = 0; i < cases / 8 + 1; ++i) {
        lfo_num x = lfo_random_number(&g, n, 1000);
        if (i % 13 == 0) x.exponent = 1000000000;
        if (i % 11 == 0) for (int w = 0; w < n; ++w) x.limb[w] = 0xffffffffu;
        for (int tb = 64; tb <= 1024; tb += 32) {
            int tn = tb / 32;
            lfo_num r = lfo_cast_words(tn, n, &x);
            mpfr_t z;
            mpfr_init2(z, bits);
            lfo_to_mpfr(n, z, &x);
            lfo_num m = lfo_from_mpfr(tn, z);
            mpfr_clear(z);
            check("cast", tn, i, &r, &m);
        }
    }
}

static void pin_divide_word(void)
{
    lfo_mt64 g;
    lfo_mt_seed(&g, 20261007);
    for (int i = 0; i < 1000000; ++i) {
        uint32_t d = (uint32_t)lfo_mt_next(&g) | 0x80000000u;
        if (i % 7 == 0) d = 0x80000000u;
        if (i % 11 == 0) d = 0xffffffffu;
        uint32_t high = (uint32_t)(lfo_mt_next(&g) % d), low = (uint32_t)lfo_mt_next(&g);
        if (i % 13 == 0) high = d - 1;
        if (i % 17 == 0) high = 0;
        if (i % 19 == 0) low = 0;
        if (i % 23 == 0) low = 0xffffffffu;
        uint32_t inv = (uint32_t)(~(uint64_t)0 / d - ((uint64_t)1 << 32));
        uint64_t qr = lfo_divide_word(high, low, d, inv), num = ((uint64_t)high << 32) | low;
        ++checks;
        if ((uint32_t)qr != num / d || (uint32_t)(qr >> 32) != num % d) {
            if (++failures < 20) fprintf(stderr, "MISMATCH divide_word case=%d\n", i);
        }
    }
}

int main(int argc, char** argv)
{
    if (argc < 3) { fprintf(stderr, "usage: mpfr_pin <bits|all> <cases> [libref.so]\n"); return 2; }
    lfo_mpfr_range();
    if (argc > 3) {
        void* h = dlopen(argv[3], RTLD_NOW);
        if (!h) { fprintf(stderr, "cannot load %s: %s\n", argv[3], dlerror()); return 2; }
        ref_elementwise = (ref_elementwise_fn)dlsym(h, "ref_elementwise");
        ref_dot = (ref_dot_fn)dlsym(h, "ref_exact_dot_add");
        if (!ref_elementwise || !ref_dot) { fprintf(stderr, "libref symbols missing\n"); return 2; }
    }
    int cases = atoi(argv[2]);
    if (!strcmp(argv[1], "all")) {
        pin_divide_word();
        for (int bits = 64; bits <= 1024; bits += 32) pin_width(bits, cases, 20261007);
    } else
        pin_width(atoi(argv[1]), cases, 20261007);
    printf("mpfr_pin: %ld MPFR checks, %ld libref checks, %ld failures\n", checks, ref_checks, failures);
    return failures ? 1 : 0;
}

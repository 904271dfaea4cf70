/*
 * mpfr_factor_pin — pins the oracle's blocked Cholesky and triangular solves (lfo_cholesky, lfo_solve:
 * the reference's right-looking host order, src/linalg.mm:411-486) against an independent left-looking
 * MPFR program of the documented rounding sequence (docs/numerics.md:687-718; the structure of the
 * reference's tests/test_cholesky.cpp:32-55): entry (i, j) is a_ij reduced by one exact-dot rounding per
 * complete column block before j's block, then the in-block dot, then mpfr_sqrt or mpfr_div.  Every
 * exact dot is 2*bits mpfr_mul products plus one mpfr_sum (lfo_mpfr_dot).  A mistaken block boundary,
 * a dot over the wrong range, a wrong pivot rule or a transposed solve operand changes bits here.
 *
 * Test infrastructure only.  usage: mpfr_factor_pin <cases-scale>
 */
#include "lf_oracle.h"
#include "lfo_cases.h"
#include "lfo_mpfr.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static long failures, checks;

typedef struct { size_t pivot; int sign; uint32_t status; } info_t;

/* RN(v - sum_{k<K} x[k*sx] y[k*sy]) */
static lfo_num dsub(int n, lfo_num v, const lfo_num* x, ptrdiff_t sx, const lfo_num* y, ptrdiff_t sy, size_t K)
{
    return lfo_mpfr_dot(n, &v, 1, x, sx, y, sy, K);
}

static info_t mpfr_cholesky(int n, const lfo_num* A, size_t m, size_t nb, lfo_num* L)
{
    info_t info = {m, 0, 0};
    if (!nb || nb > m) nb = m;
    for (size_t q = 0; q < m * m; ++q) L[q] = lfo_zero(0);
    for (size_t j = 0; j < m; ++j) {
        const size_t kb = j / nb * nb;
        for (size_t i = j; i < m; ++i) {
            lfo_num v = A[i * m + j];
            for (size_t b = 0; b < kb; b += nb) v = dsub(n, v, L + i * m + b, 1, L + j * m + b, 1, nb);
            v = dsub(n, v, L + i * m + kb, 1, L + j * m + kb, 1, j - kb);
            if (i == j) {
                if (v.status || v.sign <= 0) {
                    info.pivot = j;
                    info.sign = v.status ? 0 : v.sign;
                    info.status = v.status;
                    for (size_t r = j; r < m; ++r)
                        for (size_t c = j; c <= r; ++c) L[r * m + c] = lfo_zero(LFO_INVALID);
                    return info;
                }
                L[j * m + j] = lfo_mpfr_real(n, LFO_SQRT, &v, &v);
            } else
                L[i * m + j] = lfo_mpfr_real(n, LFO_DIV, &v, &L[j * m + j]);
        }
    }
    return info;
}

/* passes 1 forward (L X = B), 2 backward (L^T X = B), 3 both; X holds B on entry */
static void mpfr_solve(int n, const lfo_num* L, size_t m, lfo_num* X, size_t nrhs, size_t nb, int passes)
{
    if (!nb || nb > m) nb = m;
    if (passes & 1)
        for (size_t i = 0; i < m; ++i)
            for (size_t c = 0; c < nrhs; ++c) {
                lfo_num v = X[i * nrhs + c];
                const size_t r0 = i / nb * nb;
                for (size_t b = 0; b < r0; b += nb) v = dsub(n, v, L + i * m + b, 1, X + b * nrhs + c, (ptrdiff_t)nrhs, nb);
                v = dsub(n, v, L + i * m + r0, 1, X + r0 * nrhs + c, (ptrdiff_t)nrhs, i - r0);
                X[i * nrhs + c] = lfo_mpfr_real(n, LFO_DIV, &v, &L[i * m + i]);
            }
    if (passes & 2)
        for (size_t i = m; i-- > 0;)
            for (size_t c = 0; c < nrhs; ++c) {
                lfo_num v = X[i * nrhs + c];
                const size_t r0 = i / nb * nb, r1 = r0 + nb < m ? r0 + nb : m;
                for (size_t b = (m - 1) / nb * nb; b >= r1 && b > r0; b -= nb) {
                    const size_t e = b + nb < m ? b + nb : m;
                    v = dsub(n, v, L + b * m + i, (ptrdiff_t)m, X + b * nrhs + c, (ptrdiff_t)nrhs, e - b);
                }
                if (i + 1 < r1) v = dsub(n, v, L + (i + 1) * m + i, (ptrdiff_t)m, X + (i + 1) * nrhs + c, (ptrdiff_t)nrhs, r1 - i - 1);
                X[i * nrhs + c] = lfo_mpfr_real(n, LFO_DIV, &v, &L[i * m + i]);
            }
}

/* ---- inputs: A = J^T J + damping (test_cholesky.cpp:71-78), banded, or failure edits ---- */
static void store_all(int n, const lfo_num* x, size_t count, unsigned char* out)
{
    const size_t E = 4 * (size_t)n + 12;
    for (size_t q = 0; q < count; ++q) lfo_store(n, &x[q], out + q * E);
}
static void load_all(int n, const unsigned char* in, size_t count, lfo_num* x)
{
    const size_t E = 4 * (size_t)n + 12;
    for (size_t q = 0; q < count; ++q) lfo_load(n, in + q * E, &x[q]);
}

static void make_spd(lfo_mt64* g, int n, size_t m, int kind, lfo_num* A)
{
    const int bits = 32 * n;
    const size_t E = 4 * (size_t)n + 12, rows = 2 * m + 1;
    if (kind == 3) { /* banded, diagonally dominant */
        for (size_t q = 0; q < m * m; ++q) A[q] = lfo_zero(0);
        for (size_t i = 0; i < m; ++i) {
            A[i * m + i] = lfo_random_number(g, n, 0);
            A[i * m + i].sign = 1;
            A[i * m + i].exponent = 3;
            for (size_t d = 1; d <= 3 && d <= i; ++d)
                if (d != 2) {
                    lfo_num x = lfo_random_number(g, n, 0);
                    x.exponent = -1;
                    A[i * m + i - d] = A[(i - d) * m + i] = x;
                }
        }
        return;
    }
    lfo_num* J = (lfo_num*)malloc(rows * m * sizeof(lfo_num));
    for (size_t c = 0; c < m; ++c) {
        const long scale = kind == 1 ? (long)(lfo_mt_next(g) % 121) - 60 : kind == 2 ? (long)(lfo_mt_next(g) % 601) - 300 : 0;
        for (size_t r = 0; r < rows; ++r) {
            lfo_num x = lfo_random_number(g, n, kind == 1 ? 15 : 4);
            x.exponent += (int)scale;
            if (kind == 1 && lfo_mt_next(g) % 4 == 0) x = lfo_zero(0);
            J[r * m + c] = x;
        }
    }
    unsigned char* rj = (unsigned char*)calloc(rows * m, E);
    unsigned char* ra = (unsigned char*)calloc(m * m, E);
    store_all(n, J, rows * m, rj);
    lfo_syrk(bits, rj, rows, m, ra, 0, 0);
    load_all(n, ra, m * m, A);
    for (size_t i = 0; i < m; ++i) {
        lfo_num* a = &A[i * m + i];
        if (a->sign) {
            lfo_num d = lfo_pow2(n, a->exponent - 10);
            *a = lfo_mpfr_real(n, LFO_ADD, a, &d);
        }
    }
    free(J);
    free(rj);
    free(ra);
}

static void compare(int n, const char* what, int bits, size_t m, size_t nb, const lfo_num* got, const lfo_num* want, size_t count)
{
    for (size_t q = 0; q < count; ++q) {
        ++checks;
        if (!lfo_equal(n, &got[q], &want[q])) {
            if (++failures < 20)
                fprintf(stderr, "MISMATCH %s bits=%d n=%zu block=%zu entry %zu: got e=%d s=%d st=%u, want e=%d s=%d st=%u\n", what, bits, m, nb, q,
                        got[q].exponent, got[q].sign, got[q].status, want[q].exponent, want[q].sign, want[q].status);
            return;
        }
    }
}

static void run_case(lfo_mt64* g, int bits, size_t m, const lfo_num* A, size_t nb, int nrhs)
{
    const int n = bits / 32;
    const size_t E = 4 * (size_t)n + 12;
    lfo_num* want = (lfo_num*)malloc(m * m * sizeof(lfo_num) + 1);
    lfo_num* got = (lfo_num*)malloc(m * m * sizeof(lfo_num) + 1);
    unsigned char* ra = (unsigned char*)calloc(m * m * E + 1, 1);
    unsigned char* rl = (unsigned char*)malloc(m * m * E + 1);
    store_all(n, A, m * m, ra);
    int psign = 0;
    uint32_t pstat = 0;
    const size_t piv = lfo_cholesky(bits, ra, m, rl, nb, &psign, &pstat);
    load_all(n, rl, m * m, got);
    const info_t info = mpfr_cholesky(n, A, m, nb, want);
    ++checks;
    if (piv != info.pivot || psign != info.sign || pstat != info.status) {
        if (++failures < 20) fprintf(stderr, "MISMATCH pivot bits=%d n=%zu block=%zu: %zu %d %u vs %zu %d %u\n", bits, m, nb, piv, psign, pstat, info.pivot, info.sign, info.status);
    }
    compare(n, "cholesky", bits, m, nb, got, want, m * m);
    if (piv == m && nrhs > 0) {
        lfo_num* B = (lfo_num*)malloc(m * nrhs * sizeof(lfo_num));
        lfo_num* X = (lfo_num*)malloc(m * nrhs * sizeof(lfo_num));
        unsigned char* rb = (unsigned char*)malloc(m * nrhs * E);
        unsigned char* rx = (unsigned char*)malloc(m * nrhs * E);
        for (size_t q = 0; q < m * (size_t)nrhs; ++q) {
            B[q] = lfo_random_number(g, n, 8);
            if (lfo_mt_next(g) % 11 == 0) B[q] = lfo_zero(0);
        }
        if (m > 3 && nrhs > 1) B[nrhs + 1].status = LFO_INVALID;
        store_all(n, B, m * nrhs, rb);
        for (int passes = 1; passes <= 3; ++passes) {
            memcpy(X, B, m * nrhs * sizeof(lfo_num));
            mpfr_solve(n, want, m, X, (size_t)nrhs, nb, passes);
            lfo_solve(bits, rl, m, rb, (size_t)nrhs, rx, nb, passes);
            lfo_num* Y = (lfo_num*)malloc(m * nrhs * sizeof(lfo_num));
            load_all(n, rx, m * nrhs, Y);
            compare(n, passes == 1 ? "trsm" : passes == 2 ? "trsm transposed" : "cholesky_solve", bits, m, nb, Y, X, m * nrhs);
            free(Y);
        }
        free(B);
        free(X);
        free(rb);
        free(rx);
    }
    free(want);
    free(got);
    free(ra);
    free(rl);
}

int main(int argc, char** argv)
{
    if (argc < 2) { fprintf(stderr, "usage: mpfr_factor_pin <cases-scale>\n"); return 2; }
    const int scale = atoi(argv[1]);
    lfo_mpfr_range();
    lfo_mt64 g;
    lfo_mt_seed(&g, 4242);
    const int widths[] = {64, 224, 256, 384};
    for (int wi = 0; wi < 4; ++wi) {
        const int bits = widths[wi], n = bits / 32;
        const size_t sizes[] = {1, 2, 7, 17, 33};
        const size_t blocks[] = {32, 16, 7, 1, 0};
        for (int si = 0; si < 5; ++si)
            for (int kind = 0; kind < 4; ++kind)
                for (int bi = 0; bi < 5; ++bi) {
                    const size_t m = sizes[si];
                    if (bi == 3 && m > 17) continue;
                    lfo_num* A = (lfo_num*)malloc(m * m * sizeof(lfo_num));
                    make_spd(&g, n, m, kind, A);
                    run_case(&g, bits, m, A, blocks[bi], bi % 2 ? 3 : 1);
                    free(A);
                }
        /* failures: negative pivot, status entries, a zero row (test_cholesky.cpp:106-115) */
        for (int f = 0; f < 4; ++f) {
            const size_t m = 20;
            lfo_num* A = (lfo_num*)malloc(m * m * sizeof(lfo_num));
            make_spd(&g, n, m, 0, A);
            if (f == 0) { A[13 * m + 13] = lfo_pow2(n, 40); A[13 * m + 13].sign = -1; }
            if (f == 1) A[11 * m + 5].status = LFO_INVALID;
            if (f == 2) A[0].status = LFO_EXPONENT_OVERFLOW;
            if (f == 3) for (size_t j = 0; j < m; ++j) A[(j > 9 ? j : 9) * m + (j < 9 ? j : 9)] = lfo_zero(0);
            for (int bi = 0; bi < 3; ++bi) run_case(&g, bits, m, A, (size_t[]){8, 3, 0}[bi], 2);
            free(A);
        }
        for (int extra = 0; extra < scale; ++extra) {
            const size_t m = 40 + 8 * (size_t)extra;
            lfo_num* A = (lfo_num*)malloc(m * m * sizeof(lfo_num));
            make_spd(&g, n, m, extra % 3, A);
            run_case(&g, bits, m, A, 16, 2);
            free(A);
        }
    }
    printf("mpfr_factor_pin: %ld checks, %ld failures\n", checks, failures);
    return failures ? 1 : 0;
}

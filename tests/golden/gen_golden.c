/*
 * gen_golden — writes the committed golden fixtures of tests/golden/ (inputs and MPFR-expected outputs).
 * Build and run from the repository root:
 *   gcc -O2 -std=c11 -I/opt/conda/include -Ioracle tests/golden/gen_golden.c oracle/lf_oracle.c \
 *       -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmpfr -lgmp -o /tmp/gen_golden && /tmp/gen_golden tests/golden
 *
 * Expected values come from MPFR only (the reference's independent oracle semantics, tests/reference.hpp):
 * element-wise ops, fused ops, composed complex ops, 16-step multiplication chains, exact dot products
 * and small dense products (one mpfr_sum of exact products per output).  Inputs follow the reference
 * tests' seeded fixture classes (oracle/lfo_cases.h) plus the reference's known-answer cases:
 *   examples/resident.cpp:8-17         1.01^16 at 384 bits, rounded after every step
 *   tests/test_contract_names.cpp:14-19 [2^100, 1, -2^100] . [1, 1, 1] = 1 exactly at 64 bits
 *
 * Format (little endian): "LFG1", u32 kind, u32 bits, u32 sections; then per section four u32 fields and
 * the record arrays (record = 4n+12 bytes, two records for a complex value):
 *   kind 1 element-wise: op, count, has_c, 0;      a[count], b[count], (c[count]), out[count]
 *   kind 2 chain:        bits, count, steps, 0;    x0[count], f[count], out[count]
 *   kind 3 dot:          K, mode (0 dot, 1 c+dot, 2 c-dot), 0, 0;  a[K], b[K], c, out
 *   kind 4 product:      bits, then m, n, k, flags (1 transpose_a, 2 subtract, 4 syrk, 8 lower_only);
 *                        A, B (absent for syrk), C_in (m*n), C_out (m*n)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lf_oracle.h"
#include "lfo_cases.h"
#include "lfo_mpfr.h"

static FILE* out_file;
static void put_u32(uint32_t v) { fwrite(&v, 4, 1, out_file); }
static void put_num(int n, const lfo_num* x)
{
    unsigned char rec[140];
    lfo_store(n, x, rec);
    fwrite(rec, 4 * (size_t)n + 12, 1, out_file);
}
static void open_file(const char* dir, const char* name, uint32_t kind, uint32_t bits, uint32_t sections)
{
    char path[512];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    out_file = fopen(path, "wb");
    if (!out_file) { perror(path); exit(1); }
    fwrite("LFG1", 4, 1, out_file);
    put_u32(kind); put_u32(bits); put_u32(sections);
}

/* Irwin-Hall approximation of a standard normal deviate times `sigma`, rounded, in integer arithmetic */
static int normal_int(lfo_mt64* g, int sigma)
{
    int64_t s = 0;
    for (int i = 0; i < 12; ++i) s += (int64_t)(lfo_mt_next(g) >> 44); /* 20-bit uniforms */
    s -= (int64_t)6 << 20;
    int64_t v = (int64_t)sigma * s + ((int64_t)1 << 19);
    return (int)(v >= 0 ? v >> 20 : -((-v + ((1 << 20) - 1)) >> 20));
}

static void elementwise_file(const char* dir, int bits)
{
    const int n = bits / 32, R = 32, CX = 16;
    static const int ops[] = {LFO_ADD, LFO_SUB, LFO_MUL, LFO_DIV, LFO_SQUARE, LFO_SQRT, LFO_FMA, LFO_FMS, LFO_CADD, LFO_CMUL, LFO_CDIV};
    char name[64];
    snprintf(name, sizeof name, "elementwise_%d.bin", bits);
    open_file(dir, name, 1, (uint32_t)bits, 11);
    lfo_mt64 g;
    lfo_mt_seed(&g, 20261007u + (uint64_t)bits);
    for (int q = 0; q < 11; ++q) {
        int op = ops[q];
        int cplx = op >= LFO_CADD && op <= LFO_CDIV, fused = op == LFO_FMA || op == LFO_FMS;
        int count = cplx ? CX : R;
        lfo_num *a = calloc((size_t)count * 2, sizeof(lfo_num)), *b = calloc((size_t)count * 2, sizeof(lfo_num));
        lfo_num *c = calloc((size_t)count * 2, sizeof(lfo_num)), *o = calloc((size_t)count * 2, sizeof(lfo_num));
        for (int i = 0; i < count; ++i) {
            if (fused) {
                lfo_fma_case(&g, n, i, &a[i], &b[i], &c[i]);
                o[i] = lfo_mpfr_fused(n, &a[i], &b[i], &c[i], op == LFO_FMS);
            } else if (cplx) {
                for (int p = 0; p < 2; ++p) { a[2 * i + p] = lfo_random_number(&g, n, 1000); b[2 * i + p] = lfo_random_number(&g, n, 1000); }
                if (i % 5 == 1) b[2 * i + 1] = lfo_zero(0);
                if (i % 7 == 2) { b[2 * i] = lfo_zero(0); b[2 * i + 1] = lfo_zero(0); }
                if (i % 6 == 3) { a[2 * i + 1] = lfo_neg(a[2 * i]); b[2 * i + 1] = b[2 * i]; }
                lfo_mpfr_complex(n, op, &a[2 * i], &b[2 * i], &o[2 * i]);
            } else if (op == LFO_SQRT) {
                /* root-boundary fixtures live at case indices 40..103; take every other one plus the fixed ones */
                int idx = i < 6 ? i : 40 + 2 * (i - 6) + (i & 1);
                lfo_num t, u;
                lfo_binary_case(&g, n, idx, &t, &u);
                a[i] = t;
                lfo_root_case(&g, n, idx, &a[i]);
                if (i == 5 && a[i].sign) a[i].sign = -1; /* negative radicand */
                o[i] = (a[i].sign < 0 && !a[i].status) ? lfo_zero(LFO_INVALID) : lfo_mpfr_real(n, LFO_SQRT, &a[i], &a[i]);
            } else {
                int idx = i < 24 ? i : 101 + (i - 24); /* fixed fixtures 0..23 then a status case and randoms */
                lfo_binary_case(&g, n, idx, &a[i], &b[i]);
                o[i] = lfo_mpfr_real(n, op, &a[i], &b[i]);
            }
        }
        put_u32((uint32_t)op); put_u32((uint32_t)count); put_u32((uint32_t)fused); put_u32(0);
        int parts = cplx ? 2 : 1;
        for (int i = 0; i < count * parts; ++i) put_num(n, &a[i]);
        for (int i = 0; i < count * parts; ++i) put_num(n, &b[i]);
        if (fused) for (int i = 0; i < count; ++i) put_num(n, &c[i]);
        for (int i = 0; i < count * parts; ++i) put_num(n, &o[i]);
        free(a); free(b); free(c); free(o);
    }
    fclose(out_file);
}

static void chain_file(const char* dir)
{
    /* 1024-bit and 256-bit chains with the benchmark's inputs (exponent span 64), plus the 384-bit known answer */
    open_file(dir, "chain.bin", 2, 0, 3);
    static const int widths[] = {1024, 256};
    for (int w = 0; w < 2; ++w) {
        const int bits = widths[w], n = bits / 32, count = 64, steps = 16;
        lfo_mt64 g;
        lfo_mt_seed(&g, 20261007u + (uint64_t)bits);
        lfo_num x[64], f[64];
        for (int i = 0; i < count; ++i) { x[i] = lfo_random_number(&g, n, 64); f[i] = lfo_random_number(&g, n, 64); }
        put_u32((uint32_t)bits); put_u32((uint32_t)count); put_u32((uint32_t)steps); put_u32(0);
        for (int i = 0; i < count; ++i) put_num(n, &x[i]);
        for (int i = 0; i < count; ++i) put_num(n, &f[i]);
        mpfr_t r, y;
        mpfr_inits2(bits, r, y, (mpfr_ptr)0);
        for (int i = 0; i < count; ++i) {
            lfo_to_mpfr(n, r, &x[i]);
            lfo_to_mpfr(n, y, &f[i]);
            for (int s = 0; s < steps; ++s) mpfr_mul(r, r, y, MPFR_RNDN);
            lfo_num o = lfo_from_mpfr(n, r);
            put_num(n, &o);
        }
        mpfr_clears(r, y, (mpfr_ptr)0);
    }
    {
        const int bits = 384, n = 12, count = 4, steps = 16;
        mpfr_t one, f, r;
        mpfr_inits2(bits, one, f, r, (mpfr_ptr)0);
        mpfr_set_ui(one, 1, MPFR_RNDN);
        mpfr_set_str(f, "1.01", 10, MPFR_RNDN);
        mpfr_set(r, one, MPFR_RNDN);
        for (int s = 0; s < steps; ++s) mpfr_mul(r, r, f, MPFR_RNDN);
        lfo_num x1 = lfo_from_mpfr(n, one), f1 = lfo_from_mpfr(n, f), o = lfo_from_mpfr(n, r);
        put_u32((uint32_t)bits); put_u32((uint32_t)count); put_u32((uint32_t)steps); put_u32(0);
        for (int i = 0; i < count; ++i) put_num(n, &x1);
        for (int i = 0; i < count; ++i) put_num(n, &f1);
        for (int i = 0; i < count; ++i) put_num(n, &o);
        mpfr_clears(one, f, r, (mpfr_ptr)0);
    }
    fclose(out_file);
}

static void dot_file(const char* dir, int bits)
{
    const int n = bits / 32;
    static const size_t Ks[] = {0, 1, 2, 3, 5, 17, 64};
    char name[64];
    snprintf(name, sizeof name, "dot_%d.bin", bits);
    open_file(dir, name, 3, (uint32_t)bits, 7 * 3 * 2);
    lfo_mt64 g;
    lfo_mt_seed(&g, 777u + (uint64_t)bits);
    for (int t = 0; t < 14; ++t) {
        size_t K = Ks[t % 7];
        lfo_num* a = calloc(K + 1, sizeof(lfo_num));
        lfo_num* b = calloc(K + 1, sizeof(lfo_num));
        int span = t < 7 ? 64 : 3000;
        for (size_t k = 0; k < K; ++k) { a[k] = lfo_random_number(&g, n, span); b[k] = lfo_random_number(&g, n, span); }
        if (K >= 3) { a[1] = lfo_neg(a[0]); b[1] = b[0]; b[2].exponent -= bits + 10; } /* cancellation leaving a tiny term */
        lfo_num c = lfo_random_number(&g, n, span);
        for (int mode = 0; mode < 3; ++mode) {
            lfo_num o = lfo_mpfr_dot(n, mode ? &c : NULL, mode == 2, a, 1, b, 1, K);
            put_u32((uint32_t)K); put_u32((uint32_t)mode); put_u32(0); put_u32(0);
            for (size_t k = 0; k < K; ++k) put_num(n, &a[k]);
            for (size_t k = 0; k < K; ++k) put_num(n, &b[k]);
            put_num(n, &c);
            put_num(n, &o);
        }
        free(a); free(b);
    }
    fclose(out_file);
}

/* one dense product section with MPFR-expected outputs */
static void product_section(int n, size_t m, size_t nn, size_t k, uint32_t flags, const lfo_num* A, const lfo_num* B, const lfo_num* Cin)
{
    const int ta = flags & 1, sub = (flags & 2) != 0, syrk = (flags & 4) != 0, lower = (flags & 8) != 0;
    put_u32((uint32_t)m); put_u32((uint32_t)nn); put_u32((uint32_t)k); put_u32(flags);
    size_t asz = syrk ? k * m : m * k;
    for (size_t i = 0; i < asz; ++i) put_num(n, &A[i]);
    if (!syrk) for (size_t i = 0; i < k * nn; ++i) put_num(n, &B[i]);
    for (size_t i = 0; i < m * nn; ++i) put_num(n, &Cin[i]);
    lfo_num* out = calloc(m * nn, sizeof(lfo_num));
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < nn; ++j) {
            if (syrk && lower && j > i) { out[i * nn + j] = Cin[i * nn + j]; continue; }
            const lfo_num *row, *col;
            ptrdiff_t sa, sb;
            if (syrk) { row = A + i; sa = (ptrdiff_t)m; col = A + j; sb = (ptrdiff_t)m; }
            else {
                row = ta ? A + i : A + i * k; sa = ta ? (ptrdiff_t)m : 1;
                col = B + j; sb = (ptrdiff_t)nn;
            }
            out[i * nn + j] = lfo_mpfr_dot(n, sub ? &Cin[i * nn + j] : NULL, sub, row, sa, col, sb, k);
        }
    for (size_t i = 0; i < m * nn; ++i) put_num(n, &out[i]);
    free(out);
}

static lfo_num band_entry(lfo_mt64* g, int n, int centres)
{
    lfo_num x = lfo_random_number(g, n, 0);
    int c = centres == 3 ? 200 * ((int)(lfo_mt_next(g) % 3) - 1) : 0;
    x.exponent = c + normal_int(g, 8);
    return x;
}

static void product_file(const char* dir)
{
    open_file(dir, "products.bin", 4, 0, 8);
    lfo_mt64 g;
    lfo_mt_seed(&g, 4242);
    lfo_num A[512], B[512], C[512];
    /* 1: 512-bit GEMM with three exponent bands per line (centres -200, 0, +200, sigma 8), ragged sizes */
    {
        const int n = 16; size_t m = 5, nn = 7, k = 19;
        for (size_t i = 0; i < m * k; ++i) A[i] = band_entry(&g, n, 3);
        for (size_t i = 0; i < k * nn; ++i) B[i] = band_entry(&g, n, 3);
        for (size_t i = 0; i < m * nn; ++i) C[i] = lfo_zero(0);
        put_u32(512); product_section(n, m, nn, k, 0, A, B, C);
    }
    /* 2: 512-bit transposed-A update C - A^T B with a random old C */
    {
        const int n = 16; size_t m = 4, nn = 6, k = 11;
        for (size_t i = 0; i < k * m; ++i) A[i] = band_entry(&g, n, 3);
        for (size_t i = 0; i < k * nn; ++i) B[i] = band_entry(&g, n, 3);
        for (size_t i = 0; i < m * nn; ++i) { C[i] = lfo_random_number(&g, n, 0); C[i].exponent = 400 + normal_int(&g, 8); }
        put_u32(512); product_section(n, m, nn, k, 1 | 2, A, B, C);
    }
    /* 3, 4: 256-bit SYRK A^T A, A 17 x 6, lower only and full */
    {
        const int n = 8; size_t rows = 17, cols = 6;
        for (size_t i = 0; i < rows * cols; ++i) A[i] = band_entry(&g, n, 1);
        for (size_t i = 0; i < cols * cols; ++i) C[i] = lfo_zero(0);
        put_u32(256); product_section(n, cols, cols, rows, 4 | 8, A, NULL, C);
        put_u32(256); product_section(n, cols, cols, rows, 4, A, NULL, C);
    }
    /* 5: the contract-names known answer: [2^100, 1, -2^100] . [1, 1, 1] = 1 (64 bits) */
    {
        const int n = 2;
        A[0] = lfo_pow2(n, 100); A[1] = lfo_pow2(n, 0); A[2] = lfo_neg(lfo_pow2(n, 100));
        for (int i = 0; i < 3; ++i) B[i] = lfo_pow2(n, 0);
        C[0] = lfo_zero(0);
        put_u32(64); product_section(n, 1, 1, 3, 0, A, B, C);
    }
    /* 6: statuses, a zero line and an update of a status entry (96 bits) */
    {
        const int n = 3; size_t m = 3, nn = 3, k = 4;
        for (size_t i = 0; i < m * k; ++i) A[i] = lfo_random_number(&g, n, 30);
        for (size_t i = 0; i < k * nn; ++i) B[i] = lfo_random_number(&g, n, 30);
        for (size_t kk = 0; kk < k; ++kk) A[1 * k + kk] = lfo_zero(0); /* zero row 1 */
        B[2 * nn + 2].status = LFO_INVALID;                               /* status in column 2 */
        for (size_t i = 0; i < m * nn; ++i) C[i] = lfo_random_number(&g, n, 30);
        C[0].status = LFO_EXPONENT_OVERFLOW;
        put_u32(96); product_section(n, m, nn, k, 2, A, B, C);
    }
    /* 7: K = 0 plain and update (128 bits) */
    {
        const int n = 4; size_t m = 2, nn = 3;
        for (size_t i = 0; i < m * nn; ++i) C[i] = lfo_random_number(&g, n, 30);
        put_u32(128); product_section(n, m, nn, 0, 2, A, B, C);
    }
    /* 8: 512-bit lines spread far beyond the band limits (the exact fallback class) */
    {
        const int n = 16; size_t m = 3, nn = 4, k = 9;
        for (size_t i = 0; i < m * k; ++i) { A[i] = lfo_random_number(&g, n, 0); A[i].exponent = (int)(lfo_mt_next(&g) % 4001) - 2000; }
        for (size_t i = 0; i < k * nn; ++i) { B[i] = lfo_random_number(&g, n, 0); B[i].exponent = (int)(lfo_mt_next(&g) % 4001) - 2000; }
        for (size_t i = 0; i < m * nn; ++i) C[i] = lfo_zero(0);
        put_u32(512); product_section(n, m, nn, k, 0, A, B, C);
    }
    fclose(out_file);
}

int main(int argc, char** argv)
{
    const char* dir = argc > 1 ? argv[1] : "tests/golden";
    lfo_mpfr_range();
    static const int widths[] = {64, 96, 256, 384, 512, 1024};
    for (int w = 0; w < 6; ++w) elementwise_file(dir, widths[w]);
    chain_file(dir);
    dot_file(dir, 128);
    dot_file(dir, 512);
    product_file(dir);
    printf("golden fixtures written to %s\n", dir);
    return 0;
}

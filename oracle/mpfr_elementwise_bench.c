/*
 * mpfr_elementwise_bench — the reference benchmark's MPFR CPU baseline for element-wise operations
 * (benchmarks/benchmark.cpp:16-31 and benchmark_support.hpp:12-26): div, fma (coefficient b broadcast
 * with period 64), sqrt of |x|, and the composed complex division of core.hpp:416-419
 * (d = br^2 + bi^2, re = (ar br + ai bi) / d, im = (ai br - ar bi) / d, every operation rounded), on
 * preallocated mpfr_t arrays, serially and over `threads` workers (contiguous chunks).  Operands as
 * config 0 with exponents uniform in [-512, 512].  Setup is not timed; 2 warm-ups, then the median of
 * `samples` runs.  Prints one JSON object.  TEST/MEASUREMENT INFRASTRUCTURE ONLY (bench.py).
 *
 * usage: mpfr_elementwise_bench <bits> <div|fma|sqrt|cdiv> <count> <threads> [samples]
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "lf_oracle.h"
#include "lfo_cases.h"
#include "lfo_mpfr.h"

enum { DIV, FMA, SQRT, CDIV };
typedef struct { int op, bits; mpfr_t *a, *b, *c, *ai, *bi, *r, *ri; size_t lo, hi; } job_t;

static void* work(void* p)
{
    job_t* j = (job_t*)p;
    mpfr_t t0, t1, den;
    mpfr_inits2(j->bits, t0, t1, den, (mpfr_ptr)0);
    for (size_t i = j->lo; i < j->hi; ++i) {
        switch (j->op) {
        case DIV: mpfr_div(j->r[i], j->a[i], j->b[i], MPFR_RNDN); break;
        case FMA: mpfr_fma(j->r[i], j->a[i], j->b[i % 64], j->c[i], MPFR_RNDN); break;
        case SQRT: mpfr_sqrt(j->r[i], j->a[i], MPFR_RNDN); break;
        default:
            mpfr_mul(t0, j->b[i], j->b[i], MPFR_RNDN);
            mpfr_mul(t1, j->bi[i], j->bi[i], MPFR_RNDN);
            mpfr_add(den, t0, t1, MPFR_RNDN);
            mpfr_mul(t0, j->a[i], j->b[i], MPFR_RNDN);
            mpfr_mul(t1, j->ai[i], j->bi[i], MPFR_RNDN);
            mpfr_add(j->r[i], t0, t1, MPFR_RNDN);
            mpfr_div(j->r[i], j->r[i], den, MPFR_RNDN);
            mpfr_mul(t0, j->ai[i], j->b[i], MPFR_RNDN);
            mpfr_mul(t1, j->a[i], j->bi[i], MPFR_RNDN);
            mpfr_sub(j->ri[i], t0, t1, MPFR_RNDN);
            mpfr_div(j->ri[i], j->ri[i], den, MPFR_RNDN);
        }
    }
    mpfr_clears(t0, t1, den, (mpfr_ptr)0);
    return NULL;
}

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

static mpfr_t* array(size_t count, int bits, lfo_mt64* g, int n, int positive)
{
    mpfr_t* v = malloc(count * sizeof(mpfr_t));
    for (size_t i = 0; i < count; ++i) {
        lfo_num x = lfo_random_number(g, n, 512);
        if (positive) x.sign = 1;
        mpfr_init2(v[i], bits);
        lfo_to_mpfr(n, v[i], &x);
    }
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 5) { fprintf(stderr, "usage: mpfr_elementwise_bench <bits> <div|fma|sqrt|cdiv> <count> <threads> [samples]\n"); return 2; }
    const int bits = atoi(argv[1]), n = bits / 32, threads = atoi(argv[4]);
    const int op = !strcmp(argv[2], "div") ? DIV : !strcmp(argv[2], "fma") ? FMA : !strcmp(argv[2], "sqrt") ? SQRT : CDIV;
    const size_t count = (size_t)atoll(argv[3]);
    const int samples = argc > 5 ? atoi(argv[5]) : 5;
    lfo_mpfr_range();
    lfo_mt64 g;
    lfo_mt_seed(&g, 20261007u + (uint64_t)bits);
    job_t base = {op, bits, array(count, bits, &g, n, op == SQRT), array(count, bits, &g, n, 0), NULL, NULL, NULL, array(count, bits, &g, n, 0),
                  NULL, 0, count};
    if (op == FMA) base.c = array(count, bits, &g, n, 0);
    if (op == CDIV) {
        base.ai = array(count, bits, &g, n, 0);
        base.bi = array(count, bits, &g, n, 0);
        base.ri = array(count, bits, &g, n, 0);
    }
    double* ts = malloc(sizeof(double) * (samples + 2));
    double* tp = malloc(sizeof(double) * (samples + 2));
    pthread_t* th = malloc(sizeof(pthread_t) * threads);
    job_t* jobs = malloc(sizeof(job_t) * threads);
    for (int rep = 0; rep < samples + 2; ++rep) {
        double t0 = now();
        work(&base);
        ts[rep] = now() - t0;
        t0 = now();
        for (int k = 0; k < threads; ++k) {
            jobs[k] = base;
            jobs[k].lo = count * k / threads;
            jobs[k].hi = count * (k + 1) / threads;
            pthread_create(&th[k], NULL, work, &jobs[k]);
        }
        for (int k = 0; k < threads; ++k) pthread_join(th[k], NULL);
        tp[rep] = now() - t0;
    }
    qsort(ts + 2, samples, sizeof(double), cmp_d);
    qsort(tp + 2, samples, sizeof(double), cmp_d);
    printf("{\"mpfr\": \"%s\", \"bits\": %d, \"op\": \"%s\", \"count\": %zu, \"threads\": %d, \"serial_s\": %.6f, \"parallel_s\": %.6f}\n",
           mpfr_get_version(), bits, argv[2], count, threads, ts[2 + samples / 2], tp[2 + samples / 2]);
    return 0;
}

/*
 * mpfr_chain_bench — the reference benchmark's MPFR CPU baseline for the multiplication chain
 * (benchmarks/benchmark.cpp:67-94, benchmarks/benchmark_support.hpp:12-26): every value is multiplied
 * `steps` times by its own factor with mpfr_mul(RNDN), serially and over a pool of `threads` workers
 * (contiguous chunks), on preallocated mpfr_t arrays.  Conversion and setup are not timed; 2 warm-ups,
 * then the median of `samples` runs.  Prints one JSON object.  TEST/MEASUREMENT INFRASTRUCTURE ONLY
 * (run by bench.py's cpu_baseline leg).
 *
 * usage: mpfr_chain_bench <bits> <count> <steps> <threads> [samples]
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

typedef struct { mpfr_t* r; mpfr_t* f; size_t lo, hi; int steps; } job_t;

static void* work(void* p)
{
    job_t* j = (job_t*)p;
    for (int s = 0; s < j->steps; ++s)
        for (size_t i = j->lo; i < j->hi; ++i) mpfr_mul(j->r[i], j->r[i], j->f[i], MPFR_RNDN);
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

int main(int argc, char** argv)
{
    if (argc < 5) { fprintf(stderr, "usage: mpfr_chain_bench <bits> <count> <steps> <threads> [samples]\n"); return 2; }
    const int bits = atoi(argv[1]), n = bits / 32, steps = atoi(argv[3]), threads = atoi(argv[4]);
    const size_t count = (size_t)atoll(argv[2]);
    const int samples = argc > 5 ? atoi(argv[5]) : 5;
    lfo_mpfr_range();
    lfo_mt64 g;
    lfo_mt_seed(&g, 20261007u + (uint64_t)bits);
    mpfr_t *a = malloc(count * sizeof(mpfr_t)), *f = malloc(count * sizeof(mpfr_t)), *r = malloc(count * sizeof(mpfr_t));
    for (size_t i = 0; i < count; ++i) {
        lfo_num x = lfo_random_number(&g, n, 64), y = lfo_random_number(&g, n, 64);
        mpfr_inits2(bits, a[i], f[i], r[i], (mpfr_ptr)0);
        lfo_to_mpfr(n, a[i], &x);
        lfo_to_mpfr(n, f[i], &y);
    }
    double* ts = malloc(sizeof(double) * (samples + 2));
    double* tp = malloc(sizeof(double) * (samples + 2));
    pthread_t* th = malloc(sizeof(pthread_t) * threads);
    job_t* jobs = malloc(sizeof(job_t) * threads);
    for (int rep = 0; rep < samples + 2; ++rep) {
        for (size_t i = 0; i < count; ++i) mpfr_set(r[i], a[i], MPFR_RNDN);
        job_t all = {r, f, 0, count, steps};
        double t0 = now();
        work(&all);
        ts[rep] = now() - t0;
        for (size_t i = 0; i < count; ++i) mpfr_set(r[i], a[i], MPFR_RNDN);
        t0 = now();
        for (int k = 0; k < threads; ++k) {
            jobs[k] = (job_t){r, f, count * k / threads, count * (k + 1) / threads, steps};
            pthread_create(&th[k], NULL, work, &jobs[k]);
        }
        for (int k = 0; k < threads; ++k) pthread_join(th[k], NULL);
        tp[rep] = now() - t0;
    }
    qsort(ts + 2, samples, sizeof(double), cmp_d);
    qsort(tp + 2, samples, sizeof(double), cmp_d);
    printf("{\"bits\": %d, \"count\": %zu, \"steps\": %d, \"threads\": %d, \"samples\": %d, \"mpfr\": \"%s\", "
           "\"serial_s\": %.6f, \"parallel_s\": %.6f}\n",
           bits, count, steps, threads, samples, mpfr_get_version(), ts[2 + samples / 2], tp[2 + samples / 2]);
    return 0;
}

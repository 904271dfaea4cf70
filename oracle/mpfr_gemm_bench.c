/*
 * mpfr_gemm_bench — the MPFR CPU baseline of the exact dense product C = A B (the 512-bit GEMM workload).
 * TEST/MEASUREMENT INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg).
 *
 * Two CPU methods on the same operands, each timed serially and on `threads` workers (rows handed out
 * one at a time), each on a bounded sample of output rows (rows taken in a fixed spread order until
 * `budget` seconds have passed), reported per output:
 *   exact:    every product exact at 2*bits (mpfr_mul), then one correctly rounded mpfr_sum per output --
 *             the reference's MPFR oracle for these products (tests/reference.hpp:74-82); its results are
 *             the same bits as the GPU's and are written to `out` for comparison;
 *   consumer: s = RN(s + RN(a b)) with mpfr_mul / mpfr_add at `bits` (the reference benchmark's CPU MPFR loop,
 *             benchmarks/linalg.mm:57-64): not exact, timed only.
 * Operands are converted to mpfr_t before timing.  Prints one JSON object.
 *
 * usage: mpfr_gemm_bench <in.bin> <threads> <budget_seconds> [out.bin]
 * in.bin: uint32 bits, m, n, k, then A (m x k) and B (k x n) as records (n+3 words each, row-major).
 * out.bin: uint32 rows, then per row its index and the n exact results.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "lf_oracle.h"
#include "lfo_mpfr.h"

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

typedef struct {
    int bits, nw;
    size_t m, n, k;
    mpfr_t *a, *b;           /* a[i*k + q], b[q*n + j] */
    size_t* order;           /* sampled row order */
    int exact;               /* method */
    double deadline;
    size_t next;             /* next position in order (under lock) */
    pthread_mutex_t lock;
    size_t rows_done;
    lfo_num* results;        /* [position in order][n] for the exact method, may be NULL */
} job_t;

static void do_row(job_t* J, size_t pos, mpfr_t* p, mpfr_ptr* terms, mpfr_t s, mpfr_t t)
{
    const size_t i = J->order[pos];
    for (size_t j = 0; j < J->n; ++j) {
        if (J->exact) {
            for (size_t q = 0; q < J->k; ++q) mpfr_mul(p[q], J->a[i * J->k + q], J->b[q * J->n + j], MPFR_RNDN);
            mpfr_sum(s, terms, J->k, MPFR_RNDN);
            if (J->results) J->results[pos * J->n + j] = lfo_from_mpfr(J->nw, s);
        } else {
            mpfr_set_zero(s, 1);
            for (size_t q = 0; q < J->k; ++q) {
                mpfr_mul(t, J->a[i * J->k + q], J->b[q * J->n + j], MPFR_RNDN);
                mpfr_add(s, s, t, MPFR_RNDN);
            }
        }
    }
}

static void* worker(void* arg)
{
    job_t* J = (job_t*)arg;
    mpfr_t s, t;
    mpfr_inits2(J->bits, s, t, (mpfr_ptr)0);
    mpfr_t* p = malloc(J->k * sizeof(mpfr_t));
    mpfr_ptr* terms = malloc(J->k * sizeof(mpfr_ptr));
    for (size_t q = 0; q < J->k; ++q) {
        mpfr_init2(p[q], 2 * J->bits);
        terms[q] = p[q];
    }
    for (;;) {
        pthread_mutex_lock(&J->lock);
        size_t pos = J->next;
        int go = pos < J->m && now() < J->deadline;
        if (go) J->next++;
        pthread_mutex_unlock(&J->lock);
        if (!go) break;
        do_row(J, pos, p, terms, s, t);
        pthread_mutex_lock(&J->lock);
        J->rows_done++;
        pthread_mutex_unlock(&J->lock);
    }
    for (size_t q = 0; q < J->k; ++q) mpfr_clear(p[q]);
    free(p);
    free(terms);
    mpfr_clears(s, t, (mpfr_ptr)0);
    return NULL;
}

/* seconds per output of one method on `threads` workers */
static double timed(job_t* J, int threads, double budget, size_t* rows)
{
    J->next = 0;
    J->rows_done = 0;
    const double t0 = now();
    J->deadline = t0 + budget;
    pthread_t* th = malloc(sizeof(pthread_t) * threads);
    for (int q = 0; q < threads; ++q) pthread_create(&th[q], NULL, worker, J);
    for (int q = 0; q < threads; ++q) pthread_join(th[q], NULL);
    free(th);
    const double el = now() - t0;
    *rows = J->rows_done;
    return J->rows_done ? el / (double)(J->rows_done * J->n) : 0.0;
}

int main(int argc, char** argv)
{
    if (argc < 4) { fprintf(stderr, "usage: mpfr_gemm_bench <in.bin> <threads> <budget_seconds> [out.bin]\n"); return 2; }
    const int threads = atoi(argv[2]);
    const double budget = atof(argv[3]);
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t hdr[4];
    if (fread(hdr, 4, 4, f) != 4) return 3;
    job_t J;
    memset(&J, 0, sizeof J);
    J.bits = (int)hdr[0];
    J.nw = J.bits / 32;
    J.m = hdr[1];
    J.n = hdr[2];
    J.k = hdr[3];
    const size_t rw = (size_t)J.nw + 3;
    uint32_t* A = malloc(J.m * J.k * rw * 4);
    uint32_t* B = malloc(J.k * J.n * rw * 4);
    if (fread(A, 4, J.m * J.k * rw, f) != J.m * J.k * rw || fread(B, 4, J.k * J.n * rw, f) != J.k * J.n * rw) return 3;
    fclose(f);
    lfo_mpfr_range();
    J.a = malloc(J.m * J.k * sizeof(mpfr_t));
    J.b = malloc(J.k * J.n * sizeof(mpfr_t));
    lfo_num x;
    for (size_t q = 0; q < J.m * J.k; ++q) {
        mpfr_init2(J.a[q], J.bits);
        lfo_load(J.nw, A + q * rw, &x);
        lfo_to_mpfr(J.nw, J.a[q], &x);
    }
    for (size_t q = 0; q < J.k * J.n; ++q) {
        mpfr_init2(J.b[q], J.bits);
        lfo_load(J.nw, B + q * rw, &x);
        lfo_to_mpfr(J.nw, J.b[q], &x);
    }
    /* rows spread over the matrix: a stride coprime to m */
    J.order = malloc(J.m * sizeof(size_t));
    size_t step = J.m > 7 ? 7 : 1;
    for (;;) { /* a stride coprime to m makes the order a permutation */
        size_t a = J.m, b = step;
        while (b) { size_t t = a % b; a = b; b = t; }
        if (a == 1) break;
        step += 2;
    }
    for (size_t q = 0; q < J.m; ++q) J.order[q] = (q * step + step / 2) % J.m;
    pthread_mutex_init(&J.lock, NULL);
    J.results = malloc(J.m * J.n * sizeof(lfo_num));
    size_t rows_es, rows_ep, rows_cs, rows_cp;
    J.exact = 1;
    const double es = timed(&J, 1, budget, &rows_es);
    /* keep the serial exact rows for the comparison; the parallel run recomputes from the start */
    lfo_num* keep = malloc(rows_es * J.n * sizeof(lfo_num));
    memcpy(keep, J.results, rows_es * J.n * sizeof(lfo_num));
    const double ep = timed(&J, threads, budget, &rows_ep);
    J.exact = 0;
    J.results = NULL;
    const double cs = timed(&J, 1, budget / 2, &rows_cs);
    const double cp = timed(&J, threads, budget / 2, &rows_cp);
    if (argc > 4) {
        FILE* g = fopen(argv[4], "wb");
        uint32_t r = (uint32_t)rows_es;
        fwrite(&r, 4, 1, g);
        uint32_t* rec = malloc(rw * 4);
        for (size_t pos = 0; pos < rows_es; ++pos) {
            uint32_t idx = (uint32_t)J.order[pos];
            fwrite(&idx, 4, 1, g);
            for (size_t j = 0; j < J.n; ++j) {
                lfo_store(J.nw, &keep[pos * J.n + j], rec);
                fwrite(rec, 4, rw, g);
            }
        }
        fclose(g);
    }
    printf("{\"bits\": %d, \"m\": %zu, \"n\": %zu, \"k\": %zu, \"threads\": %d, \"mpfr\": \"%s\", "
           "\"exact_serial_s_per_output\": %.9g, \"exact_serial_rows\": %zu, \"exact_parallel_s_per_output\": %.9g, \"exact_parallel_rows\": %zu, "
           "\"consumer_serial_s_per_output\": %.9g, \"consumer_serial_rows\": %zu, \"consumer_parallel_s_per_output\": %.9g, "
           "\"consumer_parallel_rows\": %zu}\n",
           J.bits, J.m, J.n, J.k, threads, mpfr_get_version(), es, rows_es, ep, rows_ep, cs, rows_cs, cp, rows_cp);
    return 0;
}

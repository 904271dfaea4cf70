/*
 * mpfr_mixed_bench — the MPFR CPU baseline of the mixed-precision chain batch: value i of precision bits_i
 * runs len_i steps, step t computing x <- RN(x*f + g) (mpfr_fma) when bit t of its mask is set and
 * x <- RN(x*f) (mpfr_mul) otherwise, all at bits_i, on preconverted mpfr_t operands; serially and on
 * `threads` workers (values handed out in blocks of 16).  TEST/MEASUREMENT INFRASTRUCTURE ONLY (run by
 * bench.py's cpu_baseline leg).  Results are written back as records for comparison with the GPU.
 *
 * usage: mpfr_mixed_bench <in.bin> <threads> <repeats> [out.bin]
 * in.bin: uint32 buckets; per bucket: uint32 bits, count; x, f, g records (count each); len uint32[count];
 * mask uint64[count].  out.bin: the results of every bucket in order, as records.
 * Prints one JSON object: median seconds of `repeats` serial and parallel runs over all values.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "lf_oracle.h"
#include "lfo_mpfr.h"

typedef struct {
    int bits;
    mpfr_t x, x0, f, g;
    uint32_t len;
    uint64_t mask;
} value_t;

typedef struct {
    value_t* v;
    size_t count, next;
    pthread_mutex_t lock;
} job_t;

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void run_value(value_t* v)
{
    mpfr_set(v->x, v->x0, MPFR_RNDN);
    for (uint32_t t = 0; t < v->len; ++t) {
        if ((v->mask >> t) & 1) mpfr_fma(v->x, v->x, v->f, v->g, MPFR_RNDN);
        else mpfr_mul(v->x, v->x, v->f, MPFR_RNDN);
    }
}

static void* worker(void* arg)
{
    job_t* J = (job_t*)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        size_t lo = J->next;
        J->next = lo + 16 < J->count ? lo + 16 : J->count;
        size_t hi = J->next;
        pthread_mutex_unlock(&J->lock);
        if (lo >= hi) break;
        for (size_t i = lo; i < hi; ++i) run_value(&J->v[i]);
    }
    return NULL;
}

static double run_all(job_t* J, int threads)
{
    J->next = 0;
    const double t0 = now();
    pthread_t* th = malloc(sizeof(pthread_t) * threads);
    for (int q = 0; q < threads; ++q) pthread_create(&th[q], NULL, worker, J);
    for (int q = 0; q < threads; ++q) pthread_join(th[q], NULL);
    free(th);
    return now() - t0;
}

static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

int main(int argc, char** argv)
{
    if (argc < 4) { fprintf(stderr, "usage: mpfr_mixed_bench <in.bin> <threads> <repeats> [out.bin]\n"); return 2; }
    const int threads = atoi(argv[2]), repeats = atoi(argv[3]) > 0 ? atoi(argv[3]) : 1;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 3;
    uint32_t nb;
    if (fread(&nb, 4, 1, f) != 1) return 3;
    lfo_mpfr_range();
    value_t* v = NULL;
    size_t total = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t hb[2];
        if (fread(hb, 4, 2, f) != 2) return 3;
        const int bits = (int)hb[0], nw = bits / 32;
        const size_t cnt = hb[1], rw = (size_t)nw + 3;
        uint32_t* rec = malloc(3 * cnt * rw * 4 + 1);
        uint32_t* len = malloc(cnt * 4 + 1);
        uint64_t* mask = malloc(cnt * 8 + 1);
        if (fread(rec, 4, 3 * cnt * rw, f) != 3 * cnt * rw || fread(len, 4, cnt, f) != cnt || fread(mask, 8, cnt, f) != cnt) return 3;
        v = realloc(v, (total + cnt) * sizeof(value_t));
        for (size_t i = 0; i < cnt; ++i) {
            value_t* w = &v[total + i];
            lfo_num x;
            w->bits = bits;
            w->len = len[i];
            w->mask = mask[i];
            mpfr_inits2(bits, w->x, w->x0, w->f, w->g, (mpfr_ptr)0);
            lfo_load(nw, rec + i * rw, &x);
            lfo_to_mpfr(nw, w->x0, &x);
            lfo_load(nw, rec + (cnt + i) * rw, &x);
            lfo_to_mpfr(nw, w->f, &x);
            lfo_load(nw, rec + (2 * cnt + i) * rw, &x);
            lfo_to_mpfr(nw, w->g, &x);
        }
        total += cnt;
        free(rec);
        free(len);
        free(mask);
    }
    fclose(f);
    job_t J = {v, total, 0, PTHREAD_MUTEX_INITIALIZER};
    double* ts = malloc(sizeof(double) * repeats);
    double* tp = malloc(sizeof(double) * repeats);
    for (int r = 0; r < repeats; ++r) {
        ts[r] = run_all(&J, 1);
        tp[r] = run_all(&J, threads);
    }
    qsort(ts, repeats, sizeof(double), cmp_d);
    qsort(tp, repeats, sizeof(double), cmp_d);
    if (argc > 4) {
        FILE* g = fopen(argv[4], "wb");
        for (size_t i = 0; i < total; ++i) {
            const int nw = v[i].bits / 32;
            uint32_t rec[35 + 1];
            lfo_num y = lfo_from_mpfr(nw, v[i].x);
            lfo_store(nw, &y, rec);
            fwrite(rec, 4, (size_t)nw + 3, g);
        }
        fclose(g);
    }
    printf("{\"values\": %zu, \"threads\": %d, \"repeats\": %d, \"mpfr\": \"%s\", \"serial_s\": %.6f, \"parallel_s\": %.6f}\n", total,
           threads, repeats, mpfr_get_version(), ts[repeats / 2], tp[repeats / 2]);
    return 0;
}

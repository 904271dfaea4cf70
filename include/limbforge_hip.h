/*
 * limbforge_hip.h — the C ABI of liblimbforge_hip.so, the MI355X (gfx950) implementation of the
 * WolfNum/LimbForge data-parallel core.
 *
 * The reference exposes this path only as a C++17 API (include/limbforge/engine.hpp and linalg.hpp);
 * it has no C ABI of its own.  Each entry point below names the reference interface it replaces, so a
 * C++ facade (the headers under this repository's include/limbforge/) or any FFI (ctypes, cgo, JNI) can drive the
 * same kernels.  The ABI uses plain pointers, sizes and integer status codes: no C++ types, no
 * exceptions.  Arithmetic errors are never reported here: they are status bits inside the results
 * (division by zero 1, exponent overflow 2, invalid 4), as in the reference (docs/numerics.md:14-31).
 *
 * Element formats (engine.hpp:195): a real value of `bits` (multiple of 32 in [64, 1024]) is
 * n = bits/32 little-endian uint32 limbs, int32 exponent, int32 sign, uint32 status (4n+12 bytes);
 * a complex value is two of them (re, im).  Host arrays use that layout.  Device operands are
 * "limb planes": n limb planes, then exponent, sign and status planes (a complex operand: the real
 * planes, then the imaginary planes), each plane `pitch` uint32 elements long.
 *
 * Return codes: LF_OK, or one of the error classes below with a message in lf_last_error() (per
 * host thread).  The C++ facade maps them to std::invalid_argument / std::logic_error /
 * std::runtime_error as the reference throws them (engine.mm:20-28, linalg.mm:286,307-308).
 */
#ifndef LIMBFORGE_HIP_H
#define LIMBFORGE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LF_OK 0
#define LF_INVALID_ARGUMENT 1
#define LF_LOGIC_ERROR 2
#define LF_RUNTIME_ERROR 3

/* Operation numbering of limbforge::Operation (engine.hpp:13).  complex_fma / complex_fms compute each part
 * of a*b + c (a*b - c) with one rounding of its exact value (core.hpp:330-358). */
enum {
    LF_OP_ADD = 0, LF_OP_SUB = 1, LF_OP_MUL = 2, LF_OP_DIV = 3, LF_OP_COMPLEX_ADD = 4, LF_OP_COMPLEX_MUL = 5,
    LF_OP_COMPLEX_DIV = 6, LF_OP_SQUARE = 7, LF_OP_SQRT = 8, LF_OP_FMA = 9, LF_OP_FMS = 10, LF_OP_COMPLEX_FMA = 11,
    LF_OP_COMPLEX_FMS = 12
};

/* A device operand: limb planes and their pitch (uint32 elements per plane). */
typedef struct lf_operand {
    void* planes;
    uint64_t pitch;
} lf_operand;

/* Broadcast read: element i uses index (i / stride) % period; period 0 = no wrap (engine.hpp:19-25). */
typedef struct lf_broadcast {
    uint64_t stride;
    uint64_t period;
} lf_broadcast;

/* Message of the last failing call on this host thread (empty string if none). */
const char* lf_last_error(void);

/* Library version; replaces limbforge::library_version (version.hpp:31, src/version.cpp). */
void lf_library_version(unsigned* major, unsigned* minor, unsigned* patch);

/* ---- device, memory, streams: the HIP runtime behind Engine (engine.hpp:184-227), Buffer<T>
 * (engine.hpp:74-86), CommandBatch and Submission (engine.hpp:93-182) ---- */
int lf_device_count(int* count);
int lf_device_set(int device);
/* Engine::device_name (engine.hpp:188) */
int lf_device_name(char* out, size_t capacity);
/* Engine::allocate (engine.hpp:225): device memory for resident operands */
int lf_malloc(void** ptr, size_t bytes);
int lf_free(void* ptr);
/* pinned host staging */
int lf_host_alloc(void** ptr, size_t bytes);
int lf_host_free(void* ptr);
/* direction: 0 host->device, 1 device->host, 2 device->device */
int lf_memcpy_async(void* dst, const void* src, size_t bytes, int direction, void* stream);
int lf_stream_create(void** stream);
int lf_stream_destroy(void* stream);
int lf_stream_synchronize(void* stream);
int lf_event_create(void** event);
int lf_event_destroy(void* event);
int lf_event_record(void* event, void* stream);
int lf_event_synchronize(void* event);
/* 1 when the event has completed (Submission::ready, engine.hpp:101) */
int lf_event_query(void* event, int* done);
int lf_event_elapsed_ms(float* ms, void* start, void* end);
/* Capture the launches issued on `stream` into a graph (a dependent op batch replayed as one
 * submission; CommandBatch::submit, engine.hpp:180). */
int lf_graph_begin_capture(void* stream);
int lf_graph_end_capture(void* stream, void** graph_exec);
int lf_graph_launch(void* graph_exec, void* stream);
int lf_graph_destroy(void* graph_exec);

/* ---- layout: Buffer<T>::upload / download (engine.hpp:84-85) between the public AoS element and
 * limb planes; both arrays are device memory ---- */
uint64_t lf_plane_pitch(uint64_t count);
uint64_t lf_planes_bytes(int bits, int complex_values, uint64_t count);
int lf_aos_to_planes(int bits, int complex_values, const void* aos, const lf_operand* planes, uint64_t count, void* stream);
int lf_planes_to_aos(int bits, int complex_values, const lf_operand* planes, void* aos, uint64_t count, void* stream);

/* ---- compute ---- */
/* One element-wise operation over resident operands: the reference's CommandBatch::run overloads
 * (engine.hpp:120-147) through CommandBatch::encode (engine.mm:332-346) and the kernels
 * arithmetic / complex_arithmetic / fused_arithmetic (kernels.metal:8-33).
 * out[i] = op(a[i], b[(i/bs)%bp], c[(i/cs)%cp]); b is ignored by unary ops, c by non-fused ops;
 * b_index / c_index may be NULL (no broadcast).  Operands may alias out (element-wise). */
int lf_elementwise(int bits, int op, const lf_operand* a, const lf_operand* b, const lf_operand* c,
                   const lf_operand* out, uint64_t count, const lf_broadcast* b_index,
                   const lf_broadcast* c_index, void* stream);

/* The resident multiplication chain x <- RN(x * f[(i/fs)%fp]) repeated `steps` times in one launch:
 * replaces `steps` dependent CommandBatch::run(Operation::mul, x, y, x) calls (engine.hpp:120-124,
 * as benchmarks/benchmark.cpp:84-92 issues them).  out may alias x. */
int lf_mul_chain(int bits, const lf_operand* x, const lf_operand* f, const lf_operand* out, uint64_t count,
                 const lf_broadcast* f_index, int steps, void* stream);
/* The same chain with `lanes` lanes of a wavefront per value (no reference counterpart; the reference
 * spreads values over lanes only in its recurrence kernels, src/cooperative.metal:40-61): 1, or 2 and 4 from
 * 256 bits on (the lanes split each product's digit columns, exchange them inside a quad and round
 * together), bit for bit the one-lane result; 0 takes the choice lf_autotune_chain made for this width
 * and batch size, else 4 lanes from 512 bits for batches of at most 16384 values and one lane otherwise.
 * lf_mul_chain is lanes 0. */
int lf_mul_chain_lanes(int bits, const lf_operand* x, const lf_operand* f, const lf_operand* out, uint64_t count,
                       const lf_broadcast* f_index, int steps, unsigned lanes, void* stream);
/* Times the chain at 1, 2 and 4 lanes per value on a synthetic batch of `count` values and `steps` steps
 * (count <= 2^22, steps <= 1024), records the fastest for this device, width and power-of-two class of the
 * batch size, and returns it in *lanes; ms (may be NULL) receives the three times in milliseconds (-1 where
 * the width has no such form).  No reference counterpart (the project's autotuner). */
int lf_autotune_chain(int bits, uint64_t count, int steps, unsigned* lanes, float* ms);

/* Host-array element-wise call: Engine::run / run_unary / run_ternary (engine.hpp:202-204;
 * engine.mm:254-280, 479-499).  a, b, c, out are host arrays of AoS elements; b holds
 * broadcast_elements(count, b_index) elements (engine.hpp:22-25), likewise c.  Returns after the
 * device has finished; gpu_seconds (device time of the kernels) and wall_seconds may be NULL. */
int lf_run_host(int bits, int op, const void* a, const void* b, const void* c, void* out, uint64_t count,
                const lf_broadcast* b_index, const lf_broadcast* c_index, double* gpu_seconds, double* wall_seconds);

/* No counterpart in the reference.  lf_run_host for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc): the same operation and results, with the device staging in a
 * static per-GPU array of lf_noalloc_scratch_bytes() bytes (64 MiB) held until the process ends; no call
 * allocates or frees device memory.  A batch whose staging (a, b, c and out rounded up to 256 bytes each)
 * exceeds the array is rejected with LF_INVALID_ARGUMENT before any device work.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race.  stream may be NULL (the default stream);
 * returns after the stream has finished. */
int lf_run_host_noalloc(int bits, int op, const void* a, const void* b, const void* c, void* out, uint64_t count,
                        const lf_broadcast* b_index, const lf_broadcast* c_index, void* stream);
/* The size of lf_run_host_noalloc's per-GPU staging array, in bytes. */
int lf_noalloc_scratch_bytes(uint64_t* bytes);

/* Precision cast between resident operands: CommandBatch::cast (engine.hpp:170-176) and the kernels
 * cast_real / cast_complex (kernels.metal:251-266): exact widening, RN-even narrowing, componentwise for
 * complex values; statuses kept, exponent_overflow when rounding leaves the range. */
int lf_cast(int from_bits, int to_bits, int complex_values, const lf_operand* in, const lf_operand* out,
            uint64_t count, void* stream);

/* Host-array precision cast: Engine::cast (engine.hpp:206-207). */
int lf_cast_host(int from_bits, int to_bits, int complex_values, const void* in, void* out, uint64_t count,
                 double* gpu_seconds, double* wall_seconds);

/* ---- reductions ---- */
/* Segmented dot products: Engine::segmented_dot / CommandBatch::segmented_dot (engine.hpp:177-180, 205; kernels
 * segmented_dot_real / segmented_dot_complex, kernels.metal:197-208).  out[i] for i < segments is the fixed
 * sequential chain s <- fma(a[i*length + k], b[i*length + k], s) from s = 0 (b[k] with shared_right; cfma per
 * component for complex values), one rounding per step; length 0 gives zero.  Device form on limb planes
 * (a holds segments*length values, b length or segments*length) and host-array form. */
int lf_segmented_dot(int bits, int complex_values, uint64_t segments, uint64_t length, int shared_right, const lf_operand* a,
                     const lf_operand* b, const lf_operand* out, void* stream);
int lf_segmented_dot_host(int bits, int complex_values, uint64_t segments, uint64_t length, int shared_right, const void* a,
                          const void* b, void* out, double* gpu_seconds, double* wall_seconds);
/* Adjacent-pair tree sum: CommandBatch::tree_sum (engine.hpp:150-153; kernels tree_sum_* and global_tree_sum_*,
 * kernels.metal:61-112).  out (one value) = the fixed tree of rounded additions of pairs (0,1), (2,3), ... with
 * an unpaired last value carried unchanged, level by level; complex values componentwise; count 0 gives
 * canonical zero.  Partial sums use stream-ordered temporary device memory. */
int lf_tree_sum(int bits, int complex_values, const lf_operand* in, uint64_t count, const lf_operand* out, void* stream);

/* ---- exact dense products (include/limbforge/linalg.hpp) ---- */

/* LinalgOptions (linalg.hpp:18-25) plus `method`: 0 = exponent-band residue GEMM on int8 MFMA with CRT
 * rebuild (the reference's algorithm), 1 = the plain exact limb GEMM (every output an exact sum of
 * exact 2N-word products, one rounding): the same results, used as the baseline. */
typedef struct lf_linalg_options {
    int band_bits;     /* G: a band holds exponents [top-G, top] of one line (default 64) */
    int max_bands;     /* lines needing more bands take the exact fallback kernels (default 4) */
    int max_spread;    /* ... as do lines whose nonzero exponents span more than this (default 512) */
    unsigned host_threads; /* band-analysis threads; 0 = hardware concurrency */
    int force_fallback;    /* every output through the exact-dot kernels (testing) */
    int resident_bands;    /* kept for LinalgOptions compatibility */
    int method;
} lf_linalg_options;

typedef struct lf_linalg_pair {
    uint32_t left_band, right_band, rows, cols, moduli;
} lf_linalg_pair;

/* LinalgReport (linalg.hpp:33-45): band histograms per side ([b] = lines with exactly b bands; index 0:
 * zero or status lines; band_slots = max_bands + 1 entries, at most 65), fallback lines and outputs,
 * band pairs (first 64), extended GEMM sizes, moduli and int8 GEMMs (three per modulus), stage times
 * (host analysis, uploads, device time, assembly, and the device time per kernel stage). */
typedef struct lf_linalg_report {
    uint64_t m, n, k;
    uint32_t band_slots;
    uint64_t left_bands[65], right_bands[65];
    uint64_t left_fallback, right_fallback;
    uint64_t gpu_outputs, multi_band_outputs, fallback_outputs, trivial_outputs;
    uint32_t npairs;
    lf_linalg_pair pairs[64];
    uint64_t gemm_rows, gemm_cols, int8_gemms, moduli;
    uint32_t crt_words;
    double analysis_seconds, upload_seconds, gpu_seconds, assembly_seconds, fallback_seconds;
    /* device time of each stage inside gpu_seconds (no reference counterpart): residue digits of both
     * operands, the int8-MFMA residue products, the CRT rebuild and rounding, the exact-dot kernels */
    double digits_seconds, product_seconds, rebuild_seconds, dot_seconds;
} lf_linalg_report;

void lf_linalg_default_options(lf_linalg_options* options);

/* C (m x n) = RN(op(A) B) or, with subtract, RN(C - op(A) B), one rounding per entry: Linalg::gemm
 * (linalg.hpp:153-155; linalg.mm:300-407).  Host arrays, row-major; op(A) = A (m x k) or A^T (A is k x m).
 * Statuses: an output whose row of op(A) or column of B carries a status is zero with their OR (and the
 * old entry's for an update); a zero line gives canonical zero (an update leaves C unchanged).
 * options and report may be NULL. */
int lf_gemm_host(int bits, int transpose_a, const void* A, const void* B, uint64_t m, uint64_t n, uint64_t k, void* C,
                 int subtract, const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds,
                 double* wall_seconds);

/* C (cols x cols) = A^T A for A (rows x cols): Linalg::syrk (linalg.hpp:149-152).  lower_only writes
 * C[i][j] for i >= j only; otherwise the full symmetric matrix.  subtract requires lower_only. */
int lf_syrk_host(int bits, const void* A, uint64_t rows, uint64_t cols, void* C, int lower_only, int subtract,
                 const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, double* wall_seconds);

/* The same products on device operands (the resident Linalg::gemm / syrk on Buffer operands, linalg.hpp:187-190;
 * src/linalg.mm:849-1027): A, B and C are limb planes holding the row-major matrices of the host calls, and the
 * results equal the host-array calls bit for bit, the report included.  The band analysis runs on the device and
 * only its line table (one record per row of op(A) and column of B) is read back to plan the launches; every
 * output, fallback lines included, is computed on the device.  Returns after the product has finished on
 * `stream` (work queued earlier on it comes first).  C must not overlap A or B. */
int lf_gemm(int bits, int transpose_a, const lf_operand* A, const lf_operand* B, uint64_t m, uint64_t n, uint64_t k, const lf_operand* C,
            int subtract, const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, void* stream);
int lf_syrk(int bits, const lf_operand* A, uint64_t rows, uint64_t cols, const lf_operand* C, int lower_only, int subtract,
            const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, void* stream);

/* ---- blocked Cholesky and triangular solves (linalg.hpp:73-87, 157-163; src/linalg.mm:411-486) ----
 * The rounding sequence is fixed by `block` alone (docs/numerics.md:687-727): every entry takes
 * blk(j) + 2 roundings.  FactorOptions' host_macs, solve_host_macs and gpu choose the reference's
 * GPU/host placement of each update; placement never changes a bit, and here every step runs on the
 * GPU.  host_macs and gpu are accepted and not used; solve_host_macs picks the form of an off-block
 * solve update: direct exact dots below it or up to 2^22 multiply-add terms, else the update GEMM
 * (a negative value sends every such update to the update GEMM).  Both forms round the same exact
 * value once. */
typedef struct lf_factor_options {
    uint64_t block;         /* column block nb (default 32); 0 = one block */
    double host_macs;       /* default 5e4 */
    double solve_host_macs; /* default 4e5 */
    int gpu;                /* default 1 */
} lf_factor_options;

/* CholeskyInfo (linalg.hpp:82-87): pivot = n on success, else the first column p whose pivot
 * s_pp = RN(a_pp - sum_k l_pk^2) is zero, negative or carries a status; then pivot_sign is its sign
 * (0 for zero or a status) and pivot_status its status, columns j < p of L are final, and every lower
 * entry with j >= p is zero with status invalid. */
typedef struct lf_cholesky_info {
    uint64_t pivot;
    int pivot_sign;
    uint32_t pivot_status;
    double timing_gpu_seconds, timing_wall_seconds;
    uint64_t blocks, gpu_updates, host_updates;
    double panel_seconds, update_seconds;
} lf_cholesky_info;

void lf_factor_default_options(lf_factor_options* options);

/* Linalg::cholesky on device operands (Linalg::cholesky(const Buffer&, n, Buffer&), linalg.hpp:196-197):
 * L = lower factor of the lower triangle of A (n x n limb planes, element i*n + j); the strict upper
 * triangle of L is set to zero; L may be A.  linalg: options of the update products (NULL = defaults).
 * Returns after the device has finished. */
int lf_cholesky(int bits, const lf_operand* A, uint64_t n, const lf_operand* L, const lf_factor_options* options,
                const lf_linalg_options* linalg, lf_cholesky_info* info, void* stream);
/* Linalg::cholesky on host arrays (linalg.hpp:157-159). */
int lf_cholesky_host(int bits, const void* A, uint64_t n, void* L, const lf_factor_options* options, const lf_linalg_options* linalg,
                     lf_cholesky_info* info);
/* Triangular solves with lower L (strict upper triangle not read), X and B n x nrhs: passes 1 = L X = B
 * (Linalg::trsm), 2 = L^T X = B (trsm with transpose), 3 = both, forward first (Linalg::cholesky_solve);
 * linalg.hpp:160-163, src/linalg.mm:459-486.  X may be B.  Device form and host-array form. */
int lf_solve(int bits, int passes, const lf_operand* L, uint64_t n, const lf_operand* B, uint64_t nrhs, const lf_operand* X,
             const lf_factor_options* options, const lf_linalg_options* linalg, double* gpu_seconds, double* wall_seconds, void* stream);
int lf_solve_host(int bits, int passes, const void* L, uint64_t n, const void* B, uint64_t nrhs, void* X, const lf_factor_options* options,
                  const lf_linalg_options* linalg, double* gpu_seconds, double* wall_seconds);

/* ---- mixed-precision chain batches (no reference counterpart: the reference dispatches one precision
 * per call, README.md:212) ----
 * A bucket holds `count` values of one width as limb planes (x, f, g, out with one pitch) and, in device
 * memory, each value's chain length len[i] (0..64) and step mask: step t computes x <- RN(x * f + g)
 * (Operation::fma, core.hpp:317-326) when bit t of mask[i] is set and x <- RN(x * f) (Operation::mul)
 * otherwise, every step rounded once, as a dependent CommandBatch::run sequence would. */
typedef struct lf_chain_bucket {
    int bits;
    lf_operand x, f, g, out;
    const uint32_t* len;
    const uint64_t* mask;
    uint64_t count;
} lf_chain_bucket;
/* One launch per bucket, in order, on `stream`.  block[i]: lanes per workgroup of bucket i (64, 128 or
 * 256; NULL or 0 = the autotuned size of that width, else 256). */
int lf_mixed_chains(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, void* stream);
/* One hipGraph with one persistent kernel node per group of eight widths (2-8, 9-16, 17-24, 25-32 limbs) and
 * no edges between them; each node runs all of its group's buckets (launch with lf_graph_launch, free with
 * lf_graph_destroy).  The graph keeps the buckets' device pointers.  One lane per value unless
 * lf_autotune_mixed_lanes has run on this device. */
int lf_mixed_graph_create(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, void** graph_exec);
/* The same graph with group_lanes[k] lanes per value in group k (0 or 1 for the 2-8 and 9-16-limb groups;
 * 0, 1, 2 or 4 for 17-24 and 25-32 limbs: the lanes split each step's product columns, as
 * lf_mul_chain_lanes does; 0 = 1).
 * group_lanes NULL takes the choice of lf_autotune_mixed_lanes on this device, else one lane everywhere
 * (what lf_mixed_graph_create does).  Results are the same bit for bit whatever the lanes. */
int lf_mixed_graph_create_lanes(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, const unsigned* group_lanes,
                                void** graph_exec);
/* Lanes tuner: times the batch's graph (writing the buckets' out planes, with the results the batch
 * computes) from one lane everywhere, trying 1, 2 and 4 lanes per value in the 25-32-limb group, then in
 * the 17-24-limb group with the first choice held; keeps the fastest for this device and returns it in group_lanes[4]
 * and its time in *best_ms (may be NULL). */
int lf_autotune_mixed_lanes(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, unsigned* group_lanes, float* best_ms);
/* Launch-shape autotuner: times the bucket kernel of `bits` for each workgroup size on `sample` synthetic
 * values (32-step chains, half fma) and keeps the fastest for this device and width. */
int lf_autotune_mixed(int bits, uint64_t sample, unsigned* block, double* best_ms);

/* ---- measurement support (no reference counterpart) ---- */
/* Sustained v_mad_u64_u32 rate of the device (32x32->64 word products per second), measured with
 * `blocks` workgroups of 256 lanes, each lane running `iters` rounds of eight independent chains. */
int lf_calibrate_mad64(uint64_t blocks, int iters, double* products_per_second);
/* One copy of `bytes` (a multiple of 16) between two fresh device buffers; the rate counts bytes read plus
 * bytes written.  Profilers' FETCH_SIZE / WRITE_SIZE of this kernel against its known size scale those
 * counters for the traffic of the other kernels. */
int lf_calibrate_copy(uint64_t bytes, double* bytes_per_second);
/* Cycles (s_memtime ticks) of the factorization's per-row steps at `bits` on one wavefront with the GPU
 * otherwise idle, each averaged over `reps` dependent repetitions: cycles[0] an exact dot of 16 terms,
 * [1] of 31 terms, [2] a division by a precomputed reciprocal, [3] a Knuth division, [4] a square root,
 * [5] a reciprocal (synthetic normalized operands). */
int lf_factor_step_cycles(int bits, int reps, uint64_t* cycles);

#ifdef __cplusplus
}
#endif
#endif

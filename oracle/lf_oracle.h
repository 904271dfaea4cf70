/*
 * lf_oracle — CPU restatement of the WolfNum/LimbForge arithmetic contract.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing under limbforge_amd/ links or calls it.
 *
 * It restates, in plain C with a runtime word count n (2..32), the algorithms
 * of the reference's shared arithmetic core and its exact host dot product:
 *   include/limbforge/core.hpp    (pack, add/sub, mul, square, divide_word,
 *                                   div, sqrt, fma/fms, cadd/cmul/cdiv)
 *   include/limbforge/linalg.hpp:222-321 (pack_words, exact_dot_add)
 *   src/kernels.metal:251-266     (precision casts)
 *   src/linalg.mm:411-486         (blocked Cholesky / triangular solves, host order)
 * Every function cites the reference line range it follows.
 *
 * Pinning (see tests/test_oracle_*.py): every operation is compared bit for
 * bit with MPFR (correctly rounded by specification) on seeded random inputs
 * and on the reference tests' fixture classes; divide_word against 64-bit
 * integer division; exact_dot_add against one mpfr_sum of exact products; and
 * everything against oracle/_ref (the reference's own core.hpp compiled from
 * /root/reference) where that library was built.
 *
 * Record layout (the public AoS element, engine.hpp:195): n little-endian
 * uint32 limbs, int32 exponent, int32 sign, uint32 status = 4n+12 bytes.
 */
#ifndef LF_ORACLE_H
#define LF_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define LFO_MAX_WORDS 32
#define LFO_EMAX 1000000000

typedef struct {
    uint32_t limb[LFO_MAX_WORDS];
    int32_t exponent;
    int32_t sign;
    uint32_t status;
} lfo_num;

enum { LFO_OK = 0, LFO_DIVISION_BY_ZERO = 1, LFO_EXPONENT_OVERFLOW = 2, LFO_INVALID = 4 };

/* Operation numbering of limbforge::Operation (engine.hpp:13). */
enum {
    LFO_ADD = 0, LFO_SUB, LFO_MUL, LFO_DIV, LFO_CADD, LFO_CMUL, LFO_CDIV,
    LFO_SQUARE, LFO_SQRT, LFO_FMA, LFO_FMS, LFO_CFMA, LFO_CFMS
};

/* record <-> working form */
void lfo_load(int n, const void* rec, lfo_num* x);
void lfo_store(int n, const lfo_num* x, void* rec);

/* scalar primitives on working numbers (n words) */
lfo_num lfo_zero(uint32_t status);
lfo_num lfo_pack(int n, const uint32_t* w, int words, int64_t scale, int sign, uint32_t status);
lfo_num lfo_add(int n, lfo_num a, lfo_num b);
lfo_num lfo_sub(int n, lfo_num a, lfo_num b);
lfo_num lfo_mul(int n, lfo_num a, lfo_num b);
lfo_num lfo_square(int n, lfo_num a);
lfo_num lfo_div(int n, lfo_num a, lfo_num b);
lfo_num lfo_sqrt(int n, lfo_num a);
lfo_num lfo_fma(int n, lfo_num a, lfo_num b, lfo_num c);
lfo_num lfo_fms(int n, lfo_num a, lfo_num b, lfo_num c);
void lfo_cadd(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2]);
void lfo_cmul(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2]);
void lfo_cdiv(int n, const lfo_num a[2], const lfo_num b[2], lfo_num r[2]);
/* RN(a*b + c*d + e), one rounding (core.hpp:330-352) */
lfo_num lfo_dot2_add(int n, lfo_num a, lfo_num b, lfo_num c, lfo_num d, lfo_num e);
/* complex fused a*b + c / a*b - c, each part one rounding (core.hpp:353-358) */
void lfo_cfma(int n, const lfo_num a[2], const lfo_num b[2], const lfo_num c[2], lfo_num r[2]);
void lfo_cfms(int n, const lfo_num a[2], const lfo_num b[2], const lfo_num c[2], lfo_num r[2]);
lfo_num lfo_cast_words(int to_n, int from_n, const lfo_num* x);

/* divide_word (core.hpp:176-184): returns quotient | remainder << 32 */
uint64_t lfo_divide_word(uint32_t high, uint32_t low, uint32_t d, uint32_t inv);

/* ---- batch entry points over public AoS records ---- */
/* out[i] = op(a[i], b[(i/b_stride)%b_period], c[(i/c_stride)%c_period]); period 0 = no wrap.
 * Returns 0, or -1 for an unsupported op / width. */
int lfo_elementwise(int bits, int op, const void* a, const void* b, const void* c, void* out,
                    size_t count, size_t b_stride, size_t b_period, size_t c_stride, size_t c_period);
/* x[i] <- RN(x[i] * f[i]) `steps` times (the resident chain of benchmarks/benchmark.cpp:67-94). */
int lfo_mul_chain(int bits, const void* x0, const void* f, void* out, size_t count, int steps);
/* Precision cast, componentwise for complex (kernels.metal:251-266). */
int lfo_cast(int from_bits, int to_bits, int complex_values, const void* in, void* out, size_t count);

/* exact_dot_add (linalg.hpp:274-318): out = RN(c -/+ sum a[k*sa]*b[k*sb]); c may be NULL. */
void lfo_exact_dot_add(int bits, const void* c, int subtract, const void* a, ptrdiff_t sa,
                       const void* b, ptrdiff_t sb, size_t K, void* out);
/* Dense products (contract of linalg.hpp:149-155) computed output by output with exact_dot_add. */
int lfo_gemm(int bits, int transpose_a, const void* A, const void* B, size_t m, size_t n, size_t k,
             void* C, int subtract);
int lfo_syrk(int bits, const void* A, size_t rows, size_t cols, void* C, int lower_only, int subtract);
/* Blocked Cholesky / solves in the reference's fixed host order (linalg.mm:411-486, gpu=false).
 * Returns the failing pivot (n on success). passes: 1 forward, 2 backward (L^T), 3 both. */
/* Mixed chains of one width: value i runs len[i] steps, step t being x <- RN(x*f + g) when bit t of
 * mask[i] is set and x <- RN(x*f) otherwise (Operation::fma / mul, one rounding each). */
int lfo_mixed_chain(int bits, const void* x, const void* f, const void* g, const uint32_t* len, const uint64_t* mask,
                    void* out, size_t count);
/* out[i] = sequential fma (cfma) chain over a[i*length + k] * b[i*length + k] (b[k] with shared_right) */
int lfo_segmented_dot(int bits, int complex_values, size_t segments, size_t length, int shared_right, const void* a,
                      const void* b, void* out);
/* out (one value) = the adjacent-pair tree of rounded additions of in[0..count) */
int lfo_tree_sum(int bits, int complex_values, const void* in, size_t count, void* out);
/* Host threads of lfo_cholesky / lfo_solve (default 1); the results do not depend on it. */
void lfo_set_threads(int threads);
size_t lfo_cholesky(int bits, const void* A, size_t n, void* L, size_t block, int* pivot_sign, uint32_t* pivot_status);
int lfo_solve(int bits, const void* L, size_t n, const void* B, size_t nrhs, void* X, size_t block, int passes);

#ifdef __cplusplus
}
#endif
#endif

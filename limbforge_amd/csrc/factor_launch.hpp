// factor_launch.hpp — launch entry points of the Cholesky / triangular-solve kernels (instantiated per
// width in factor_w.hip) and the argument records they share with the host driver (factor.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lf {

struct FactorInfo {
    int32_t pivot;   // first failing column, -1 while none
    int32_t sign;    // sign of the failing s_pp (0 for zero or a status)
    uint32_t status; // its status
    int32_t error;   // cluster scratch overflow (never expected)
};

struct FactorArgs {
    uint32_t* w;   // work matrix planes, element (i, j) at i*n + j
    uint64_t pitch;
    uint32_t n, k0, k1;
    FactorInfo* info;
    uint32_t* scratch; // per thread: 2 * swords words
    uint64_t swords;
    uint32_t* recip;   // reciprocals of the pivots l_jj (N+1 planes, entry j), written by the diagonal kernel
    uint64_t rpitch;
};

struct SolveArgs {
    const uint32_t* f; // factor planes (n x n)
    uint64_t fpitch;
    uint32_t* x;       // right-hand sides / solution planes (n x nrhs)
    uint64_t xpitch;
    uint32_t n, nrhs, r0, r1;
    uint32_t* scratch;
    uint64_t swords;
    int32_t* error;
    const uint32_t* recip; // reciprocals of the diagonal l_ii (N+1 planes, entry i), from launch_reciprocals
    uint64_t rpitch;
};

// Window words of the register path of the exact dot: 2N product words plus 24 words (768 bits) of
// exponent spread; wider spreads take the cluster path.
constexpr int factor_window(int N) { return 2 * N + 24; }

// Scratch words per accumulator of the cluster path for dots of up to K terms (plus the addend): a
// cluster spans at most 64N bits plus (K+1) * (64N + gap) bits, gap = 32N + ceil(log2(K+1)) + 4.
inline uint64_t factor_scratch_words(int N, uint64_t K)
{
    int logk = 0;
    while ((uint64_t(1) << logk) < K + 1) ++logk;
    const uint64_t bits = 64ull * N + (K + 1) * (64ull * N + 32ull * N + logk + 4);
    return bits / 32 + 4;
}

template <int N> hipError_t launch_chol_diag(const FactorArgs& f, uint32_t threads, hipStream_t s);
// one wavefront timing the per-row steps (factor_steps_kernel)
template <int N>
hipError_t launch_factor_steps(const uint32_t* v, uint64_t pitch, uint32_t* scratch, uint64_t swords, int32_t* err, int reps, uint64_t* out,
                               uint32_t* sink, hipStream_t s);
template <int N> hipError_t launch_chol_panel(const FactorArgs& f, hipStream_t s);
// entries (i, j), j <= i, of columns [k1, k2) and rows [k1, n) updated by the block [k0, k1); at most
// `waves` wavefronts (scratch slots)
template <int N> hipError_t launch_chol_lookahead(const FactorArgs& f, uint32_t k2, uint32_t waves, hipStream_t s);
template <int N> hipError_t launch_chol_invalid(uint32_t* w, uint64_t pitch, uint32_t n, uint32_t p, hipStream_t s);
template <int N> hipError_t launch_zero_upper(uint32_t* w, uint64_t pitch, uint32_t n, hipStream_t s);
template <int N> hipError_t launch_solve(int backward, const SolveArgs& a, hipStream_t s);
// the off-block update of block [r0, r1) by exact dots (rows [r1, n) forward, [0, r0) backward); at most
// `waves` wavefronts (scratch slots)
template <int N> hipError_t launch_solve_update(int backward, const SolveArgs& a, uint32_t waves, hipStream_t s);
// recip entry i = reciprocal of the mantissa of the diagonal element (i, i) of the n x n planes, i < n
template <int N> hipError_t launch_reciprocals(const uint32_t* w, uint64_t pitch, uint32_t n, uint32_t* recip, uint64_t rpitch, hipStream_t s);

} // namespace lf

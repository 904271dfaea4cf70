#pragma once
// Exact dense products of the LimbForge/WolfNum API on AMD Instinct MI355X: C[i][j] = RN(sum_k L[i][k] R[k][j])
// or RN(C[i][j] - sum ...) as an update, one rounding per output (ties to even), for any K, exponent spread
// or cancellation.  The public surface follows the reference's include/limbforge/linalg.hpp for the products
// (LinalgOptions, LinalgReport, Linalg::gemm / syrk and their *_exact names, report(), options(),
// exact_dot_add / exact_dot).  Behind it are the C ABI entries lf_gemm_host / lf_syrk_host of
// liblimbforge_hip.so: exponent-band residue GEMMs on int8 MFMA with CRT reconstruction, and exact-dot
// kernels for lines outside the bands (include/limbforge_hip.h).  Statuses propagate: an output whose row of
// L or column of R carries a status is zero with the OR of those statuses (and the old entry's for an update).
// Empty K gives canonical zero (an update leaves C unchanged).
#include "core.hpp"
#include "engine.hpp"

#include <cstddef>
#include <cstdint>
#include <future>
#include <memory>
#include <vector>

namespace limbforge {
struct LinalgOptions {
    int band_bits = 64;        // G: a band holds exponents [top-G, top] of one row/column
    int max_bands = 4;         // rows/columns needing more bands use the exact dot kernels
    int max_spread = 512;      // ... as do rows/columns whose nonzero exponents span more than this
    unsigned host_threads = 0; // band-analysis threads; 0 = hardware concurrency
    bool force_fallback = false; // every output through the exact dot kernels (testing)
    int resident_bands = 2;
};
struct LinalgPrewarm {
    std::vector<int> bits;
    std::vector<std::size_t> inner_sizes;
    bool resident = false;
};
struct LinalgPair {
    unsigned left_band, right_band, rows, cols, moduli;
};
// Diagnostics of the last call. "Lines" are rows of the left operand and columns of the right operand.
struct LinalgReport {
    std::size_t m = 0, n = 0, k = 0;
    std::vector<std::size_t> left_bands, right_bands; // [b] = lines with exactly b bands (index 0: zero or status lines)
    std::size_t left_fallback = 0, right_fallback = 0;
    std::size_t gpu_outputs = 0, multi_band_outputs = 0, fallback_outputs = 0, trivial_outputs = 0;
    std::vector<LinalgPair> pairs;
    std::size_t gemm_rows = 0, gemm_cols = 0, int8_gemms = 0;
    bool zero_copy_output = false;
    bool resident_overflow = false;
    double analysis_seconds = 0, upload_seconds = 0, gpu_seconds = 0, assembly_seconds = 0, fallback_seconds = 0;
};

// Blocked Cholesky and triangular solves (the reference's linalg.hpp:70-87).  The result depends on
// `block` (the documented rounding sequence) and on nothing else; host_macs, solve_host_macs and gpu
// choose the reference's host/GPU placement of updates, which never changes a bit; here every step runs
// on the GPU, and solve_host_macs picks direct exact dots (below it or up to 2^22 terms) or the update
// GEMM (above; every off-block solve update when negative) for the solves' off-block updates.
struct FactorOptions {
    std::size_t block = 32;       // column block nb; 0 selects a single block
    double host_macs = 5e4;
    double solve_host_macs = 4e5;
    bool gpu = true;
};
// pivot = n on success; otherwise the first column p whose pivot s_pp = RN(a_pp - sum_k l_pk^2) is zero,
// negative or carries a status (pivot_sign / pivot_status describe it).  On failure L[i][j] is final for
// j < p and every lower entry with j >= p is zero with status invalid.
struct CholeskyInfo {
    std::size_t pivot = 0;
    int pivot_sign = 0;
    word pivot_status = 0;
    Timing timing{0, 0};
    std::size_t blocks = 0, gpu_updates = 0, host_updates = 0;
    double panel_seconds = 0, update_seconds = 0;
};

namespace detail {
struct LinalgPass;
template <class T> constexpr int real_bits()
{
    static_assert(Format<T>::bits && !Format<T>::complex, "dense linear algebra uses real Float<bits> buffers");
    return Format<T>::bits;
}
} // namespace detail

// Handle of a resident product (Linalg::syrk / gemm on a CommandBatch).  C is final once the submission is
// waited; report() then gives the analysis and placement exactly as the host-array call reports them (no
// timings).  Copies share state.
class LinalgTicket {
    std::shared_ptr<detail::LinalgPass> pass_;
    explicit LinalgTicket(std::shared_ptr<detail::LinalgPass> p) : pass_(std::move(p)) {}
    friend class Linalg;

public:
    LinalgTicket() = default;
    bool resolved() const;              // the submission was waited: C is final
    const LinalgReport& report() const; // throws std::logic_error before resolved()
    bool provisional_reads() const;     // a later operation of the batch read C; throws before resolved()
};

// One Linalg per host thread.  C must not overlap A or B.  Calls return after the GPU has finished.
class Linalg {
public:
    explicit Linalg(LinalgOptions options = {});
    explicit Linalg(Engine& engine, LinalgOptions options = {});
    ~Linalg();
    Linalg(const Linalg&) = delete;
    Linalg& operator=(const Linalg&) = delete;
    void prewarm(const LinalgPrewarm&);
    std::future<void> prewarm_async(LinalgPrewarm);
    std::string device_name() const;
    // C (cols x cols) = A^T A for A (rows x cols).  lower_only writes C[i][j] for i >= j and leaves the rest
    // untouched; otherwise the full symmetric matrix.  With subtract, C[i][j] = RN(C[i][j] - (A^T A)[i][j])
    // (lower triangle only: subtract requires lower_only).
    Timing syrk(int bits, const void* A, std::size_t rows, std::size_t cols, void* C, bool lower_only = true, bool subtract = false);
    // C (m x n) = op(A) B with B (k x n); op(A) = A (m x k) or, with transpose_a, A^T for A (k x m).
    // With subtract, C = RN(C - op(A) B), one rounding per entry.
    Timing gemm(int bits, bool transpose_a, const void* A, const void* B, std::size_t m, std::size_t n, std::size_t k, void* C,
                bool subtract = false);
    // Lower L (n x n, row-major) with A = L L^T from the lower triangle of A; the strict upper triangle of
    // L is set to zero.  L may equal A.
    CholeskyInfo cholesky(int bits, const void* A, std::size_t n, void* L, const FactorOptions& options = {});
    // X (n x nrhs) solves L X = B, or L^T X = B with transpose (strict upper triangle of L not read).
    // X may equal B.
    Timing trsm(int bits, bool transpose, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X,
                const FactorOptions& options = {});
    // X solves (L L^T) X = B: trsm, then trsm with transpose.
    Timing cholesky_solve(int bits, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X,
                          const FactorOptions& options = {});
    CholeskyInfo cholesky_blocked(int bits, const void* A, std::size_t n, void* L, const FactorOptions& o = {})
    {
        return cholesky(bits, A, n, L, o);
    }
    Timing cholesky_solve_blocked(int bits, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X,
                                  const FactorOptions& o = {})
    {
        return cholesky_solve(bits, L, n, B, nrhs, X, o);
    }
    const LinalgReport& report() const;
    const LinalgOptions& options() const;

    // ---- Resident products: syrk / gemm encoded into an Engine's CommandBatch on Buffer<Float<bits>> operands
    // of that engine (row-major, sizes as the host calls; buffers may be larger).  After Submission::wait() C
    // equals the host-array call bit for bit.  C is provisional within the batch: a later write to it in the
    // same batch throws std::logic_error and a later read sets the ticket's provisional_reads().  C must not be A
    // or B.  The Linalg's own report() is not changed.
    template <class T>
    LinalgTicket syrk(CommandBatch& batch, const Buffer<T>& A, std::size_t rows, std::size_t cols, Buffer<T>& C, bool lower_only = true,
                      bool subtract = false)
    {
        return encode_product(batch, detail::real_bits<T>(), true, false, detail::Access::operand(A), detail::Access::operand(A), cols, cols, rows,
                              detail::Access::operand(C), lower_only, subtract);
    }
    template <class T>
    LinalgTicket gemm(CommandBatch& batch, bool transpose_a, const Buffer<T>& A, const Buffer<T>& B, std::size_t m, std::size_t n, std::size_t k,
                      Buffer<T>& C, bool subtract = false)
    {
        return encode_product(batch, detail::real_bits<T>(), false, transpose_a, detail::Access::operand(A), detail::Access::operand(B), m, n, k,
                              detail::Access::operand(C), false, subtract);
    }
    template <class T>
    LinalgTicket gemm_exact(CommandBatch& b, bool transpose_a, const Buffer<T>& A, const Buffer<T>& B, std::size_t m, std::size_t n, std::size_t k,
                            Buffer<T>& C, bool subtract = false)
    {
        return gemm(b, transpose_a, A, B, m, n, k, C, subtract);
    }
    template <class T>
    LinalgTicket syrk_exact(CommandBatch& b, const Buffer<T>& A, std::size_t rows, std::size_t cols, Buffer<T>& C, bool lower_only = true,
                            bool subtract = false)
    {
        return syrk(b, A, rows, cols, C, lower_only, subtract);
    }
    // Factorizations and solves on Buffer operands: synchronous, outside any batch.  The buffers must be idle;
    // the GPU uses them in place (no copies) and the results equal the host-array calls bit for bit.
    template <class T> CholeskyInfo cholesky(const Buffer<T>& A, std::size_t n, Buffer<T>& L, const FactorOptions& options = {})
    {
        return cholesky_buffers(detail::real_bits<T>(), detail::Access::operand(A), n, detail::Access::operand(L), options);
    }
    template <class T>
    Timing trsm(bool transpose, const Buffer<T>& L, std::size_t n, const Buffer<T>& B, std::size_t nrhs, Buffer<T>& X,
                const FactorOptions& options = {})
    {
        return solve_buffers(detail::real_bits<T>(), transpose ? 2 : 1, detail::Access::operand(L), n, detail::Access::operand(B), nrhs,
                             detail::Access::operand(X), options);
    }
    template <class T>
    Timing cholesky_solve(const Buffer<T>& L, std::size_t n, const Buffer<T>& B, std::size_t nrhs, Buffer<T>& X, const FactorOptions& options = {})
    {
        return solve_buffers(detail::real_bits<T>(), 3, detail::Access::operand(L), n, detail::Access::operand(B), nrhs, detail::Access::operand(X),
                             options);
    }
    Timing gemm_exact(int bits, bool transpose_a, const void* A, const void* B, std::size_t m, std::size_t n, std::size_t k, void* C,
                      bool subtract = false)
    {
        return gemm(bits, transpose_a, A, B, m, n, k, C, subtract);
    }
    Timing syrk_exact(int bits, const void* A, std::size_t rows, std::size_t cols, void* C, bool lower_only = true, bool subtract = false)
    {
        return syrk(bits, A, rows, cols, C, lower_only, subtract);
    }

private:
    LinalgTicket encode_product(CommandBatch& batch, int bits, bool syrk, bool transpose_a, const detail::Operand& A, const detail::Operand& B,
                                std::size_t m, std::size_t n, std::size_t k, const detail::Operand& C, bool lower_only, bool subtract);
    CholeskyInfo cholesky_buffers(int bits, const detail::Operand& A, std::size_t n, const detail::Operand& L, const FactorOptions& options);
    Timing solve_buffers(int bits, int passes, const detail::Operand& L, std::size_t n, const detail::Operand& B, std::size_t nrhs,
                         const detail::Operand& X, const FactorOptions& options);
    struct Impl;
    std::unique_ptr<Impl> impl;
};

namespace detail {
// One exact dot product on the host: RN(c -/+ sum a[k*sa] b[k*sb]) (c may be null), records of bits/32 limbs.
void exact_dot_host(int bits, const void* c, bool subtract, const void* a, std::ptrdiff_t sa, const void* b, std::ptrdiff_t sb,
                    std::size_t K, void* out);
} // namespace detail

// RN(c - sum_k a[k*sa]*b[k*sb]) (subtract) or RN(c + sum ...) with one rounding; c == nullptr is a zero
// addend (the reference's linalg.hpp:268-318 contract), computed on the host.  Any status among c, a[k], b[k]
// gives zero with the OR of those statuses; without a nonzero product the result is c itself; exact
// cancellation gives canonical zero.
template <int N>
Number<N> exact_dot_add(const Number<N>* c, bool subtract, const Number<N>* a, std::ptrdiff_t sa, const Number<N>* b, std::ptrdiff_t sb,
                        std::size_t K)
{
    Number<N> r;
    detail::exact_dot_host(32 * N, c, subtract, a, sa, b, sb, K, &r);
    return r;
}
// RN(sum_k a[k*sa]*b[k*sb]) with one rounding; empty K and exact cancellation give canonical zero.
template <int N> Number<N> exact_dot(const Number<N>* a, std::ptrdiff_t sa, const Number<N>* b, std::ptrdiff_t sb, std::size_t K)
{
    return exact_dot_add<N>(nullptr, false, a, sa, b, sb, K);
}
} // namespace limbforge

// linalg.cpp — the C++17 Linalg facade over the C ABI's products and factorizations (include/limbforge_hip.h):
// host-array calls (lf_gemm_host / lf_syrk_host / lf_cholesky_host / lf_solve_host), resident products encoded
// into a CommandBatch (lf_gemm / lf_syrk on the batch's stream, tickets resolved at wait) and the Buffer forms
// of the factorization and solves (lf_cholesky / lf_solve in place).  Option validation, report translation
// and the error classes follow the reference's src/linalg.mm:255-307.
#include "limbforge/linalg.hpp"
#include "engine_internal.hpp"

#include <atomic>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace limbforge {

struct Linalg::Impl {
    LinalgOptions options;
    LinalgReport report;
    lf_linalg_options c_options() const
    {
        lf_linalg_options o;
        lf_linalg_default_options(&o);
        o.band_bits = options.band_bits;
        o.max_bands = options.max_bands;
        o.max_spread = options.max_spread;
        o.host_threads = options.host_threads;
        o.force_fallback = options.force_fallback;
        o.resident_bands = options.resident_bands;
        return o;
    }
    void take(const lf_linalg_report& r) { report = translate(r); }
    static LinalgReport translate(const lf_linalg_report& r)
    {
        LinalgReport out;
        out.m = r.m;
        out.n = r.n;
        out.k = r.k;
        out.left_bands.assign(r.left_bands, r.left_bands + r.band_slots);
        out.right_bands.assign(r.right_bands, r.right_bands + r.band_slots);
        out.left_fallback = r.left_fallback;
        out.right_fallback = r.right_fallback;
        out.gpu_outputs = r.gpu_outputs;
        out.multi_band_outputs = r.multi_band_outputs;
        out.fallback_outputs = r.fallback_outputs;
        out.trivial_outputs = r.trivial_outputs;
        for (uint32_t i = 0; i < r.npairs; ++i)
            out.pairs.push_back({r.pairs[i].left_band, r.pairs[i].right_band, r.pairs[i].rows, r.pairs[i].cols, r.pairs[i].moduli});
        out.gemm_rows = r.gemm_rows;
        out.gemm_cols = r.gemm_cols;
        out.int8_gemms = r.int8_gemms;
        out.analysis_seconds = r.analysis_seconds;
        out.upload_seconds = r.upload_seconds;
        out.gpu_seconds = r.gpu_seconds;
        out.assembly_seconds = r.assembly_seconds;
        out.fallback_seconds = r.fallback_seconds;
        return out;
    }
};

namespace detail {
struct LinalgPass {
    LinalgReport report;
    std::atomic<bool> resolved{false};
    std::shared_ptr<std::atomic<bool>> reads = std::make_shared<std::atomic<bool>>(false);
};
} // namespace detail

bool LinalgTicket::resolved() const { return pass_ && pass_->resolved.load(); }
const LinalgReport& LinalgTicket::report() const
{
    if (!resolved()) throw std::logic_error("LinalgTicket: the product's submission has not been waited");
    return pass_->report;
}
bool LinalgTicket::provisional_reads() const
{
    if (!resolved()) throw std::logic_error("LinalgTicket: the product's submission has not been waited");
    return pass_->reads->load();
}

static void check_bits(int bits)
{
    if (bits < 64 || bits > 1024 || bits % 32) throw std::invalid_argument("bits must be a multiple of 32 in [64,1024]");
}

Linalg::Linalg(LinalgOptions options) : impl(std::make_unique<Impl>())
{
    if (options.band_bits < 0 || options.band_bits > 1024 || options.max_bands < 0 || options.max_spread < 0 || options.max_spread > 8192)
        throw std::invalid_argument("linalg options: band_bits in [0,1024], max_bands >= 0, max_spread in [0,8192]");
    int devices = 0;
    lf_device_count(&devices);
    if (!devices) throw std::runtime_error("no AMD GPU available");
    impl->options = options;
}

Linalg::Linalg(Engine&, LinalgOptions options) : Linalg(options) {}
Linalg::~Linalg() = default;

void Linalg::prewarm(const LinalgPrewarm& r)
{
    for (int bits : r.bits) check_bits(bits);
    for (auto k : r.inner_sizes)
        if (k > 65472) throw std::invalid_argument("linalg prewarm: resident K exceeds 65472");
}

std::future<void> Linalg::prewarm_async(LinalgPrewarm r)
{
    prewarm(r);
    std::promise<void> done;
    done.set_value();
    return done.get_future();
}

std::string Linalg::device_name() const
{
    char buf[256] = {};
    detail::check(lf_device_name(buf, sizeof buf), "device_name");
    return buf;
}

const LinalgReport& Linalg::report() const { return impl->report; }
const LinalgOptions& Linalg::options() const { return impl->options; }

Timing Linalg::syrk(int bits, const void* A, std::size_t rows, std::size_t cols, void* C, bool lower_only, bool subtract)
{
    check_bits(bits);
    if (subtract && !lower_only) throw std::invalid_argument("syrk: subtract updates the lower triangle only (lower_only)");
    lf_linalg_options o = impl->c_options();
    lf_linalg_report r;
    Timing t{0, 0};
    detail::check(lf_syrk_host(bits, A, rows, cols, C, lower_only, subtract, &o, &r, &t.gpu_seconds, &t.wall_seconds), "Linalg::syrk");
    impl->take(r);
    return t;
}

Timing Linalg::gemm(int bits, bool transpose_a, const void* A, const void* B, std::size_t m, std::size_t n, std::size_t k, void* C,
                    bool subtract)
{
    check_bits(bits);
    lf_linalg_options o = impl->c_options();
    lf_linalg_report r;
    Timing t{0, 0};
    detail::check(lf_gemm_host(bits, transpose_a, A, B, m, n, k, C, subtract, &o, &r, &t.gpu_seconds, &t.wall_seconds), "Linalg::gemm");
    impl->take(r);
    return t;
}

static lf_factor_options c_factor(const FactorOptions& o)
{
    lf_factor_options f;
    lf_factor_default_options(&f);
    f.block = o.block;
    f.host_macs = o.host_macs;
    f.solve_host_macs = o.solve_host_macs;
    f.gpu = o.gpu;
    return f;
}

CholeskyInfo Linalg::cholesky(int bits, const void* A, std::size_t n, void* L, const FactorOptions& options)
{
    check_bits(bits);
    lf_linalg_options o = impl->c_options();
    lf_factor_options f = c_factor(options);
    lf_cholesky_info ci;
    detail::check(lf_cholesky_host(bits, A, n, L, &f, &o, &ci), "Linalg::cholesky");
    CholeskyInfo info;
    info.pivot = ci.pivot;
    info.pivot_sign = ci.pivot_sign;
    info.pivot_status = ci.pivot_status;
    info.timing = {ci.timing_gpu_seconds, ci.timing_wall_seconds};
    info.blocks = ci.blocks;
    info.gpu_updates = ci.gpu_updates;
    info.host_updates = ci.host_updates;
    info.panel_seconds = ci.panel_seconds;
    info.update_seconds = ci.update_seconds;
    return info;
}

static Timing solve(int bits, int passes, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X, const FactorOptions& options,
                    const lf_linalg_options& o, const char* where)
{
    if (bits < 64 || bits > 1024 || bits % 32) throw std::invalid_argument("bits must be a multiple of 32 in [64,1024]");
    lf_factor_options f = c_factor(options);
    Timing t{0, 0};
    detail::check(lf_solve_host(bits, passes, L, n, B, nrhs, X, &f, &o, &t.gpu_seconds, &t.wall_seconds), where);
    return t;
}

Timing Linalg::trsm(int bits, bool transpose, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X, const FactorOptions& options)
{
    return solve(bits, transpose ? 2 : 1, L, n, B, nrhs, X, options, impl->c_options(), "Linalg::trsm");
}

Timing Linalg::cholesky_solve(int bits, const void* L, std::size_t n, const void* B, std::size_t nrhs, void* X, const FactorOptions& options)
{
    return solve(bits, 3, L, n, B, nrhs, X, options, impl->c_options(), "Linalg::cholesky_solve");
}

// ---- resident products ----
static void check_operand_bits(const detail::Operand& o, int bits, const char* name)
{
    if (!o.storage) throw std::invalid_argument(std::string("null buffer ") + name);
    if (o.storage->bits != bits || o.storage->complex) throw std::invalid_argument(std::string("buffer ") + name + " has another format");
}

LinalgTicket Linalg::encode_product(CommandBatch& batch, int bits, bool syrk, bool transpose_a, const detail::Operand& A,
                                    const detail::Operand& B, std::size_t m, std::size_t n, std::size_t k, const detail::Operand& C,
                                    bool lower_only, bool subtract)
{
    check_bits(bits);
    if (syrk && subtract && !lower_only) throw std::invalid_argument("syrk: subtract updates the lower triangle only (lower_only)");
    check_operand_bits(A, bits, "A");
    check_operand_bits(B, bits, "B");
    check_operand_bits(C, bits, "C");
    if (C.storage == A.storage || C.storage == B.storage) throw std::invalid_argument("resident product: C must not be A or B");
    const bool fits = syrk ? (A.size >= k * m && C.size >= m * m) : (A.size >= m * k && B.size >= k * n && C.size >= m * n);
    if (!fits) throw std::invalid_argument("resident product: a buffer is smaller than the matrix it holds");
    detail::BatchAccess::retain(batch, A.storage, false);
    if (!syrk) detail::BatchAccess::retain(batch, B.storage, false);
    detail::BatchAccess::retain(batch, C.storage, true);
    auto pass = std::make_shared<detail::LinalgPass>();
    const lf_linalg_options o = impl->c_options();
    auto a = A.storage, b = B.storage, c = C.storage;
    detail::BatchAccess::enqueue(batch, [=](void* stream) {
        lf_linalg_report r;
        const lf_operand oa = a->operand(), ob = b->operand(), oc = c->operand();
        if (syrk) detail::check(lf_syrk(bits, &oa, k, m, &oc, lower_only, subtract, &o, &r, nullptr, stream), "Linalg::syrk (resident)");
        else
            detail::check(lf_gemm(bits, transpose_a, &oa, &ob, m, n, k, &oc, subtract, &o, &r, nullptr, stream), "Linalg::gemm (resident)");
        LinalgReport rep = Impl::translate(r);
        rep.analysis_seconds = rep.upload_seconds = rep.gpu_seconds = rep.assembly_seconds = rep.fallback_seconds = 0;
        pass->report = std::move(rep);
    });
    detail::BatchAccess::mark_provisional(batch, c, pass->reads);
    detail::BatchAccess::on_completion(batch, [pass] { pass->resolved.store(true); });
    return LinalgTicket(pass);
}

// ---- factorization and solves on Buffer operands ----
namespace {
// Claims the buffers for a synchronous call (as a submission does), writes back host-side changes first,
// and marks the written ones' host shadows stale afterwards.
struct BufferCall {
    std::vector<std::pair<std::shared_ptr<detail::BufferStorage>, bool>> held;
    void claim(const detail::Operand& o, int bits, bool written, const char* name)
    {
        check_operand_bits(o, bits, name);
        for (auto& h : held)
            if (h.first == o.storage) {
                h.second = h.second || written;
                return;
            }
        bool expected = false;
        if (!o.storage->busy.compare_exchange_strong(expected, true)) throw std::logic_error("buffer belongs to an unwaited submission");
        held.emplace_back(o.storage, written);
        o.storage->flush();
    }
    ~BufferCall()
    {
        for (auto& h : held) {
            if (h.second) {
                std::lock_guard<std::mutex> lock(h.first->mutex);
                h.first->host_valid = false;
            }
            h.first->busy.store(false);
        }
    }
};
} // namespace

CholeskyInfo Linalg::cholesky_buffers(int bits, const detail::Operand& A, std::size_t n, const detail::Operand& L, const FactorOptions& options)
{
    check_bits(bits);
    if (A.size < n * n || L.size < n * n) throw std::invalid_argument("cholesky: a buffer is smaller than n x n");
    BufferCall call;
    call.claim(A, bits, false, "A");
    call.claim(L, bits, true, "L");
    lf_linalg_options o = impl->c_options();
    lf_factor_options f = c_factor(options);
    lf_cholesky_info ci;
    const lf_operand oa = A.storage->operand(), ol = L.storage->operand();
    detail::check(lf_cholesky(bits, &oa, n, &ol, &f, &o, &ci, nullptr), "Linalg::cholesky");
    CholeskyInfo info;
    info.pivot = ci.pivot;
    info.pivot_sign = ci.pivot_sign;
    info.pivot_status = ci.pivot_status;
    info.timing = {ci.timing_gpu_seconds, ci.timing_wall_seconds};
    info.blocks = ci.blocks;
    info.gpu_updates = ci.gpu_updates;
    info.host_updates = ci.host_updates;
    info.panel_seconds = ci.panel_seconds;
    info.update_seconds = ci.update_seconds;
    return info;
}

Timing Linalg::solve_buffers(int bits, int passes, const detail::Operand& L, std::size_t n, const detail::Operand& B, std::size_t nrhs,
                             const detail::Operand& X, const FactorOptions& options)
{
    check_bits(bits);
    if (L.size < n * n || B.size < n * nrhs || X.size < n * nrhs) throw std::invalid_argument("solve: a buffer is smaller than its matrix");
    BufferCall call;
    call.claim(L, bits, false, "L");
    call.claim(B, bits, false, "B");
    call.claim(X, bits, true, "X");
    lf_linalg_options o = impl->c_options();
    lf_factor_options f = c_factor(options);
    Timing t{0, 0};
    const lf_operand ol = L.storage->operand(), ob = B.storage->operand(), ox = X.storage->operand();
    detail::check(lf_solve(bits, passes, &ol, n, &ob, nrhs, &ox, &f, &o, &t.gpu_seconds, &t.wall_seconds, nullptr),
                  passes == 3 ? "Linalg::cholesky_solve" : "Linalg::trsm");
    return t;
}

} // namespace limbforge

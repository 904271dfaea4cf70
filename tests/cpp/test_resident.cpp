// C++ API test of the resident products and the Buffer forms of the factorization (the reference's
// tests/test_resident_linalg.cpp:51-64 and tests/test_cholesky.cpp Buffer cases): Linalg::gemm / syrk encoded
// into a CommandBatch after a producer of A in the same batch equal the host-array calls bit for bit, report
// included; C is provisional within its batch (a later write throws, a later read is flagged); tickets resolve
// at wait; Linalg::cholesky / cholesky_solve / trsm on Buffers equal the host-array calls.  Built and run by
// tests/test_cpp_api.py on the GPU.
#include <limbforge/linalg.hpp>

#include <cstring>
#include <iostream>
#include <random>
#include <vector>

using namespace limbforge;
using F = Float<256>;

static int failures = 0;
#define REQUIRE(cond, msg)                                                                                             \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            ++failures;                                                                                                \
            std::cerr << "FAIL " << msg << " (" << __LINE__ << ")\n";                                                  \
        }                                                                                                              \
    } while (0)

static F random_value(std::mt19937_64& g, int lo, int hi)
{
    F x = zero<8>();
    for (auto& w : x.limb) w = uint32_t(g());
    x.limb[7] |= 0x80000000u;
    x.exponent = lo + int(g() % uint64_t(hi - lo + 1));
    x.sign = g() % 2 ? 1 : -1;
    return x;
}

static bool same(const std::vector<F>& a, const std::vector<F>& b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(F)) == 0;
}

static bool same_report(const LinalgReport& a, const LinalgReport& b)
{
    bool ok = a.m == b.m && a.n == b.n && a.k == b.k && a.left_bands == b.left_bands && a.right_bands == b.right_bands &&
              a.left_fallback == b.left_fallback && a.right_fallback == b.right_fallback && a.gpu_outputs == b.gpu_outputs &&
              a.multi_band_outputs == b.multi_band_outputs && a.fallback_outputs == b.fallback_outputs &&
              a.trivial_outputs == b.trivial_outputs && a.gemm_rows == b.gemm_rows && a.gemm_cols == b.gemm_cols &&
              a.int8_gemms == b.int8_gemms && a.pairs.size() == b.pairs.size();
    for (std::size_t i = 0; ok && i < a.pairs.size(); ++i)
        ok = a.pairs[i].left_band == b.pairs[i].left_band && a.pairs[i].right_band == b.pairs[i].right_band &&
             a.pairs[i].rows == b.pairs[i].rows && a.pairs[i].cols == b.pairs[i].cols && a.pairs[i].moduli == b.pairs[i].moduli;
    return ok;
}

int main()
{
    try {
        std::mt19937_64 g(42);
        Engine gpu;
        Linalg la(gpu);
        const std::size_t m = 37, n = 29, k = 70;
        // entries: three exponent clusters, a zero row and a line beyond max_spread (exact-dot fallback)
        std::vector<F> X(m * k), Y(m * k), B(k * n), C0(m * n);
        for (std::size_t i = 0; i < X.size(); ++i) {
            const int c = int(g() % 3) * 150 - 150;
            X[i] = random_value(g, c - 10, c + 10);
            Y[i] = random_value(g, -3, 3);
        }
        for (auto& v : B) {
            const int c = int(g() % 3) * 150 - 150;
            v = random_value(g, c - 10, c + 10);
        }
        for (auto& v : C0) v = random_value(g, -100, 100);
        for (std::size_t q = 0; q < k; ++q) X[5 * k + q] = zero<8>();
        B[3 * n + 2].exponent = 900; // column 2 spreads beyond max_spread
        // the host-array reference: A = X * Y element-wise, then C = C0 - A B
        std::vector<F> A(m * k), Ch = C0;
        gpu.run(256, Operation::mul, X.data(), Y.data(), A.data(), A.size());
        la.gemm(256, false, A.data(), B.data(), m, n, k, Ch.data(), true);
        const LinalgReport host = la.report();
        REQUIRE(host.fallback_outputs > 0 && host.gpu_outputs > 0, "the data exercises both the residue path and the fallback");
        // resident: the producer of A and the product in one batch
        auto bx = gpu.make_buffer<F>(X.size()), by = gpu.make_buffer<F>(Y.size()), ba = gpu.make_buffer<F>(A.size());
        auto bb = gpu.make_buffer<F>(B.size()), bc = gpu.make_buffer<F>(C0.size()), bd = gpu.make_buffer<F>(C0.size());
        bx.upload(X.data(), X.size());
        by.upload(Y.data(), Y.size());
        bb.upload(B.data(), B.size());
        bc.upload(C0.data(), C0.size());
        auto batch = gpu.batch();
        batch.run(Operation::mul, bx, by, ba);
        LinalgTicket t = la.gemm(batch, false, ba, bb, m, n, k, bc, true);
        batch.run(Operation::add, bc, bc, bd); // a reader of C: flagged
        bool threw = false;
        try {
            batch.run(Operation::add, bd, bd, bc); // a writer of C: rejected
        } catch (const std::logic_error&) {
            threw = true;
        }
        REQUIRE(threw, "a later write to a resident product's output is rejected");
        REQUIRE(!t.resolved(), "unresolved before submission");
        bool early = false;
        try {
            (void)t.report();
        } catch (const std::logic_error&) {
            early = true;
        }
        REQUIRE(early, "report() before wait throws");
        batch.submit().wait();
        REQUIRE(t.resolved(), "resolved after wait");
        std::vector<F> Cr(C0.size());
        bc.download(Cr.data(), Cr.size());
        REQUIRE(same(Cr, Ch), "resident gemm equals the host-array call");
        REQUIRE(same_report(t.report(), host), "ticket report equals the host-array report");
        REQUIRE(t.provisional_reads(), "the later reader was flagged");
        std::vector<F> Dw(C0.size()), Dr(C0.size());
        gpu.run(256, Operation::add, Ch.data(), Ch.data(), Dw.data(), Dw.size());
        bd.download(Dr.data(), Dr.size());
        REQUIRE(same(Dr, Dw), "the reader saw the product's final values");
        // resident SYRK (full and lower) against the host call
        const std::size_t rows = 90, cols = 24;
        std::vector<F> S(rows * cols), Sh(cols * cols, zero<8>()), Sl(cols * cols, zero<8>());
        for (auto& v : S) v = random_value(g, -20, 20);
        la.syrk(256, S.data(), rows, cols, Sh.data(), false);
        const LinalgReport hs = la.report();
        auto bs = gpu.make_buffer<F>(S.size()), bsc = gpu.make_buffer<F>(cols * cols);
        bs.upload(S.data(), S.size());
        bsc.upload(Sl.data(), Sl.size());
        auto b2 = gpu.batch();
        LinalgTicket ts = la.syrk(b2, bs, rows, cols, bsc, false);
        b2.submit().wait();
        std::vector<F> Sr(cols * cols);
        bsc.download(Sr.data(), Sr.size());
        REQUIRE(same(Sr, Sh), "resident syrk equals the host-array call");
        REQUIRE(same_report(ts.report(), hs), "syrk ticket report equals the host-array report");
        REQUIRE(!ts.provisional_reads(), "no later reader");
        // Buffer forms of cholesky / cholesky_solve / trsm against the host-array calls: S + I is SPD
        std::vector<F> one(cols);
        for (auto& v : one) {
            v = zero<8>();
            v.limb[7] = 0x80000000u;
            v.sign = 1;
        }
        std::vector<F> diag(cols), sum(cols);
        for (std::size_t i = 0; i < cols; ++i) diag[i] = Sh[i * cols + i];
        gpu.run(256, Operation::add, diag.data(), one.data(), sum.data(), cols);
        for (std::size_t i = 0; i < cols; ++i) Sh[i * cols + i] = sum[i];
        std::vector<F> Lh(cols * cols), R(cols * 5), Xh(cols * 5), Th(cols * 5);
        for (auto& v : R) v = random_value(g, -8, 8);
        const CholeskyInfo hi = la.cholesky(256, Sh.data(), cols, Lh.data());
        la.cholesky_solve(256, Lh.data(), cols, R.data(), 5, Xh.data());
        la.trsm(256, true, Lh.data(), cols, R.data(), 5, Th.data());
        auto bA = gpu.make_buffer<F>(cols * cols), bL = gpu.make_buffer<F>(cols * cols), bR = gpu.make_buffer<F>(cols * 5),
             bX = gpu.make_buffer<F>(cols * 5), bT = gpu.make_buffer<F>(cols * 5);
        bA.upload(Sh.data(), Sh.size());
        bR.upload(R.data(), R.size());
        const CholeskyInfo bi = la.cholesky(bA, cols, bL);
        la.cholesky_solve(bL, cols, bR, 5, bX);
        la.trsm(true, bL, cols, bR, 5, bT);
        std::vector<F> Lb(cols * cols), Xb(cols * 5), Tb(cols * 5);
        bL.download(Lb.data(), Lb.size());
        bX.download(Xb.data(), Xb.size());
        bT.download(Tb.data(), Tb.size());
        REQUIRE(bi.pivot == cols && hi.pivot == cols, "SPD");
        REQUIRE(same(Lb, Lh), "Buffer cholesky equals the host-array call");
        REQUIRE(same(Xb, Xh), "Buffer cholesky_solve equals the host-array call");
        REQUIRE(same(Tb, Th), "Buffer trsm (transpose) equals the host-array call");
        // in place: L may be A
        la.cholesky(bA, cols, bA);
        std::vector<F> La(cols * cols);
        bA.download(La.data(), La.size());
        REQUIRE(same(La, Lh), "Buffer cholesky in place");
    } catch (const std::exception& e) {
        std::cerr << "exception: " << e.what() << "\n";
        ++failures;
    }
    std::cout << "resident products: passed " << failures << " failures\n";
    return failures ? 1 : 0;
}

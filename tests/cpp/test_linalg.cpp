// C++ API parity test of the Linalg facade on the GPU, written as the reference's tests use the API
// (tests/test_linalg.cpp, tests/test_contract_names.cpp): every golden product section through
// Linalg::gemm / syrk (and their *_exact names) under the reference's option configurations, the
// [2^100, 1, -2^100] . [1, 1, 1] = 1 known answer through exact_dot / exact_dot_add, the report, and
// argument errors.  Built and run by tests/test_cpp_api.py.
#include <limbforge/linalg.hpp>

#include "lf_oracle.h"

#include <cstring>
#include <fstream>
#include <functional>
#include <random>
#include <iostream>
#include <string>
#include <vector>

using namespace limbforge;

// a random normalized value with exponent in [-span, span]
template <int Bits> Float<Bits> random_value(std::mt19937_64& g, int span)
{
    Float<Bits> x = zero<Bits / 32>();
    for (auto& w : x.limb) w = uint32_t(g());
    x.limb[Bits / 32 - 1] |= 0x80000000u;
    x.exponent = int(g() % uint64_t(2 * span + 1)) - span;
    x.sign = g() % 2 ? 1 : -1;
    return x;
}

static int failures = 0;
#define REQUIRE(cond, msg)                                                                                             \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            ++failures;                                                                                                \
            std::cerr << "FAIL " << msg << " (" << __LINE__ << ")\n";                                                  \
        }                                                                                                              \
    } while (0)

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    try {
        std::ifstream f(dir + "/products.bin", std::ios::binary);
        std::vector<unsigned char> bytes((std::istreambuf_iterator<char>(f)), {});
        uint32_t nsec;
        std::memcpy(&nsec, &bytes[12], 4);
        std::vector<LinalgOptions> configs(4);
        configs[1].band_bits = 8;       // narrow bands
        configs[2].band_bits = 0;       // zero-width bands
        configs[3].force_fallback = true;
        size_t off = 16;
        for (uint32_t s = 0; s < nsec; ++s) {
            uint32_t h[5];
            std::memcpy(h, &bytes[off], 20);
            off += 20;
            const uint32_t bits = h[0], m = h[1], n = h[2], k = h[3], flags = h[4];
            const size_t E = bits / 8 + 12;
            const bool ta = flags & 1, sub = flags & 2, syrk = flags & 4, lower = flags & 8;
            const unsigned char* A = &bytes[off];
            off += (syrk ? k * m : m * k) * E;
            const unsigned char* B = nullptr;
            if (!syrk) {
                B = &bytes[off];
                off += k * n * E;
            }
            const unsigned char* Cin = &bytes[off];
            off += m * n * E;
            const unsigned char* Cout = &bytes[off];
            off += m * n * E;
            for (size_t c = 0; c < configs.size(); ++c) {
                Linalg la(configs[c]);
                std::vector<unsigned char> C(Cin, Cin + m * n * E);
                if (syrk) la.syrk_exact(int(bits), A, k, m, C.data(), lower, sub);
                else la.gemm_exact(int(bits), ta, A, B, m, n, k, C.data(), sub);
                REQUIRE(std::memcmp(C.data(), Cout, C.size()) == 0, "golden product " << s << " config " << c);
                REQUIRE(la.report().m == m && la.report().k == k, "report shape " << s);
            }
        }
        // contract names known answer (test_contract_names.cpp:14-19)
        using F = Float<64>;
        F one = zero<2>();
        one.limb[1] = 0x80000000u;
        one.sign = 1;
        F large = one;
        large.exponent = 100;
        std::vector<F> a{large, one, negate(large)}, b(3, one);
        F exact = exact_dot<2>(a.data(), 1, b.data(), 1, 3);
        REQUIRE(std::memcmp(&exact, &one, sizeof(F)) == 0, "exact_dot [2^100,1,-2^100].[1,1,1] = 1");
        F c = one;
        c.exponent = 3; // 8
        F up = exact_dot_add<2>(&c, false, a.data(), 1, b.data(), 1, 3), down = exact_dot_add<2>(&c, true, a.data(), 1, b.data(), 1, 3);
        F nine, seven, eight = c;
        lfo_num o_up, o_down, l8, l1;
        lfo_load(2, &eight, &l8);
        lfo_load(2, &one, &l1);
        o_up = lfo_add(2, l8, l1);
        o_down = lfo_sub(2, l8, l1);
        lfo_store(2, &o_up, &nine);
        lfo_store(2, &o_down, &seven);
        REQUIRE(std::memcmp(&up, &nine, sizeof(F)) == 0, "exact_dot_add 8 + 1");
        REQUIRE(std::memcmp(&down, &seven, sizeof(F)) == 0, "exact_dot_add 8 - 1");
        // report of a multi-band product
        Linalg la;
        std::vector<F> C(1);
        la.gemm(64, false, a.data(), b.data(), 1, 1, 3, C.data());
        REQUIRE(std::memcmp(&C[0], &one, sizeof(F)) == 0, "Linalg::gemm known answer");
        REQUIRE(la.report().left_bands.size() == 5, "histogram has max_bands + 1 slots");
        // Cholesky, trsm and cholesky_solve against the oracle's host order (n = 24, block 8, 256 bits)
        {
            using G = Float<256>;
            const std::size_t n = 24, m = 2 * n + 1, nrhs = 3;
            std::mt19937_64 g(77);
            std::vector<G> J(m * n), S(n * n), L(n * n), W(n * n), B(n * nrhs), X(n * nrhs), Y(n * nrhs);
            for (auto& x : J) x = random_value<256>(g, 4);
            for (auto& x : B) x = random_value<256>(g, 8);
            lfo_syrk(256, J.data(), m, n, S.data(), 0, 0);
            for (std::size_t i = 0; i < n; ++i) S[i * n + i].exponent += 1; // 2 (J^T J)_ii: diagonally heavier
            FactorOptions fo;
            fo.block = 8;
            CholeskyInfo info = la.cholesky(256, S.data(), n, L.data(), fo);
            int ps = 0;
            uint32_t pst = 0;
            const std::size_t piv = lfo_cholesky(256, S.data(), n, W.data(), 8, &ps, &pst);
            REQUIRE(info.pivot == piv && piv == n && info.blocks == 3, "cholesky pivot " << info.pivot);
            REQUIRE(std::memcmp(L.data(), W.data(), L.size() * sizeof(G)) == 0, "cholesky factor");
            for (int passes = 1; passes <= 3; ++passes) {
                if (passes == 3) la.cholesky_solve(256, L.data(), n, B.data(), nrhs, X.data(), fo);
                else la.trsm(256, passes == 2, L.data(), n, B.data(), nrhs, X.data(), fo);
                lfo_solve(256, W.data(), n, B.data(), nrhs, Y.data(), 8, passes);
                REQUIRE(std::memcmp(X.data(), Y.data(), X.size() * sizeof(G)) == 0, "solve passes " << passes);
            }
            // in place: L may equal A
            std::vector<G> A2 = S;
            la.cholesky(256, A2.data(), n, A2.data(), fo);
            REQUIRE(std::memcmp(A2.data(), W.data(), A2.size() * sizeof(G)) == 0, "cholesky in place");
        }
        // errors
        bool threw = false;
        try {
            LinalgOptions bad;
            bad.band_bits = 2000;
            Linalg x(bad);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        REQUIRE(threw, "invalid options rejected");
        threw = false;
        try {
            la.syrk(64, a.data(), 3, 1, C.data(), false, true);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        REQUIRE(threw, "full syrk update rejected");
    } catch (const std::exception& e) {
        std::cerr << "exception: " << e.what() << "\n";
        return 2;
    }
    std::cout << (failures ? "FAILED " : "passed ") << failures << " failures\n";
    return failures ? 1 : 0;
}

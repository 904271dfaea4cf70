// Self-checking program for include/limbforge/mpfr_bridge.hpp (built and run by tests/test_mpfr_bridge.py):
// round trips, RN-even narrowing against the oracle's MPFR conversion (lfo_from_mpfr, the restatement of the
// reference's conversion), ties, special values, the status rule, the array and runtime-width forms on
// several thread counts, and the mpc_t forms.  Prints "<checks> checks, <failures> failures".
#include <limbforge/mpfr_bridge.hpp>

extern "C" {
#include "lf_oracle.h"
#include "lfo_mpfr.h"
}

#include <cstdio>
#include <random>
#include <vector>

using namespace limbforge;
static long checks = 0, failures = 0;
#define CHECK(cond, what)                                                                                             \
    do {                                                                                                              \
        ++checks;                                                                                                     \
        if (!(cond) && ++failures <= 20) std::printf("FAIL %s (line %d)\n", what, __LINE__);                          \
    } while (0)

template <int Bits> bool same(const Float<Bits>& a, const Float<Bits>& b) { return std::memcmp(&a, &b, sizeof a) == 0; }
template <int Bits> bool same_oracle(const Float<Bits>& a, const lfo_num& o)
{
    constexpr int N = Bits / 32;
    if (a.sign != o.sign || a.exponent != o.exponent || a.status != o.status) return false;
    for (int i = 0; i < N; ++i)
        if (a.limb[i] != o.limb[i]) return false;
    return true;
}

template <int Bits> Float<Bits> random_value(std::mt19937_64& g, int span)
{
    constexpr int N = Bits / 32;
    Float<Bits> r{};
    for (int i = 0; i < N; ++i) r.limb[i] = word(g());
    r.limb[N - 1] |= 0x80000000u;
    r.exponent = int(g() % (2 * span + 1)) - span;
    r.sign = (g() & 1) ? 1 : -1;
    return r;
}

template <int Bits> void width(std::mt19937_64& g)
{
    constexpr int N = Bits / 32;
    mpfr_t x, wide, narrow;
    mpfr_inits2(Bits, x, (mpfr_ptr)0);
    mpfr_init2(wide, Bits + 97);
    mpfr_init2(narrow, Bits / 2 + 3);
    for (int t = 0; t < 400; ++t) {
        const Float<Bits> a = random_value<Bits>(g, t % 3 ? 1000 : 999999990);
        to_mpfr<Bits>(x, a);
        CHECK(same<Bits>(from_mpfr<Bits>(x), a), "round trip at Bits");
        // a wider value with random low bits, and with an exact tie below the last kept bit
        to_mpfr<Bits>(wide, a);
        mpfr_t tail;
        mpfr_init2(tail, 128);
        mpfr_set_ui(tail, (unsigned long)(g() | 1), MPFR_RNDN);
        const bool tie = t % 4 == 0;
        if (tie) mpfr_set_ui(tail, 1, MPFR_RNDN);
        mpfr_mul_2si(tail, tail, (long)a.exponent - Bits - (tie ? 0 : 64), MPFR_RNDN);
        if (a.sign < 0) mpfr_neg(tail, tail, MPFR_RNDN);
        if (t % 8 == 1) mpfr_neg(tail, tail, MPFR_RNDN);
        mpfr_add(wide, wide, tail, MPFR_RNDN);
        mpfr_clear(tail);
        CHECK(same_oracle<Bits>(from_mpfr<Bits>(wide), lfo_from_mpfr(N, wide)), "RN-even from a wider value");
        // narrowing into a smaller MPFR precision is one RN-even rounding of the exact value
        to_mpfr<Bits>(narrow, a);
        mpfr_t want;
        mpfr_init2(want, Bits / 2 + 3);
        mpfr_set(want, x, MPFR_RNDN);
        CHECK(mpfr_equal_p(narrow, want), "to_mpfr into a narrower precision");
        mpfr_clear(want);
    }
    // special values and limits
    mpfr_set_nan(x);
    CHECK(from_mpfr<Bits>(x).status == invalid, "NaN");
    mpfr_set_inf(x, -1);
    CHECK(from_mpfr<Bits>(x).status == invalid, "Inf");
    mpfr_set_zero(x, -1);
    CHECK(same<Bits>(from_mpfr<Bits>(x), zero<N>()), "signed zero is canonical zero");
    mpfr_set_ui_2exp(x, 1, 1000000000, MPFR_RNDN);
    Float<Bits> top = from_mpfr<Bits>(x);
    CHECK(top.status == ok && top.exponent == 1000000000, "largest exponent");
    mpfr_set_ui_2exp(x, 1, 1000000001, MPFR_RNDN);
    CHECK(from_mpfr<Bits>(x).status == exponent_overflow, "exponent overflow");
    mpfr_set_ui_2exp(x, 1, -1000000001, MPFR_RNDN);
    CHECK(from_mpfr<Bits>(x).status == exponent_overflow, "exponent underflow");
    bool threw = false;
    try {
        to_mpfr<Bits>(x, zero<N>(division_by_zero));
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw, "to_mpfr throws on a status");
    // arrays and the runtime-width forms agree with the scalar conversions on any thread count
    const std::size_t n = 3000;
    std::vector<Float<Bits>> v(n), back(n), rt(n);
    std::vector<mpfr_t> m(n);
    for (auto& y : m) mpfr_init2(y, Bits);
    for (auto& y : v) y = random_value<Bits>(g, 5000);
    for (unsigned threads : {1u, 4u, 0u}) {
        to_mpfr_array<Bits>(m.data(), v.data(), n, threads);
        from_mpfr_array<Bits>(m.data(), back.data(), n, threads);
        bool ok_all = true;
        for (std::size_t i = 0; i < n; ++i) ok_all &= same<Bits>(back[i], v[i]);
        CHECK(ok_all, "array round trip");
        from_mpfr_array(Bits, m.data(), rt.data(), n, threads);
        CHECK(std::memcmp(rt.data(), v.data(), n * sizeof(Float<Bits>)) == 0, "runtime-width array");
    }
    for (auto& y : m) mpfr_clear(y);
#ifdef LIMBFORGE_HAS_MPC
    mpc_t z;
    mpc_init2(z, Bits);
    const Complex<N> c{random_value<Bits>(g, 100), random_value<Bits>(g, 100)};
    to_mpc<Bits>(z, c);
    const Complex<N> c2 = from_mpc<Bits>(z);
    CHECK(same<Bits>(c2.re, c.re) && same<Bits>(c2.im, c.im), "mpc round trip");
    Complex<N> c3;
    from_mpc(Bits, z, &c3);
    CHECK(same<Bits>(c3.re, c.re) && same<Bits>(c3.im, c.im), "runtime-width mpc");
    mpc_clear(z);
#endif
    mpfr_clears(x, wide, narrow, (mpfr_ptr)0);
}

int main()
{
    lfo_mpfr_range();
    std::mt19937_64 g(20261203);
    width<64>(g);
    width<96>(g);
    width<256>(g);
    width<512>(g);
    width<1024>(g);
    const Float<64> h = from_decimal<64>("1.5");
    CHECK(h.limb[1] == 0xC0000000u && h.limb[0] == 0 && h.exponent == 0 && h.sign == 1, "decimal 1.5");
    const Float<64> t = from_decimal<64>("-0.1");
    CHECK(t.sign == -1 && t.exponent == -4, "decimal -0.1");
    std::printf("%ld checks, %ld failures\n", checks, failures);
    return failures != 0;
}

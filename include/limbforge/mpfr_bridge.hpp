#pragma once
// Conversions between the LimbForge number formats and MPFR / MPC values on the host (the reference's
// include/limbforge/mpfr_bridge.hpp contract).  Header-only; link MPFR and GMP when using it.  The mpc_t
// forms need <mpc.h> on the include path and no libmpc symbols; define LIMBFORGE_NO_MPC to leave them out.
//
//   from_mpfr<Bits>(x)   RN-even to Bits; NaN/Inf -> zero with status invalid, zero -> canonical zero,
//                        exponent (MPFR exponent - 1) outside [-1e9, 1e9] -> zero with exponent_overflow
//   to_mpfr<Bits>(y, x)  exact when y's precision is >= Bits, otherwise RN to that precision; throws
//                        std::runtime_error for a value that carries a status
//   *_array              the same per element over contiguous arrays, split over host threads (0: chosen
//                        from the array size); workers use the caller's MPFR exponent range
//   runtime-width forms  the same on untyped element storage of bridge_element_bytes(bits, complex) bytes
#include "core.hpp"

#include <mpfr.h>
#if !defined(LIMBFORGE_NO_MPC) && __has_include(<mpc.h>)
#include <mpc.h>
#define LIMBFORGE_HAS_MPC 1
#endif

#include <algorithm>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace limbforge {
namespace bridge_detail {
struct Range {
    mpfr_exp_t emin, emax;
};
inline Range caller_range() { return {mpfr_get_emin(), mpfr_get_emax()}; }

template <int Bits> inline Float<Bits> from_mpfr_value(mpfr_srcptr in)
{
    static_assert(Bits >= 64 && Bits <= 1024 && Bits % 32 == 0, "precision must be a multiple of 32 in [64,1024]");
    constexpr int N = Bits / 32;
    if (!mpfr_number_p(in)) return zero<N>(invalid);
    if (mpfr_zero_p(in)) return zero<N>();
    // round to Bits in an unbounded exponent range, then read the integer significand
    const mpfr_exp_t emin = mpfr_get_emin(), emax = mpfr_get_emax();
    mpfr_set_emin(mpfr_get_emin_min());
    mpfr_set_emax(mpfr_get_emax_max());
    mpfr_t x;
    mpfr_init2(x, Bits);
    mpfr_set(x, in, MPFR_RNDN);
    const mpfr_exp_t e = mpfr_get_exp(x) - 1;
    Float<Bits> r = zero<N>();
    if (e > 1000000000 || e < -1000000000) {
        r = zero<N>(exponent_overflow);
    } else {
        mpz_t z;
        mpz_init(z);
        mpfr_get_z_2exp(z, x); // x = z * 2^k with |z| < 2^Bits and the top bit of |z| set
        mpz_abs(z, z);
        std::size_t words = 0;
        mpz_export(r.limb, &words, -1, sizeof(word), 0, 0, z);
        mpz_clear(z);
        r.sign = mpfr_sgn(x) < 0 ? -1 : 1;
        r.exponent = int(e);
    }
    mpfr_clear(x);
    mpfr_set_emin(emin);
    mpfr_set_emax(emax);
    return r;
}

template <int Bits> inline void to_mpfr_value(mpfr_ptr out, const Float<Bits>& x)
{
    constexpr int N = Bits / 32;
    if (x.status) throw std::runtime_error("limbforge arithmetic status " + std::to_string(x.status));
    if (!x.sign) {
        mpfr_set_zero(out, 1);
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpz_import(z, N, -1, sizeof(word), 0, 0, x.limb);
    if (x.sign < 0) mpz_neg(z, z);
    // one rounding (none when the output holds Bits bits): z * 2^(exponent - (Bits - 1))
    mpfr_set_z_2exp(out, z, mpfr_exp_t(x.exponent) - (Bits - 1), MPFR_RNDN);
    mpz_clear(z);
}

// Runs f(begin, end) over contiguous chunks of [0, n) on `threads` host threads (0: one thread per `grain`
// elements, at most the hardware concurrency); each worker takes the caller's MPFR exponent range, and the
// first chunk's failure (in chunk order) is rethrown once every chunk has finished.
template <class F> inline void chunks(std::size_t n, unsigned threads, std::size_t grain, F&& f)
{
    if (!n) return;
    const Range range = caller_range();
    if (!threads) {
        const std::size_t by_size = std::max<std::size_t>(1, n / std::max<std::size_t>(grain, 1));
        threads = unsigned(std::min<std::size_t>(by_size, std::max(1u, std::thread::hardware_concurrency())));
    }
    threads = unsigned(std::min<std::size_t>(threads, n));
    if (threads <= 1) {
        f(std::size_t(0), n);
        return;
    }
    std::vector<std::exception_ptr> failed(threads);
    std::vector<std::thread> pool;
    const std::size_t per = (n + threads - 1) / threads;
    for (unsigned t = 1; t < threads; ++t) {
        const std::size_t b = t * per, e = std::min(n, b + per);
        if (b >= e) break;
        pool.emplace_back([&f, &failed, range, t, b, e] {
            try {
                mpfr_set_emin(range.emin);
                mpfr_set_emax(range.emax);
                f(b, e);
            } catch (...) {
                failed[t] = std::current_exception();
            }
            mpfr_free_cache2(MPFR_FREE_LOCAL_CACHE);
        });
    }
    try {
        f(std::size_t(0), std::min(n, per));
    } catch (...) {
        failed[0] = std::current_exception();
    }
    for (auto& w : pool) w.join();
    for (auto& e : failed)
        if (e) std::rethrow_exception(e);
}

template <int Bits = 64, class F> inline void by_width(int bits, F&& f)
{
    if (bits == Bits) {
        f(std::integral_constant<int, Bits>{});
        return;
    }
    if constexpr (Bits < 1024) by_width<Bits + 32>(bits, f);
}
template <class T> inline T load(const void* p)
{
    T v;
    std::memcpy(&v, p, sizeof v);
    return v;
}
template <class T> inline void store(void* p, const T& v) { std::memcpy(p, &v, sizeof v); }
} // namespace bridge_detail

template <int Bits> inline Float<Bits> from_mpfr(mpfr_srcptr x) { return bridge_detail::from_mpfr_value<Bits>(x); }
template <int Bits> inline void to_mpfr(mpfr_ptr out, const Float<Bits>& x) { bridge_detail::to_mpfr_value<Bits>(out, x); }

template <int Bits> inline void from_mpfr_array(const mpfr_t* in, Float<Bits>* out, std::size_t n, unsigned threads = 0)
{
    bridge_detail::chunks(n, threads, (std::size_t(1) << 22) / Bits, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) out[i] = from_mpfr<Bits>(in[i]);
    });
}
// On a value with a status the elements of other chunks may already be written.
template <int Bits> inline void to_mpfr_array(mpfr_t* out, const Float<Bits>* in, std::size_t n, unsigned threads = 0)
{
    bridge_detail::chunks(n, threads, (std::size_t(1) << 22) / Bits, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) to_mpfr<Bits>(out[i], in[i]);
    });
}

#ifdef LIMBFORGE_HAS_MPC
template <int Bits> inline Complex<Bits / 32> from_mpc(mpc_srcptr z) { return {from_mpfr<Bits>(mpc_realref(z)), from_mpfr<Bits>(mpc_imagref(z))}; }
// Throws before writing anything when either part carries a status.
template <int Bits> inline void to_mpc(mpc_ptr out, const Complex<Bits / 32>& x)
{
    if (x.re.status | x.im.status) throw std::runtime_error("limbforge arithmetic status " + std::to_string(x.re.status | x.im.status));
    to_mpfr<Bits>(mpc_realref(out), x.re);
    to_mpfr<Bits>(mpc_imagref(out), x.im);
}
template <int Bits> inline void from_mpc_array(const mpc_t* in, Complex<Bits / 32>* out, std::size_t n, unsigned threads = 0)
{
    bridge_detail::chunks(n, threads, (std::size_t(1) << 21) / Bits, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) out[i] = from_mpc<Bits>(in[i]);
    });
}
template <int Bits> inline void to_mpc_array(mpc_t* out, const Complex<Bits / 32>* in, std::size_t n, unsigned threads = 0)
{
    bridge_detail::chunks(n, threads, (std::size_t(1) << 21) / Bits, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) to_mpc<Bits>(out[i], in[i]);
    });
}
#endif

// Runtime-width forms: element storage holds the same Float<bits> / Complex<bits/32> records as the typed API
// (no alignment required).
inline std::size_t bridge_element_bytes(int bits, bool complex = false)
{
    if (bits < 64 || bits > 1024 || bits % 32) throw std::invalid_argument("bridge bits must be a multiple of 32 in [64,1024]");
    return std::size_t(bits / 8 + 12) * (complex ? 2 : 1);
}
inline void from_mpfr(int bits, mpfr_srcptr in, void* out)
{
    bridge_element_bytes(bits);
    if (!in || !out) throw std::invalid_argument("null bridge operand");
    bridge_detail::by_width(bits, [&](auto b) { bridge_detail::store(out, from_mpfr<b.value>(in)); });
}
inline void to_mpfr(int bits, mpfr_ptr out, const void* in)
{
    bridge_element_bytes(bits);
    if (!in || !out) throw std::invalid_argument("null bridge operand");
    bridge_detail::by_width(bits, [&](auto b) { to_mpfr<b.value>(out, bridge_detail::load<Float<b.value>>(in)); });
}
inline void from_mpfr_array(int bits, const mpfr_t* in, void* out, std::size_t count, unsigned threads = 0)
{
    const std::size_t bytes = bridge_element_bytes(bits);
    if (count && (!in || !out)) throw std::invalid_argument("null bridge array");
    if (count > std::size_t(-1) / bytes) throw std::invalid_argument("bridge array size overflow");
    bridge_detail::by_width(bits, [&](auto b) {
        bridge_detail::chunks(count, threads, (std::size_t(1) << 22) / b.value, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t i = lo; i < hi; ++i) bridge_detail::store(static_cast<char*>(out) + i * bytes, from_mpfr<b.value>(in[i]));
        });
    });
}
inline void to_mpfr_array(int bits, mpfr_t* out, const void* in, std::size_t count, unsigned threads = 0)
{
    const std::size_t bytes = bridge_element_bytes(bits);
    if (count && (!in || !out)) throw std::invalid_argument("null bridge array");
    if (count > std::size_t(-1) / bytes) throw std::invalid_argument("bridge array size overflow");
    bridge_detail::by_width(bits, [&](auto b) {
        bridge_detail::chunks(count, threads, (std::size_t(1) << 22) / b.value, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t i = lo; i < hi; ++i)
                to_mpfr<b.value>(out[i], bridge_detail::load<Float<b.value>>(static_cast<const char*>(in) + i * bytes));
        });
    });
}
#ifdef LIMBFORGE_HAS_MPC
inline void from_mpc(int bits, mpc_srcptr in, void* out)
{
    bridge_element_bytes(bits, true);
    if (!in || !out) throw std::invalid_argument("null bridge operand");
    bridge_detail::by_width(bits, [&](auto b) { bridge_detail::store(out, from_mpc<b.value>(in)); });
}
inline void to_mpc(int bits, mpc_ptr out, const void* in)
{
    bridge_element_bytes(bits, true);
    if (!in || !out) throw std::invalid_argument("null bridge operand");
    bridge_detail::by_width(bits, [&](auto b) { to_mpc<b.value>(out, bridge_detail::load<Complex<b.value / 32>>(in)); });
}
inline void from_mpc_array(int bits, const mpc_t* in, void* out, std::size_t count, unsigned threads = 0)
{
    const std::size_t bytes = bridge_element_bytes(bits, true);
    if (count && (!in || !out)) throw std::invalid_argument("null bridge array");
    if (count > std::size_t(-1) / bytes) throw std::invalid_argument("bridge array size overflow");
    bridge_detail::by_width(bits, [&](auto b) {
        bridge_detail::chunks(count, threads, (std::size_t(1) << 21) / b.value, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t i = lo; i < hi; ++i) bridge_detail::store(static_cast<char*>(out) + i * bytes, from_mpc<b.value>(in[i]));
        });
    });
}
inline void to_mpc_array(int bits, mpc_t* out, const void* in, std::size_t count, unsigned threads = 0)
{
    const std::size_t bytes = bridge_element_bytes(bits, true);
    if (count && (!in || !out)) throw std::invalid_argument("null bridge array");
    if (count > std::size_t(-1) / bytes) throw std::invalid_argument("bridge array size overflow");
    bridge_detail::by_width(bits, [&](auto b) {
        bridge_detail::chunks(count, threads, (std::size_t(1) << 21) / b.value, [&](std::size_t lo, std::size_t hi) {
            for (std::size_t i = lo; i < hi; ++i)
                to_mpc<b.value>(out[i], bridge_detail::load<Complex<b.value / 32>>(static_cast<const char*>(in) + i * bytes));
        });
    });
}
#endif

// A decimal string rounded once to Bits (RN-even).
template <int Bits> inline Float<Bits> from_decimal(const std::string& text)
{
    mpfr_t x;
    mpfr_init2(x, Bits);
    if (mpfr_set_str(x, text.c_str(), 10, MPFR_RNDN)) {
        mpfr_clear(x);
        throw std::invalid_argument("invalid decimal number");
    }
    const Float<Bits> r = from_mpfr<Bits>(x);
    mpfr_clear(x);
    return r;
}
} // namespace limbforge

#pragma once
// Number formats and scalar arithmetic of the LimbForge/WolfNum API (the reference's
// include/limbforge/core.hpp contract):
//   value = sign * integer(limb) * 2^(exponent - (32N - 1)), limbs little-endian, the top bit of limb[N-1] set
//   for every nonzero value, exponent in [-1e9, 1e9], sign in {-1, 0, 1}, canonical zero (all fields 0 except
//   the status bitmask).  Every operation rounds to nearest, ties to even; arithmetic failures are status bits
//   in the result (division by zero 1, exponent overflow 2, invalid 4), never exceptions.
// The scalar functions below (add ... cdiv, dot2_add, cfma, cfms) run on the host, in liblimbforge.so: they are the arithmetic core
// of the GPU kernels built for the CPU, so a host result is the same bits as the GPU's for the same operands.
// N is the limb count, 2..32 (64 to 1024 bits).
#include <cstdint>
#include "version.hpp"

namespace limbforge {
using word = std::uint32_t;
using dword = std::uint64_t;
using exponent_type = std::int64_t;

enum Status : word { ok = 0, division_by_zero = 1, exponent_overflow = 2, invalid = 4 };

template <int N> struct Number {
    word limb[N];  // little endian; top bit set for every nonzero value
    int exponent;  // value = sign * mantissa * 2^(exponent - (32*N-1))
    int sign;      // -1, 0, +1; canonical unsigned zero
    word status;
};

template <int Bits> struct FloatType {
    static_assert(Bits >= 64 && Bits <= 1024 && Bits % 32 == 0, "precision must be a multiple of 32 in [64,1024]");
    using type = Number<Bits / 32>;
};
template <int Bits> using Float = typename FloatType<Bits>::type;
template <int N> struct Complex {
    Number<N> re, im;
};

template <int N> inline Number<N> zero(word status = ok)
{
    Number<N> r{};
    r.status = status;
    return r;
}
template <int N> inline Number<N> negate(Number<N> a)
{
    a.sign = -a.sign;
    return a;
}

// Compares |a| and |b| (exponent, then limbs from the top): -1, 0 or 1.  Zero is the smallest magnitude.
template <int N> inline int magnitude_compare(const Number<N>& a, const Number<N>& b)
{
    if (!a.sign || !b.sign) return (a.sign != 0) - (b.sign != 0);
    if (a.exponent != b.exponent) return a.exponent < b.exponent ? -1 : 1;
    for (int i = N - 1; i >= 0; --i)
        if (a.limb[i] != b.limb[i]) return a.limb[i] < b.limb[i] ? -1 : 1;
    return 0;
}

// Quotient and remainder of the two-word value (high, low) by a normalized word d (d >= 2^31, high < d),
// with the reused reciprocal inv = floor((2^64 - 1) / d) - 2^32: the 2/1 division step of Moller and
// Granlund, "Improved division by invariant integers" (IEEE Trans. Computers, 2011).
struct WordDivision {
    word quotient, remainder;
};
inline WordDivision divide_word(word high, word low, word d, word inv)
{
    const dword p = dword(high) * inv + ((dword(high) << 32) | low);
    word q = word(p >> 32) + 1;
    word r = low - q * d;
    if (r > word(p)) {
        --q;
        r += d;
    }
    if (r >= d) {
        ++q;
        r -= d;
    }
    return {q, r};
}

namespace detail {
// The host arithmetic of liblimbforge.so: out = op(a, b, c) on records of `words` limbs, op numbered as
// limbforge::Operation (add 0, sub 1, mul 2, div 3, complex add/mul/div 4-6, square 7, sqrt 8, fma 9, fms 10,
// complex fma / fms 11-12).
void host_arithmetic(int op, int words, const void* a, const void* b, const void* c, void* out);
// out = RN(a b + c d + e), one rounding, on records of `words` limbs
void host_dot2_add(int words, const void* a, const void* b, const void* c, const void* d, const void* e, void* out);
template <int N> inline Number<N> host_real(int op, const Number<N>& a, const Number<N>& b, const Number<N>& c)
{
    static_assert(N >= 2 && N <= 32, "limb count must be in [2, 32]");
    Number<N> r;
    host_arithmetic(op, N, &a, &b, &c, &r);
    return r;
}
template <int N> inline Complex<N> host_complex(int op, const Complex<N>& a, const Complex<N>& b, const Complex<N>* c = nullptr)
{
    static_assert(N >= 2 && N <= 32, "limb count must be in [2, 32]");
    Complex<N> r;
    host_arithmetic(op, N, &a, &b, c, &r);
    return r;
}
} // namespace detail

// RN(a + b), RN(a - b)
template <int N> inline Number<N> add(Number<N> a, Number<N> b) { return detail::host_real<N>(0, a, b, a); }
template <int N> inline Number<N> sub(Number<N> a, Number<N> b) { return detail::host_real<N>(1, a, b, a); }
// RN(a * b).  Padded selects a Metal register layout in the reference and does not change the result.
template <int N, bool Padded = true> inline Number<N> mul(const Number<N>& a, const Number<N>& b)
{
    return detail::host_real<N>(2, a, b, a);
}
template <int N> inline Number<N> square(const Number<N>& a) { return detail::host_real<N>(7, a, a, a); }
// RN(a / b); b = 0 gives zero with status division_by_zero.
template <int N, bool Padded = true> inline Number<N> div(const Number<N>& a, const Number<N>& b)
{
    return detail::host_real<N>(3, a, b, a);
}
// RN(sqrt(a)); a < 0 gives zero with status invalid.
template <int N> inline Number<N> sqrt(const Number<N>& a) { return detail::host_real<N>(8, a, a, a); }
// RN(a * b + c) and RN(a * b - c), one rounding.
template <int N> inline Number<N> fma(const Number<N>& a, const Number<N>& b, const Number<N>& c)
{
    return detail::host_real<N>(9, a, b, c);
}
template <int N> inline Number<N> fms(const Number<N>& a, const Number<N>& b, const Number<N>& c)
{
    return detail::host_real<N>(10, a, b, c);
}
// Composed complex operations, every real primitive rounded (the reference's cadd/cmul/cdiv).
template <int N> inline Complex<N> cadd(Complex<N> a, Complex<N> b) { return detail::host_complex<N>(4, a, b); }
template <int N> inline Complex<N> cmul(Complex<N> a, Complex<N> b) { return detail::host_complex<N>(5, a, b); }
template <int N> inline Complex<N> cdiv(Complex<N> a, Complex<N> b) { return detail::host_complex<N>(6, a, b); }
// RN(a*b + c*d + e) with one rounding.
template <int N>
inline Number<N> dot2_add(const Number<N>& a, const Number<N>& b, const Number<N>& c, const Number<N>& d, const Number<N>& e)
{
    static_assert(N >= 2 && N <= 32, "limb count must be in [2, 32]");
    Number<N> r;
    detail::host_dot2_add(N, &a, &b, &c, &d, &e, &r);
    return r;
}
// Complex a*b + c and a*b - c, each part one rounding of its exact value.
template <int N> inline Complex<N> cfma(Complex<N> a, Complex<N> b, Complex<N> c) { return detail::host_complex<N>(11, a, b, &c); }
template <int N> inline Complex<N> cfms(Complex<N> a, Complex<N> b, Complex<N> c) { return detail::host_complex<N>(12, a, b, &c); }
} // namespace limbforge

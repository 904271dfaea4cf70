#pragma once
// Number formats of the LimbForge/WolfNum API (the reference's include/limbforge/core.hpp:10-50 contract):
//   value = sign * integer(limb) * 2^(exponent - (32N - 1)), limbs little-endian, the top bit of limb[N-1] set
//   for every nonzero value, exponent in [-1e9, 1e9], sign in {-1, 0, 1}, canonical zero (all fields 0 except
//   the status bitmask).  Every operation rounds to nearest, ties to even.
// The arithmetic itself runs on the GPU (Engine, Linalg): this MI355X build does not provide the reference's
// host-side scalar functions (add, mul, div, ... on the CPU).
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
} // namespace limbforge

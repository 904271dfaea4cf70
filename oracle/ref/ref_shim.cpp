// ref_shim.cpp — a C entry layer over the reference's OWN header-only CPU arithmetic
// (include/limbforge/core.hpp and the exact_dot_add template of include/limbforge/linalg.hpp), compiled
// from /root/reference by oracle/Makefile into oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: it
// validates the C restatement in lf_oracle.c and serves as the reference CPU baseline.  No reference
// source is copied here; the headers are included from where they lie.
#include "limbforge/core.hpp"
#include "limbforge/linalg.hpp"
#include <cstring>
#include <type_traits>

using namespace limbforge;

namespace {
template <int N, class F> auto with_words(int words, F&& f)
{
    if constexpr (N < 32) {
        if (words != N) return with_words<N + 1>(words, f);
    }
    return f(std::integral_constant<int, N>{});
}

inline std::size_t index_of(std::size_t i, std::size_t stride, std::size_t period)
{
    std::size_t j = stride > 1 ? i / stride : i;
    return period ? j % period : j;
}

template <int N>
void elementwise(int op, const unsigned char* a, const unsigned char* b, const unsigned char* c, unsigned char* out,
                 std::size_t count, std::size_t bs, std::size_t bp, std::size_t cs, std::size_t cp)
{
    using T = Number<N>;
    using C = Complex<N>;
    const bool cplx = op == 4 || op == 5 || op == 6 || op == 11 || op == 12;
    for (std::size_t i = 0; i < count; ++i) {
        std::size_t jb = index_of(i, bs, bp), jc = index_of(i, cs, cp);
        if (cplx) {
            C x, y, z;
            std::memcpy(&x, a + i * sizeof(C), sizeof(C));
            std::memcpy(&y, b + jb * sizeof(C), sizeof(C));
            if (op == 11 || op == 12) {
                C w;
                std::memcpy(&w, c + jc * sizeof(C), sizeof(C));
                z = op == 11 ? cfma(x, y, w) : cfms(x, y, w);
            } else
                z = op == 4 ? cadd(x, y) : op == 5 ? cmul(x, y) : cdiv(x, y);
            std::memcpy(out + i * sizeof(C), &z, sizeof(C));
            continue;
        }
        T x, y, z, r;
        std::memcpy(&x, a + i * sizeof(T), sizeof(T));
        std::memcpy(&y, b + jb * sizeof(T), sizeof(T));
        std::memcpy(&z, c + jc * sizeof(T), sizeof(T));
        switch (op) {
        case 0: r = add(x, y); break;
        case 1: r = sub(x, y); break;
        case 2: r = mul(x, y); break;
        case 3: r = div(x, y); break;
        case 7: r = square(x); break;
        case 8: r = limbforge::sqrt(x); break;
        case 9: r = limbforge::fma(x, y, z); break;
        default: r = limbforge::fms(x, y, z); break;
        }
        std::memcpy(out + i * sizeof(T), &r, sizeof(T));
    }
}
} // namespace

extern "C" int ref_elementwise(int bits, int op, const void* a, const void* b, const void* c, void* out, std::size_t count,
                               std::size_t bs, std::size_t bp, std::size_t cs, std::size_t cp)
{
    if (bits < 64 || bits > 1024 || bits % 32 || op < 0 || op > 12) return -1;
    with_words<2>(bits / 32, [&](auto w) {
        elementwise<decltype(w)::value>(op, static_cast<const unsigned char*>(a), static_cast<const unsigned char*>(b),
                                        static_cast<const unsigned char*>(c), static_cast<unsigned char*>(out), count, bs, bp, cs, cp);
    });
    return 0;
}

extern "C" int ref_mul_chain(int bits, const void* x0, const void* f, void* out, std::size_t count, int steps)
{
    if (bits < 64 || bits > 1024 || bits % 32) return -1;
    with_words<2>(bits / 32, [&](auto w) {
        constexpr int N = decltype(w)::value;
        using T = Number<N>;
        for (std::size_t i = 0; i < count; ++i) {
            T x, y;
            std::memcpy(&x, static_cast<const unsigned char*>(x0) + i * sizeof(T), sizeof(T));
            std::memcpy(&y, static_cast<const unsigned char*>(f) + i * sizeof(T), sizeof(T));
            for (int s = 0; s < steps; ++s) x = mul(x, y);
            std::memcpy(static_cast<unsigned char*>(out) + i * sizeof(T), &x, sizeof(T));
        }
    });
    return 0;
}

extern "C" void ref_exact_dot_add(int bits, const void* c, int subtract, const void* a, std::ptrdiff_t sa, const void* b,
                                  std::ptrdiff_t sb, std::size_t K, void* out)
{
    with_words<2>(bits / 32, [&](auto w) {
        constexpr int N = decltype(w)::value;
        using T = Number<N>;
        T cc{}, r;
        if (c) std::memcpy(&cc, c, sizeof(T));
        r = exact_dot_add<N>(c ? &cc : nullptr, subtract != 0, static_cast<const T*>(a), sa, static_cast<const T*>(b), sb, K);
        std::memcpy(out, &r, sizeof(T));
    });
}

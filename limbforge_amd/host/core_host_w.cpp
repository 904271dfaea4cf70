// One width of the host arithmetic behind include/limbforge/core.hpp: the GPU kernels' arithmetic core
// (limb_core.hpp) built as plain C++ (LF_HOST_BUILD), compiled once per LF_WORDS in [2, 32] so the 31
// widths build in parallel.  Records are the public element layout (N limbs, exponent, sign, status;
// a complex value is two of them), which is lf::Num<N>'s layout.
#include "dot2_core.hpp"
#include "limb_core.hpp"
#include "core_host.hpp"

#include <cstring>

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {
namespace {
template <int N> Num<N> get(const void* p)
{
    Num<N> r;
    std::memcpy(&r, p, sizeof r);
    return r;
}
template <int N> Cpx<N> get_cpx(const void* p)
{
    Cpx<N> r;
    std::memcpy(&r.re, p, sizeof(Num<N>));
    std::memcpy(&r.im, (const char*)p + sizeof(Num<N>), sizeof(Num<N>));
    return r;
}
template <int N> void put(void* p, const Num<N>& r) { std::memcpy(p, &r, sizeof r); }
template <int N> void put_cpx(void* p, const Cpx<N>& r)
{
    std::memcpy(p, &r.re, sizeof(Num<N>));
    std::memcpy((char*)p + sizeof(Num<N>), &r.im, sizeof(Num<N>));
}
} // namespace

template <int N> int host_arithmetic(int op, const void* a, const void* b, const void* c, void* out)
{
    static_assert(sizeof(Num<N>) == 4 * N + 12, "lf::Num is the public element record");
    switch (op) {
    case HOST_ADD: put<N>(out, add<N>(get<N>(a), get<N>(b))); return 0;
    case HOST_SUB: put<N>(out, sub<N>(get<N>(a), get<N>(b))); return 0;
    case HOST_MUL: put<N>(out, mul<N>(get<N>(a), get<N>(b))); return 0;
    case HOST_DIV: put<N>(out, div<N>(get<N>(a), get<N>(b))); return 0;
    case HOST_CADD: put_cpx<N>(out, cadd<N>(get_cpx<N>(a), get_cpx<N>(b))); return 0;
    case HOST_CMUL: put_cpx<N>(out, cmul<N>(get_cpx<N>(a), get_cpx<N>(b))); return 0;
    case HOST_CDIV: put_cpx<N>(out, cdiv<N>(get_cpx<N>(a), get_cpx<N>(b))); return 0;
    case HOST_SQUARE: {
        const Num<N> x = get<N>(a);
        put<N>(out, mul<N>(x, x));
        return 0;
    }
    case HOST_SQRT: put<N>(out, lf::sqrt<N>(get<N>(a))); return 0;
    case HOST_FMA: put<N>(out, lf::fma<N>(get<N>(a), get<N>(b), get<N>(c))); return 0;
    case HOST_FMS: put<N>(out, fms<N>(get<N>(a), get<N>(b), get<N>(c))); return 0;
    case HOST_CFMA: put_cpx<N>(out, lf::cfma<N>(get_cpx<N>(a), get_cpx<N>(b), get_cpx<N>(c))); return 0;
    case HOST_CFMS: put_cpx<N>(out, lf::cfms<N>(get_cpx<N>(a), get_cpx<N>(b), get_cpx<N>(c))); return 0;
    default: return -1;
    }
}

template <int N> void host_dot2(const void* const* x, void* out)
{
    put<N>(out, dot2_add<N>(get<N>(x[0]), get<N>(x[1]), get<N>(x[2]), get<N>(x[3]), get<N>(x[4])));
}

template int host_arithmetic<LF_WORDS>(int op, const void* a, const void* b, const void* c, void* out);
template void host_dot2<LF_WORDS>(const void* const* x, void* out);
} // namespace lf

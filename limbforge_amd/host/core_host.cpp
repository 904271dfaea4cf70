// core_host.cpp — dispatch of the public scalar arithmetic (include/limbforge/core.hpp) to the per-width
// host builds of the arithmetic core (core_host_w.cpp).
#include "limbforge/core.hpp"
#include "core_host.hpp"

#include <array>
#include <stdexcept>
#include <string>
#include <utility>

namespace lf {
namespace {
using Fn = int (*)(int, const void*, const void*, const void*, void*);
using Dot2 = void (*)(const void* const*, void*);
template <int... I> constexpr auto table(std::integer_sequence<int, I...>)
{
    return std::array<Fn, sizeof...(I)>{&host_arithmetic<I + 2>...};
}
template <int... I> constexpr auto dot2_table(std::integer_sequence<int, I...>)
{
    return std::array<Dot2, sizeof...(I)>{&host_dot2<I + 2>...};
}
} // namespace
} // namespace lf

namespace limbforge {
namespace detail {
void host_arithmetic(int op, int words, const void* a, const void* b, const void* c, void* out)
{
    static const auto fns = lf::table(std::make_integer_sequence<int, 31>{});
    if (words < 2 || words > 32) throw std::invalid_argument("limbforge: limb count must be in [2, 32]");
    if (fns[words - 2](op, a, b, c, out)) throw std::invalid_argument("limbforge: unknown operation " + std::to_string(op));
}

void host_dot2_add(int words, const void* a, const void* b, const void* c, const void* d, const void* e, void* out)
{
    static const auto fns = lf::dot2_table(std::make_integer_sequence<int, 31>{});
    if (words < 2 || words > 32) throw std::invalid_argument("limbforge: limb count must be in [2, 32]");
    const void* x[5] = {a, b, c, d, e};
    fns[words - 2](x, out);
}
} // namespace detail
} // namespace limbforge

// mixed_chain.hpp — one value of the mixed-precision chains: x <- RN(x f + c_t) for t < len, c_t = g where
// bit t of the value's mask is set and zero otherwise (an fma step or a mul step), one rounding per step.
//
// Widths up to 14 limbs run each step as a call: inlined into the chain loop, the step's divergent cases
// (zero and status operands, the addend far above or below the product, the window) made the register
// allocator keep several copies of the window live around the loop back edge, which cost 240-256
// registers and 260-1,924 bytes of spills per lane at 7-14 limbs (against 85-105 registers for one
// step without the loop).  The called step needs 50-170 registers and no spills; from 15 limbs on the
// inlined loop needs no spills and the call would only add its argument traffic.
#pragma once
#include "coop_core.hpp"
#include "elementwise_impl.hpp"
#include "lf_internal.hpp"

namespace lf {

template <int N> struct ChainDigits {
    u32 d[use_digits<N>() ? digit_count<N>() : 1];
};

template <int N> __device__ __noinline__ Num<N> chain_step(Num<N> x, const Num<N> f, const ChainDigits<N> fd, const Num<N> c)
{
    if constexpr (use_digits<N>()) return fma_by_digits<N>(x, f, fd.d, c);
    else return fma<N>(x, f, c);
}

template <int N> __device__ __forceinline__ void chain_value(const MixedArgs& a, u32 i)
{
    Num<N> x = load_num<N>(a.x, a.pitch, i);
    const Num<N> f = load_num<N>(a.f, a.pitch, i);
    const Num<N> g = load_num<N>(a.g, a.pitch, i);
    const u32 len = a.len[i];
    const u64 mask = a.mask[i];
    if constexpr (N <= 14) {
        ChainDigits<N> fd{};
        if constexpr (use_digits<N>()) to_digits<N>(f.w, fd.d);
        for (u32 t = 0; t < len; ++t) x = chain_step<N>(x, f, fd, ((mask >> t) & 1u) ? g : zero<N>(ST_OK));
    } else {
        u32 fd[digit_count<N>()]; // the factor's digits, formed once for the chain's multiplications
        to_digits<N>(f.w, fd);
        for (u32 t = 0; t < len; ++t) x = fma_by_digits<N>(x, f, fd, ((mask >> t) & 1u) ? g : zero<N>(ST_OK));
    }
    store_num<N>(a.out, a.pitch, i, x);
}

// Value i's chain on lane k of a group of G lanes (coop_core.hpp): every step's product columns are
// split over the group; lane 0 stores the result.
template <int N, int G> __device__ __forceinline__ void chain_value_coop(const MixedArgs& a, u32 i, u32 k)
{
    static_assert(use_digits<N>(), "cooperative chains need the digit product");
    Num<N> x = load_num<N>(a.x, a.pitch, i);
    const Num<N> f = load_num<N>(a.f, a.pitch, i);
    const Num<N> g = load_num<N>(a.g, a.pitch, i);
    const u32 len = a.len[i];
    const u64 mask = a.mask[i];
    u32 fd[digit_count<N>()], fs[coop_factor_digits<N, G>()];
    to_digits<N>(f.w, fd);
    coop_shift_factor<N, G>(fd, k, fs);
    for (u32 t = 0; t < len; ++t) x = coop_fma_by_digits<N, G>(x, f, fs, ((mask >> t) & 1u) ? g : zero<N>(ST_OK));
    if (k == 0) store_num<N>(a.out, a.pitch, i, x);
}

} // namespace lf

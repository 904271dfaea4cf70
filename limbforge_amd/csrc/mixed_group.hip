// Persistent precision-group kernels of the mixed-precision chains: compiled once per group of widths
// (LF_GROUP 0..3 covers 2-8, 9-16, 17-24 and 25-32 limbs).  One launch of a group kernel runs every bucket
// of its widths: workgroup b finds its bucket in the table of per-bucket block offsets and runs that
// bucket's width.  The four group kernels of a batch are the kernel nodes of one hipGraph, so a batch of
// up to 31 precisions is four concurrent launches, each sized by its own widths' register needs.
#include "mixed_chain.hpp"

#ifndef LF_GROUP
#error "compile with -DLF_GROUP=<0..3>"
#endif

namespace lf {

namespace {
constexpr int GLO = LF_GROUP == 0 ? 2 : 8 * LF_GROUP + 1;
constexpr int GHI = 8 * LF_GROUP + 8;

template <int N, int L> __device__ __forceinline__ void dispatch(int n, const MixedArgs& a, u32 i, u32 k)
{
    if (n == N) {
        if constexpr (L == 1) chain_value<N>(a, i);
        else chain_value_coop<N, L>(a, i, k);
    } else if constexpr (N < GHI) dispatch<N + 1, L>(n, a, i, k);
}
} // namespace

// table: nb buckets of this group; first[k] = first workgroup of bucket k (first[nb] = grid size); a bucket
// of count values has count * L threads, L consecutive lanes per value (coop_core.hpp when L > 1)
template <int G, int L>
__global__ void __launch_bounds__(256) mixed_group_kernel(const MixedArgs* buckets, const int32_t* words, const uint32_t* first, uint32_t nb)
{
    u32 lo = 0, hi = nb; // the bucket k with first[k] <= blockIdx.x < first[k+1]
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) / 2;
        if (first[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const MixedArgs a = buckets[lo];
    const u64 t = (u64)(blockIdx.x - first[lo]) * blockDim.x + threadIdx.x;
    if (t / L >= a.count) return;
    dispatch<GLO, L>(words[lo], a, (u32)(t / L), (u32)(t % L));
}

template <> const void* mixed_group_entry<LF_GROUP>(int lanes)
{
    if (lanes == 1) return reinterpret_cast<const void*>(&mixed_group_kernel<LF_GROUP, 1>);
#if LF_GROUP >= 2 // at 9-16 limbs the cooperative chain loop spills (2 KB of scratch per lane)
    if (lanes == 2) return reinterpret_cast<const void*>(&mixed_group_kernel<LF_GROUP, 2>);
    if (lanes == 4) return reinterpret_cast<const void*>(&mixed_group_kernel<LF_GROUP, 4>);
#endif
    return nullptr;
}

} // namespace lf

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

template <int N> __device__ __forceinline__ void dispatch(int n, const MixedArgs& a, u32 i)
{
    if (n == N) chain_value<N>(a, i);
    else if constexpr (N < GHI) dispatch<N + 1>(n, a, i);
}
} // namespace

// table: nb buckets of this group; first[k] = first workgroup of bucket k (first[nb] = grid size)
template <int G>
__global__ void __launch_bounds__(256) mixed_group_kernel(const MixedArgs* buckets, const int32_t* words, const uint32_t* first, uint32_t nb)
{
    u32 lo = 0, hi = nb; // the bucket k with first[k] <= blockIdx.x < first[k+1]
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) / 2;
        if (first[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const MixedArgs a = buckets[lo];
    const u32 i = (blockIdx.x - first[lo]) * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    dispatch<GLO>(words[lo], a, i);
}

template <> const void* mixed_group_entry<LF_GROUP>() { return reinterpret_cast<const void*>(&mixed_group_kernel<LF_GROUP>); }

} // namespace lf

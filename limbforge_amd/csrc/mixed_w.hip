// One width of the mixed-precision chain kernel: compiled once per LF_WORDS in [2, 32].
//
// The batching pattern of config "mixed precisions": values of 64-1024 bits with chains of uneven length
// are bucketed by precision on the host (and ordered by chain length inside a bucket, so the lanes of a
// wave run similar trip counts); each bucket is one launch of this kernel for its width, and the buckets
// of a batch become independent kernel nodes of one hipGraph (lf_mixed_graph_create), so a single graph
// launch runs every precision concurrently.
#include "mixed_chain.hpp"

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {

template <int N> __global__ void __launch_bounds__(256) mixed_chain_kernel(MixedArgs a)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    chain_value<N>(a, i);
}

template <int N> hipError_t launch_mixed_chain(const MixedArgs& g, unsigned block, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL((mixed_chain_kernel<N>), dim3(lf::grid_blocks(g.count, block)), dim3(block), 0, s, g);
    return hipGetLastError();
}

template <int N> const void* mixed_chain_entry() { return reinterpret_cast<const void*>(&mixed_chain_kernel<N>); }

template hipError_t launch_mixed_chain<LF_WORDS>(const MixedArgs&, unsigned, hipStream_t);
template const void* mixed_chain_entry<LF_WORDS>();

} // namespace lf

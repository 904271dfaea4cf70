// One width of the mixed-precision chain kernel: compiled once per LF_WORDS in [2, 32].
//
// The batching pattern of config "mixed precisions": values of 64-1024 bits with chains of uneven length
// are bucketed by precision on the host (and ordered by chain length inside a bucket, so the lanes of a
// wave run similar trip counts); each bucket is one launch of this kernel for its width, and the buckets
// of a batch become independent kernel nodes of one hipGraph (lf_mixed_graph_create), so a single graph
// launch runs every precision concurrently.
#include "elementwise_impl.hpp"
#include "lf_internal.hpp"

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {

template <int N> __global__ void __launch_bounds__(256) mixed_chain_kernel(MixedArgs a)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    Num<N> x = load_num<N>(a.x, a.pitch, i);
    const Num<N> f = load_num<N>(a.f, a.pitch, i);
    const Num<N> g = load_num<N>(a.g, a.pitch, i);
    const u32 len = a.len[i];
    const u64 mask = a.mask[i];
    if constexpr (use_digits<N>()) {
        u32 fd[digit_count<N>()]; // the factor's digits, formed once for the chain's multiplications
        to_digits<N>(f.w, fd);
        for (u32 t = 0; t < len; ++t) x = fma_by_digits<N>(x, f, fd, ((mask >> t) & 1u) ? g : zero<N>(ST_OK));
    } else
        for (u32 t = 0; t < len; ++t) x = fma<N>(x, f, ((mask >> t) & 1u) ? g : zero<N>(ST_OK));
    store_num<N>(a.out, a.pitch, i, x);
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

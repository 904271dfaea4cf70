// One width of the element-wise kernels: compiled once per LF_WORDS in [2, 32] so the 31 widths build in
// parallel; capi.hip dispatches to the explicit instantiations below.
#include "elementwise_impl.hpp"

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {
template hipError_t launch_elementwise<LF_WORDS>(int op, const ElemArgs& g, hipStream_t s);
template hipError_t launch_mul_chain<LF_WORDS>(const ChainArgs& g, hipStream_t s);
template hipError_t launch_mul_chain_lanes<LF_WORDS>(const ChainArgs& g, unsigned lanes, hipStream_t s);
} // namespace lf

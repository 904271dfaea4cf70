// One width of the reductions over limb planes: segmented dot products and the adjacent-pair tree sum
// (the reference's segmented_dot_* and tree_sum_* kernels, src/kernels.metal:61-112,197-208), compiled once
// per LF_WORDS in [2, 32].
//
// segmented_dot: out[i] = the fixed sequential chain s <- fma(a[i K + k], b[i K + k] or b[k], s) from s = 0
// (cfma per component for complex values), one rounding per step (docs/numerics.md "Segmented dot
// products").  One lane per segment; the chain is serial by contract, so the parallelism is across segments.
//
// tree_sum: the fixed adjacent-pair tree of rounded additions: at every level pairs (0,1), (2,3), ... are
// added and an unpaired last value is carried unchanged, until one value remains (docs/numerics.md
// "Pairwise reductions").  One pass reduces 2 * 256 consecutive values per workgroup by nine levels of that
// tree: an aligned block of 2^9 values is a subtree of the global tree, and the carry rule is applied with
// the global count of each level, so the block's result is the global tree's node.  Passes repeat over the
// partial sums until one value remains.
#include "elementwise_impl.hpp"
#include "reduce_launch.hpp"

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {

template <int N, bool CPLX> __global__ void __launch_bounds__(256) segmented_dot_kernel(SegDotArgs g)
{
    const u64 i = (u64)blockIdx.x * 256u + threadIdx.x;
    if (i >= g.segments) return;
    const u64 base = i * g.length;
    if constexpr (CPLX) {
        Cpx<N> s{zero<N>(ST_OK), zero<N>(ST_OK)};
        for (u64 k = 0; k < g.length; ++k) {
            const Cpx<N> x = load_cpx<N>(g.a, g.pa, base + k);
            const Cpx<N> y = load_cpx<N>(g.b, g.pb, g.shared_right ? k : base + k);
            s = lf::cfma<N>(x, y, s);
        }
        store_cpx<N>(g.out, g.po, i, s);
    } else {
        Num<N> s = zero<N>(ST_OK);
        for (u64 k = 0; k < g.length; ++k) {
            const Num<N> x = load_num<N>(g.a, g.pa, base + k);
            const Num<N> y = load_num<N>(g.b, g.pb, g.shared_right ? k : base + k);
            s = lf::fma<N>(x, y, s);
        }
        store_num<N>(g.out, g.po, i, s);
    }
}

// the value type of one reduction element
template <int N, bool CPLX> struct TreeValue;
template <int N> struct TreeValue<N, false> {
    Num<N> v;
    __device__ static TreeValue load(const u32* p, u64 pitch, u64 i) { return {load_num<N>(p, pitch, i)}; }
    __device__ void store(u32* p, u64 pitch, u64 i) const { store_num<N>(p, pitch, i, v); }
    __device__ static TreeValue sum(const TreeValue& a, const TreeValue& b) { return {add<N>(a.v, b.v)}; }
};
template <int N> struct TreeValue<N, true> {
    Cpx<N> v;
    __device__ static TreeValue load(const u32* p, u64 pitch, u64 i) { return {load_cpx<N>(p, pitch, i)}; }
    __device__ void store(u32* p, u64 pitch, u64 i) const { store_cpx<N>(p, pitch, i, v); }
    __device__ static TreeValue sum(const TreeValue& a, const TreeValue& b) { return {cadd<N>(a.v, b.v)}; }
};

// Values per level-0 block and tree levels per pass
constexpr int TREE_LANES = 256, TREE_LEVELS = 9;

template <int N, bool CPLX> __global__ void __launch_bounds__(TREE_LANES) tree_sum_kernel(TreeArgs g)
{
    using T = TreeValue<N, CPLX>;
    __shared__ T part[TREE_LANES];
    const u32 t = threadIdx.x;
    const u64 c0 = g.count, c1 = (c0 + 1) / 2;
    const u64 node = (u64)blockIdx.x * TREE_LANES + t; // level-1 node of this lane
    if (node < c1) {
        T x = T::load(g.in, g.pin, 2 * node);
        if (2 * node + 1 < c0) x = T::sum(x, T::load(g.in, g.pin, 2 * node + 1));
        part[t] = x;
    }
    __syncthreads();
    u64 prev = c1; // count of the level below
#pragma unroll 1
    for (int level = 2; level <= TREE_LEVELS; ++level) {
        const u64 cnt = (prev + 1) / 2;
        const u32 stride = 1u << (level - 1);           // lanes between this level's nodes, in level-1 units
        const u64 first = (u64)blockIdx.x * (TREE_LANES >> (level - 1)); // this block's first node at this level
        if ((t & (stride - 1)) == 0) {
            const u64 j = first + (t >> (level - 1));
            if (j < cnt && 2 * j + 1 < prev) part[t] = T::sum(part[t], part[t + stride / 2]);
        }
        __syncthreads();
        prev = cnt;
    }
    if (t == 0 && blockIdx.x < prev) part[0].store(g.out, g.pout, blockIdx.x);
}

template <int N> hipError_t launch_segmented_dot(const SegDotArgs& g, bool complex_values, hipStream_t s)
{
    if (!g.segments) return hipSuccess;
    const dim3 grid(grid_blocks(g.segments, 256)), block(256);
    if (complex_values) hipLaunchKernelGGL((segmented_dot_kernel<N, true>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((segmented_dot_kernel<N, false>), grid, block, 0, s, g);
    return hipGetLastError();
}

template <int N> hipError_t launch_tree_pass(const TreeArgs& g, bool complex_values, hipStream_t s)
{
    const dim3 grid(grid_blocks(g.count, 2 * TREE_LANES)), block(TREE_LANES);
    if (complex_values) hipLaunchKernelGGL((tree_sum_kernel<N, true>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((tree_sum_kernel<N, false>), grid, block, 0, s, g);
    return hipGetLastError();
}

template hipError_t launch_segmented_dot<LF_WORDS>(const SegDotArgs&, bool, hipStream_t);
template hipError_t launch_tree_pass<LF_WORDS>(const TreeArgs&, bool, hipStream_t);

} // namespace lf

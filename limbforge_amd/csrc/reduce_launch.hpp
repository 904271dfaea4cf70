// reduce_launch.hpp — arguments and per-width launchers of the reductions (reduce_w.hip) used by the C ABI
// entries in reduce.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lf {
// out[i] = sequential fma chain over a[i*length + k] * b[i*length + k] (b[k] with shared_right)
struct SegDotArgs {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* out;
    uint64_t pa, pb, po;
    uint64_t segments, length;
    int shared_right;
};
// one tree-sum pass: `count` values of `in` reduced to ceil(count / 512) values of `out`
struct TreeArgs {
    const uint32_t* in;
    uint32_t* out;
    uint64_t pin, pout;
    uint64_t count;
};
template <int N> hipError_t launch_segmented_dot(const SegDotArgs& g, bool complex_values, hipStream_t s);
template <int N> hipError_t launch_tree_pass(const TreeArgs& g, bool complex_values, hipStream_t s);
constexpr uint64_t TREE_SPAN = 512; // values reduced to one by a tree-sum pass
} // namespace lf

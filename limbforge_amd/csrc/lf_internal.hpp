// lf_internal.hpp — declarations shared by the HIP translation units of liblimbforge_hip.so.
//
// Device layout of a resident operand ("limb planes", structure of arrays): for `count` values of N
// words, plane k (k < N) holds limb k of every value, then one plane each of exponents, signs and
// statuses; a complex operand stores the real planes and then the imaginary planes.  Every plane has
// `pitch` 32-bit elements (count rounded up to 64, so each plane starts on a 256-byte boundary) and a
// wave's 64 lanes read one plane word each: every limb load is one coalesced 256-byte transaction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lf {

// Workgroups for `count` items at `per` items each, summed in 64 bits (counts reach 2^32 - 1).
inline unsigned grid_blocks(uint64_t count, uint64_t per) { return unsigned((count + per - 1) / per); }

struct ElemArgs {
    const uint32_t* a;
    const uint32_t* b;
    const uint32_t* c;
    uint32_t* out;
    uint64_t pa, pb, pc, po; // plane pitches in elements
    uint32_t count;
    uint32_t b_stride, b_period, c_stride, c_period;
};

struct ChainArgs {
    const uint32_t* x;
    const uint32_t* f;
    uint32_t* out;
    uint64_t px, pf, po;
    uint32_t count;
    uint32_t f_stride, f_period;
    int steps;
};

// One precision bucket of the mixed-precision chains: value i runs len[i] steps, step t being
// x <- RN(x * f + g) when bit t of mask[i] is set and x <- RN(x * f) otherwise; x, f, g and out are
// limb planes of `count` values with one pitch.
struct MixedArgs {
    const uint32_t* x;
    const uint32_t* f;
    const uint32_t* g;
    uint32_t* out;
    uint64_t pitch;
    const uint32_t* len;
    const uint64_t* mask;
    uint32_t count;
};

// Element-wise operation `op` (numbering of limbforge::Operation) on N-word values.
template <int N> hipError_t launch_elementwise(int op, const ElemArgs& g, hipStream_t s);
// x <- RN(x * f) `steps` times with x held in registers (one launch for the whole chain).
template <int N> hipError_t launch_mul_chain(const ChainArgs& g, hipStream_t s);
// the same chain with `lanes` (1, 2 or 4; 2 and 4 from 8 limbs on) lanes per value, bit for bit
template <int N> hipError_t launch_mul_chain_lanes(const ChainArgs& g, unsigned lanes, hipStream_t s);
// The mixed chains of one bucket with `block` lanes per workgroup (64, 128 or 256).
template <int N> hipError_t launch_mixed_chain(const MixedArgs& g, unsigned block, hipStream_t s);
// The kernel entry of launch_mixed_chain, for graph kernel nodes (arguments: one MixedArgs).
template <int N> const void* mixed_chain_entry();
// The persistent kernel of width group G (0: 2-8 limbs, 1: 9-16, 2: 17-24, 3: 25-32); arguments: a device
// table of MixedArgs, the limb count of each, and the first workgroup of each bucket (plus the grid size).
// the group kernel of group G with `lanes` lanes per value (1; 2 and 4 for groups 2 and 3), or null
template <int G> const void* mixed_group_entry(int lanes);
constexpr int mixed_group_of(int N) { return N <= 8 ? 0 : (N - 1) / 8; }

} // namespace lf

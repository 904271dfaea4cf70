// elementwise_impl.hpp — element-wise and chain kernels over limb planes (one value per lane).
// The reference's `arithmetic`, `complex_arithmetic` and `fused_arithmetic` kernels
// (src/kernels.metal:8-33) with their broadcast read of b and c (kernels.metal:7), rebuilt for wave64:
// operands are loaded plane by plane (coalesced), computed in VGPRs and stored plane by plane.
#pragma once
#include "lf_internal.hpp"
#include "coop_core.hpp"
#include "dot2_core.hpp"
#include "limb_core.hpp"

namespace lf {

constexpr int OP_ADD = 0, OP_SUB = 1, OP_MUL = 2, OP_DIV = 3, OP_CADD = 4, OP_CMUL = 5, OP_CDIV = 6, OP_SQUARE = 7,
              OP_SQRT = 8, OP_FMA = 9, OP_FMS = 10, OP_CFMA = 11, OP_CFMS = 12;

template <int N> __device__ __forceinline__ Num<N> load_num(const u32* base, u64 pitch, u64 i)
{
    Num<N> r;
#pragma unroll
    for (int k = 0; k < N; ++k) r.w[k] = base[k * pitch + i];
    r.e = (int)base[N * pitch + i];
    r.s = (int)base[(N + 1) * pitch + i];
    r.st = base[(N + 2) * pitch + i];
    return r;
}

template <int N> __device__ __forceinline__ void store_num(u32* base, u64 pitch, u64 i, const Num<N>& r)
{
#pragma unroll
    for (int k = 0; k < N; ++k) base[k * pitch + i] = r.w[k];
    base[N * pitch + i] = (u32)r.e;
    base[(N + 1) * pitch + i] = (u32)r.s;
    base[(N + 2) * pitch + i] = r.st;
}

template <int N> __device__ __forceinline__ Cpx<N> load_cpx(const u32* base, u64 pitch, u64 i)
{
    return {load_num<N>(base, pitch, i), load_num<N>(base + (u64)(N + 3) * pitch, pitch, i)};
}

template <int N> __device__ __forceinline__ void store_cpx(u32* base, u64 pitch, u64 i, const Cpx<N>& r)
{
    store_num<N>(base, pitch, i, r.re);
    store_num<N>(base + (u64)(N + 3) * pitch, pitch, i, r.im);
}

// broadcast index (i / stride) % period; stride <= 1 and period 0 leave i unchanged (kernels.metal:7)
__device__ __forceinline__ u32 operand_index(u32 i, u32 stride, u32 period)
{
    u32 j = stride > 1 ? i / stride : i;
    return period ? j % period : j;
}

template <int N, int OP> __global__ void __launch_bounds__(256) elementwise_kernel(ElemArgs g)
{
    const u32 i = blockIdx.x * 256u + threadIdx.x;
    if (i >= g.count) return;
    if constexpr (OP == OP_CADD || OP == OP_CMUL || OP == OP_CDIV || OP == OP_CFMA || OP == OP_CFMS) {
        const u32 jb = operand_index(i, g.b_stride, g.b_period);
        Cpx<N> x = load_cpx<N>(g.a, g.pa, i);
        Cpx<N> y = load_cpx<N>(g.b, g.pb, jb);
        Cpx<N> z;
        if constexpr (OP == OP_CADD) z = cadd<N>(x, y);
        else if constexpr (OP == OP_CMUL) z = cmul<N>(x, y);
        else if constexpr (OP == OP_CDIV) z = cdiv<N>(x, y);
        else {
            const Cpx<N> w = load_cpx<N>(g.c, g.pc, operand_index(i, g.c_stride, g.c_period));
            if constexpr (OP == OP_CFMA) z = lf::cfma<N>(x, y, w);
            else z = lf::cfms<N>(x, y, w);
        }
        store_cpx<N>(g.out, g.po, i, z);
    } else {
        Num<N> x = load_num<N>(g.a, g.pa, i), z;
        if constexpr (OP == OP_SQUARE) z = mul<N>(x, x);
        else if constexpr (OP == OP_SQRT) z = lf::sqrt<N>(x);
        else {
            const u32 jb = operand_index(i, g.b_stride, g.b_period);
            Num<N> y = load_num<N>(g.b, g.pb, jb);
            if constexpr (OP == OP_ADD) z = add<N>(x, y);
            else if constexpr (OP == OP_SUB) z = sub<N>(x, y);
            else if constexpr (OP == OP_MUL) z = mul<N>(x, y);
            else if constexpr (OP == OP_DIV) z = div<N>(x, y);
            else {
                const u32 jc = operand_index(i, g.c_stride, g.c_period);
                Num<N> w = load_num<N>(g.c, g.pc, jc);
                if constexpr (OP == OP_FMA) z = lf::fma<N>(x, y, w);
                else z = fms<N>(x, y, w);
            }
        }
        store_num<N>(g.out, g.po, i, z);
    }
}

// The resident multiplication chain (benchmarks/benchmark.cpp:84-92 encodes it as `steps` dependent
// dispatches with barriers): x stays in VGPRs across all steps, rounded after every multiplication.
template <int N> __global__ void __launch_bounds__(256) mul_chain_kernel(ChainArgs g)
{
    const u32 i = blockIdx.x * 256u + threadIdx.x;
    if (i >= g.count) return;
    Num<N> x = load_num<N>(g.x, g.px, i);
    const Num<N> f = load_num<N>(g.f, g.pf, operand_index(i, g.f_stride, g.f_period));
    if constexpr (use_digits<N>()) {
        u32 fd[digit_count<N>()]; // the factor's digits, formed once for the whole chain
        to_digits<N>(f.w, fd);
        for (int s = 0; s < g.steps; ++s) x = mul_by_digits<N>(x, f, fd);
    } else
        for (int s = 0; s < g.steps; ++s) x = mul<N>(x, f);
    store_num<N>(g.out, g.po, i, x);
}

// The chain with G lanes per value (coop_core.hpp): lanes G i .. G i + G - 1 hold value i, split each
// product's digit columns and round together; lane G i stores the result.  Bit for bit the single-lane
// chain; it shortens each step (about 1/G of the multiply-adds per lane) for batches too small to fill
// the GPU with one lane per value.
template <int N, int G> __global__ void __launch_bounds__(256) mul_chain_coop_kernel(ChainArgs g)
{
    const u64 t = (u64)blockIdx.x * 256u + threadIdx.x;
    const u32 i = (u32)(t / G), k = (u32)(t % G);
    if (i >= g.count) return; // whole groups leave together: G divides the wavefront
    Num<N> x = load_num<N>(g.x, g.px, i);
    const Num<N> f = load_num<N>(g.f, g.pf, operand_index(i, g.f_stride, g.f_period));
    u32 fd[digit_count<N>()], fs[coop_factor_digits<N, G>()];
    to_digits<N>(f.w, fd);
    coop_shift_factor<N, G>(fd, k, fs);
    for (int s = 0; s < g.steps; ++s) x = coop_mul_by_digits<N, G>(x, f, fd, fs);
    if (k == 0) store_num<N>(g.out, g.po, i, x);
}

template <int N, int OP> inline hipError_t launch_op(const ElemArgs& g, hipStream_t s)
{
    const dim3 grid(grid_blocks(g.count, 256)), block(256);
    hipLaunchKernelGGL((elementwise_kernel<N, OP>), grid, block, 0, s, g);
    return hipGetLastError();
}

template <int N> hipError_t launch_elementwise(int op, const ElemArgs& g, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    switch (op) {
    case OP_ADD: return launch_op<N, OP_ADD>(g, s);
    case OP_SUB: return launch_op<N, OP_SUB>(g, s);
    case OP_MUL: return launch_op<N, OP_MUL>(g, s);
    case OP_DIV: return launch_op<N, OP_DIV>(g, s);
    case OP_CADD: return launch_op<N, OP_CADD>(g, s);
    case OP_CMUL: return launch_op<N, OP_CMUL>(g, s);
    case OP_CDIV: return launch_op<N, OP_CDIV>(g, s);
    case OP_SQUARE: return launch_op<N, OP_SQUARE>(g, s);
    case OP_SQRT: return launch_op<N, OP_SQRT>(g, s);
    case OP_FMA: return launch_op<N, OP_FMA>(g, s);
    case OP_FMS: return launch_op<N, OP_FMS>(g, s);
    default: return hipErrorInvalidValue;
    }
}

template <int N> hipError_t launch_mul_chain(const ChainArgs& g, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    const dim3 grid(grid_blocks(g.count, 256)), block(256);
    hipLaunchKernelGGL((mul_chain_kernel<N>), grid, block, 0, s, g);
    return hipGetLastError();
}

// The chain with `lanes` lanes per value: 1, or 2 and 4 from 8 limbs on (the digit products).
template <int N> hipError_t launch_mul_chain_lanes(const ChainArgs& g, unsigned lanes, hipStream_t s)
{
    if (lanes == 1) return launch_mul_chain<N>(g, s);
    if constexpr (use_digits<N>()) {
        if (!g.count) return hipSuccess;
        const dim3 grid(grid_blocks((u64)g.count * lanes, 256)), block(256);
        if (lanes == 2) hipLaunchKernelGGL((mul_chain_coop_kernel<N, 2>), grid, block, 0, s, g);
        else if (lanes == 4) hipLaunchKernelGGL((mul_chain_coop_kernel<N, 4>), grid, block, 0, s, g);
        else return hipErrorInvalidValue;
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

} // namespace lf

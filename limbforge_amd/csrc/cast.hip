// cast.hip — precision casts between limb-plane operands (exact widening, round-to-nearest-even
// narrowing, componentwise for complex): the reference's cast_real / cast_complex
// (src/kernels.metal:251-266; docs/numerics.md "Precision casts").  The kernel is specialised on the
// target width only: the source width is a run-time value because source limbs are read straight from
// their planes (a memory index, not a register index).
#include "elementwise_impl.hpp"
#include "limbforge_hip.h"

namespace lf {

template <int N>
__global__ void __launch_bounds__(256) cast_kernel(const u32* in, u64 pin, int M, u32* out, u64 pout, u32 count, int parts)
{
    const u32 i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    for (int p = 0; p < parts; ++p) {
        const u32* src = in + (u64)p * (M + 3) * pin;
        u32* dst = out + (u64)p * (N + 3) * pout;
        const int e = (int)src[(u64)M * pin + i], s = (int)src[(u64)(M + 1) * pin + i];
        const u32 st = src[(u64)(M + 2) * pin + i];
        Num<N> r;
        if (st || !s) r = zero<N>(st);
        else if (M <= N) { // exact widening: the source limbs become the top M words
#pragma unroll
            for (int k = 0; k < N; ++k) r.w[k] = k >= N - M ? src[(u64)(k - (N - M)) * pin + i] : 0u;
            r.e = e;
            r.s = s;
            r.st = ST_OK;
        } else { // RN narrowing of the normalized source (msb = e)
            u32 m[N];
#pragma unroll
            for (int k = 0; k < N; ++k) m[k] = src[(u64)(M - N + k) * pin + i];
            const u32 g = src[(u64)(M - N - 1) * pin + i];
            bool sticky = (g << 1) != 0;
            for (int k = 0; k < M - N - 1 && !sticky; ++k) sticky = src[(u64)k * pin + i] != 0;
            r = round_words<N>(m, g >> 31, sticky, (i64)e, s, ST_OK);
        }
        store_num<N>(dst, pout, i, r);
    }
}

template <int N> hipError_t launch_cast(const u32* in, u64 pin, int M, u32* out, u64 pout, u32 count, int parts, hipStream_t s)
{
    hipLaunchKernelGGL((cast_kernel<N>), dim3(lf::grid_blocks(count, 256)), dim3(256), 0, s, in, pin, M, out, pout, count, parts);
    return hipGetLastError();
}

typedef hipError_t (*cast_fn)(const u32*, u64, int, u32*, u64, u32, int, hipStream_t);
template <int N> void fill_cast(cast_fn* t)
{
    t[N] = launch_cast<N>;
    if constexpr (N < 32) fill_cast<N + 1>(t);
}

hipError_t dispatch_cast(int to_words, const u32* in, u64 pin, int M, u32* out, u64 pout, u32 count, int parts, hipStream_t s)
{
    static const struct Table {
        cast_fn t[33] = {};
        Table() { fill_cast<2>(t); }
    } table;
    if (!count) return hipSuccess;
    return table.t[to_words](in, pin, M, out, pout, count, parts, s);
}

} // namespace lf

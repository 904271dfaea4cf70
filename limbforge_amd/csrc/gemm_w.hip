// One width of the dense-product kernels: compiled once per LF_WORDS in [2, 32].
#include "gemm_kernels.hpp"
#include "gemm_launch.hpp"

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {

template <int N> hipError_t launch_digits(const DigitsArgs& g, hipStream_t s)
{
    if (!g.K || !g.nlines || !g.L) return hipSuccess;
    if (g.Kp % 16u) return hipErrorInvalidValue; // rows are written in 16-byte chunks
    hipLaunchKernelGGL((digits_kernel<N>), dim3(((g.Kp + 255u) / 256u) * g.nlines), dim3(256), 0, s, g);
    return hipGetLastError();
}

template <int N> hipError_t launch_rebuild(int wx_class, const RebuildArgs& g, hipStream_t s)
{
    if (!g.rows || !g.n) return hipSuccess;
    if (g.L && g.pairs) { // stage 1: the CRT of every band pair
        const dim3 cgrid((unsigned)(((u64)g.rows * g.n + 255u) / 256u), g.pairs), cblock(256);
        if (wx_class == 0)
            hipLaunchKernelGGL((crt_pairs_kernel<rebuild_words(N, 0)>), cgrid, cblock, crt_table_bytes(g.L, rebuild_words(N, 0)), s, g);
        else hipLaunchKernelGGL((crt_pairs_kernel<rebuild_words(N, 1)>), cgrid, cblock, crt_table_bytes(g.L, rebuild_words(N, 1)), s, g);
        if (hipError_t e = hipGetLastError()) return e;
    }
    const dim3 grid(((g.n + REBUILD_BLOCK - 1) / REBUILD_BLOCK) * g.rows), block(REBUILD_BLOCK);
    if (wx_class == 0) hipLaunchKernelGGL((rebuild_kernel<N, rebuild_words(N, 0), rebuild_words(N, 0) + 33>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((rebuild_kernel<N, rebuild_words(N, 1), 1>), grid, block, 0, s, g);
    return hipGetLastError();
}

template <int N> hipError_t launch_dot(int window_class, const DotArgs& g, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    const dim3 grid((g.count + 255u) / 256u), block(256);
    if (window_class == 0) hipLaunchKernelGGL((dot_window_kernel<N, 64>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((dot_window_kernel<N, 128>), grid, block, 0, s, g);
    return hipGetLastError();
}

template <int N> hipError_t launch_cluster_size(const DotArgs& g, uint32_t* words, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    hipLaunchKernelGGL((cluster_size_kernel<N>), dim3((g.count + 255u) / 256u), dim3(256), 0, s, g, words);
    return hipGetLastError();
}

template <int N>
hipError_t launch_cluster_sum(const DotArgs& g, const uint32_t* words, const uint64_t* offsets, uint32_t* scratch, hipStream_t s)
{
    if (!g.count) return hipSuccess;
    hipLaunchKernelGGL((cluster_sum_kernel<N>), dim3((g.count + 255u) / 256u), dim3(256), 0, s, g, words, offsets, scratch);
    return hipGetLastError();
}

template hipError_t launch_digits<LF_WORDS>(const DigitsArgs&, hipStream_t);
template hipError_t launch_rebuild<LF_WORDS>(int, const RebuildArgs&, hipStream_t);
template hipError_t launch_dot<LF_WORDS>(int, const DotArgs&, hipStream_t);
template hipError_t launch_cluster_size<LF_WORDS>(const DotArgs&, uint32_t*, hipStream_t);
template hipError_t launch_cluster_sum<LF_WORDS>(const DotArgs&, const uint32_t*, const uint64_t*, uint32_t*, hipStream_t);

} // namespace lf

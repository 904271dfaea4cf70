// One width of the Cholesky / triangular-solve kernels: compiled once per LF_WORDS in [2, 32].
#include "factor_kernels.hpp"
#include "factor_launch.hpp"

#include <algorithm>

#ifndef LF_WORDS
#error "compile with -DLF_WORDS=<2..32>"
#endif

namespace lf {

template <int N> hipError_t launch_chol_diag(const FactorArgs& f, uint32_t threads, hipStream_t s)
{
    (void)threads;
    if (f.k1 <= f.k0) return hipSuccess;
    hipLaunchKernelGGL((chol_diag_kernel<N>), dim3(1), dim3(diag_waves<N>() * WAVE), 0, s, f);
    return hipGetLastError();
}

template <int N> hipError_t launch_chol_panel(const FactorArgs& f, hipStream_t s)
{
    if (f.k1 >= f.n) return hipSuccess;
    hipLaunchKernelGGL((chol_panel_kernel<N>), dim3(f.n - f.k1), dim3(WAVE), 0, s, f);
    return hipGetLastError();
}

template <int N> hipError_t launch_chol_lookahead(const FactorArgs& f, uint32_t k2, uint32_t waves, hipStream_t s)
{
    if (f.k1 >= f.n || k2 <= f.k1 || !waves) return hipSuccess;
    const uint64_t entries = (uint64_t)(f.n - f.k1) * (k2 - f.k1);
    hipLaunchKernelGGL((chol_lookahead_kernel<N>), dim3((unsigned)std::min<uint64_t>(entries, waves)), dim3(WAVE), 0, s, f, k2);
    return hipGetLastError();
}

template <int N> hipError_t launch_chol_invalid(uint32_t* w, uint64_t pitch, uint32_t n, uint32_t p, hipStream_t s)
{
    const uint64_t m = n - p;
    if (!m) return hipSuccess;
    hipLaunchKernelGGL((chol_invalid_kernel<N>), dim3((unsigned)((m * m + 255u) / 256u)), dim3(256), 0, s, w, pitch, n, p);
    return hipGetLastError();
}

template <int N> hipError_t launch_zero_upper(uint32_t* w, uint64_t pitch, uint32_t n, hipStream_t s)
{
    const uint64_t total = (uint64_t)n * n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL((zero_upper_kernel<N>), dim3((unsigned)((total + 255u) / 256u)), dim3(256), 0, s, w, pitch, n);
    return hipGetLastError();
}

template <int N> hipError_t launch_solve(int backward, const SolveArgs& a, hipStream_t s)
{
    if (!a.nrhs || a.r1 <= a.r0) return hipSuccess;
    const dim3 grid(a.nrhs), block(WAVE);
    // blocks of up to STAGED_ROWS rows whose staged copy fits 64 KB of LDS beside the exact-dot rows run
    // from LDS; larger blocks and wider values read global memory
    const uint32_t B = a.r1 - a.r0;
    const uint64_t dyn = staged_words<N>(B) * 4;
    if (B <= STAGED_ROWS && dyn + sizeof(WaveLds<N>) <= (uint64_t(64) << 10)) {
        if (backward) hipLaunchKernelGGL((solve_backward_staged_kernel<N>), grid, block, (unsigned)dyn, s, a);
        else hipLaunchKernelGGL((solve_forward_staged_kernel<N>), grid, block, (unsigned)dyn, s, a);
        return hipGetLastError();
    }
    if (backward) hipLaunchKernelGGL((solve_backward_kernel<N>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((solve_forward_kernel<N>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <int N> hipError_t launch_solve_update(int backward, const SolveArgs& a, uint32_t waves, hipStream_t s)
{
    const uint64_t rows = backward ? a.r0 : a.n - a.r1;
    if (!a.nrhs || !rows || a.r1 <= a.r0 || !waves) return hipSuccess;
    const uint64_t entries = rows * a.nrhs;
    hipLaunchKernelGGL((solve_update_kernel<N>), dim3((unsigned)std::min<uint64_t>(entries, waves)), dim3(WAVE), 0, s, a,
                       backward ? 1u : 0u);
    return hipGetLastError();
}

template <int N> hipError_t launch_reciprocals(const uint32_t* w, uint64_t pitch, uint32_t n, uint32_t* recip, uint64_t rpitch, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL((reciprocals_kernel<N>), dim3((n + 63) / 64), dim3(64), 0, s, w, pitch, n, recip, rpitch);
    return hipGetLastError();
}

template <int N>
hipError_t launch_factor_steps(const uint32_t* v, uint64_t pitch, uint32_t* scratch, uint64_t swords, int32_t* err, int reps, uint64_t* out,
                               uint32_t* sink, hipStream_t s)
{
    hipLaunchKernelGGL((factor_steps_kernel<N, 0>), dim3(1), dim3(WAVE), 0, s, v, pitch, scratch, swords, err, reps, out, sink);
    hipLaunchKernelGGL((factor_steps_kernel<N, 1>), dim3(1), dim3(WAVE), 0, s, v, pitch, scratch, swords, err, reps, out, sink);
    hipLaunchKernelGGL((factor_steps_kernel<N, 2>), dim3(1), dim3(WAVE), 0, s, v, pitch, scratch, swords, err, reps, out, sink);
    hipLaunchKernelGGL((factor_steps_kernel<N, 3>), dim3(1), dim3(WAVE), 0, s, v, pitch, scratch, swords, err, reps, out, sink);
    hipLaunchKernelGGL((factor_steps_kernel<N, 4>), dim3(1), dim3(WAVE), 0, s, v, pitch, scratch, swords, err, reps, out, sink);
    return hipGetLastError();
}

template hipError_t launch_chol_diag<LF_WORDS>(const FactorArgs&, uint32_t, hipStream_t);
template hipError_t launch_factor_steps<LF_WORDS>(const uint32_t*, uint64_t, uint32_t*, uint64_t, int32_t*, int, uint64_t*, uint32_t*, hipStream_t);
template hipError_t launch_chol_panel<LF_WORDS>(const FactorArgs&, hipStream_t);
template hipError_t launch_chol_lookahead<LF_WORDS>(const FactorArgs&, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_chol_invalid<LF_WORDS>(uint32_t*, uint64_t, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_zero_upper<LF_WORDS>(uint32_t*, uint64_t, uint32_t, hipStream_t);
template hipError_t launch_solve<LF_WORDS>(int, const SolveArgs&, hipStream_t);
template hipError_t launch_solve_update<LF_WORDS>(int, const SolveArgs&, uint32_t, hipStream_t);
template hipError_t launch_reciprocals<LF_WORDS>(const uint32_t*, uint64_t, uint32_t, uint32_t*, uint64_t, hipStream_t);

} // namespace lf

// gemm_launch.hpp — launch entry points of the dense-product kernels (instantiated per width in gemm_w.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "gemm_common.hpp"

namespace lf {

// Words of the CRT integer X' for the two compiled size classes: class 0 holds products of band integers
// up to band_bits = 64 (need <= 64N + 146 bits), class 1 up to band_bits = 512 (need <= 64N + 1042 bits);
// one word of headroom holds X' < L M' before the reduction.
constexpr int rebuild_words(int N, int cls) { return cls == 0 ? 2 * N + 7 : 2 * N + 35; }
// Exact-dot window classes: 0 -> 64 words, 1 -> 128 words; wider outputs use exponent clusters.
constexpr int dot_window_words(int cls) { return cls == 0 ? 64 : 128; }

template <int N> hipError_t launch_digits(const DigitsArgs& g, hipStream_t s);
template <int N> hipError_t launch_rebuild(int wx_class, const RebuildArgs& g, hipStream_t s);
template <int N> hipError_t launch_dot(int window_class, const DotArgs& g, hipStream_t s);
template <int N> hipError_t launch_cluster_size(const DotArgs& g, uint32_t* words, hipStream_t s);
template <int N>
hipError_t launch_cluster_sum(const DotArgs& g, const uint32_t* words, const uint64_t* offsets, uint32_t* scratch, hipStream_t s);
hipError_t launch_product(const ProductArgs& g, hipStream_t s);

} // namespace lf

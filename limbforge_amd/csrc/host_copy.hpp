// host_copy.hpp — copies of pageable host arrays to and from the device for the host-array entries
// (lf_run_host, lf_cast_host, lf_gemm_host / lf_syrk_host, lf_cholesky_host / lf_solve_host,
// lf_segmented_dot_host).
//
// A DMA needs pinned memory, so each chunk of a host array passes through a ring of pinned chunks kept per
// host thread.  The host side of a chunk (pageable <-> pinned) is split over a small process-wide pool of
// copy threads, and the chunk's DMA runs while the next chunk is being copied, so a transfer takes about
// the longer of the DMA time and the parallel host copy time instead of their sum with a one-thread copy.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace lf {

// host -> device on stream s: returns once every byte of `host` has been copied into pinned memory (the
// host array may then be reused; the DMAs may still be running, ordered on s before later work on s).
hipError_t copy_to_device(void* dev, const void* host, std::size_t bytes, hipStream_t s);

// device -> host after the work queued on s: returns once every byte is in `host`.
hipError_t copy_to_host(void* host, const void* dev, std::size_t bytes, hipStream_t s);

// `bytes` from src to dst (host memory) on the copy pool; returns when done.
void parallel_host_copy(void* dst, const void* src, std::size_t bytes);

} // namespace lf

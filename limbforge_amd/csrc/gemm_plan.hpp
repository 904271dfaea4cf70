// gemm_plan.hpp — one exact product on device-resident limb planes, for the host drivers built on the
// update GEMM (the blocked Cholesky and the triangular solves): the planner of gemm.hip behind
// lf_gemm_host / lf_syrk_host.
#pragma once
#include "gemm_common.hpp"
#include "limbforge_hip.h"

#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

namespace lf {

// C (m x n, element (i, j) at plane element i*out_stride + j of C.planes) = or -= left * right^T, where
// left / right are device views of the m rows of op(A) and the n columns of B (the band analysis reads their
// exponent, sign and status planes on the device).  syrk: right = left, outputs i >= j (mirrored unless
// lower_only).
struct Product {
    int bits;
    bool syrk, lower_only, subtract;
    MatView left, right;
    std::size_t m, n, K;
    lf_operand C;
    lf_linalg_options opt;
    hipStream_t stream;
    uint64_t out_stride; // 0: n
};

// Runs the product on P.stream and waits for it; LF_OK or an lf error code (message set).
int run_product(const Product& P, lf_linalg_report* rep, float* device_ms);
int check_linalg_options(const lf_linalg_options* o);

} // namespace lf

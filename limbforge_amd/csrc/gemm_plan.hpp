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

// Host field access for the band analysis: element (line, k) is record line*ls + k*ks; its exponent,
// sign and status words are at the given pointers plus record * stride words.
struct Fields {
    const uint32_t* exponent;
    const uint32_t* sign;
    const uint32_t* status;
    uint64_t stride, ls, ks;
    void get(std::size_t line, std::size_t k, int& e, int& s, uint32_t& st) const
    {
        const uint64_t i = (uint64_t(line) * ls + uint64_t(k) * ks) * stride;
        e = int(exponent[i]);
        s = int(sign[i]);
        st = status[i];
    }
};

// C (m x n, element (i, j) at plane element i*out_stride + j of C.planes) = or -= left * right^T, where
// left / right are device views of the m rows of op(A) and the n columns of B; lf_ / rf_ read the same
// elements on the host for the analysis.  syrk: right = left, outputs i >= j (mirrored unless lower_only).
struct Product {
    int bits;
    bool syrk, lower_only, subtract;
    MatView left, right;
    Fields lf_, rf_;
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

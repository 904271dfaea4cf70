// gemm_common.hpp — data shared by the host planner (gemm.hip) and the dense-product kernels.
//
// Terms (the reference's, src/linalg.mm:70-111 and docs/numerics.md "Dense products"): a *line* is a row
// of the left operand op(A) or a column of the right operand B (for SYRK, a column of A on both sides).
// Each line is split into exponent *bands* [lo, hi] (greedy from the largest exponent, width band_bits);
// an entry (mantissa m, exponent e) of band [lo, hi] is the exact integer X = m * 2^(e-lo).  All bands of
// one side are concatenated into an *extended* operand (one extended row per band member), so one integer
// GEMM per modulus yields every band pair as a sub-block.
#pragma once
#include <stdint.h>

namespace lf {

constexpr int GEMM_MAX_BANDS = 8; // bands per line on the GPU residue path
constexpr int GEMM_MAX_MODULI = 192;

// Line classes (linalg.mm:77): 0 residue GEMM bands, 1 zero or status line (trivial output),
// 2 exact fallback (too many bands, too wide a spread, or every line when forced / K too large).
struct GemmLine {
    int32_t cls;
    int32_t nb;                       // bands (class 0)
    int32_t low;                      // smallest band lo (the line's lowest exponent)
    uint32_t status;                  // OR of the line's entry statuses
    int32_t top, bottom;              // largest / smallest exponent of a nonzero status-free entry
    int32_t member[GEMM_MAX_BANDS];   // extended row (left) / column (right) index of band b
    int32_t lo[GEMM_MAX_BANDS];
    int32_t hi[GEMM_MAX_BANDS];
};

// An operand matrix in limb planes: element (line, k) is plane element line*ls + k*ks.
struct MatView {
    const uint32_t* planes;
    uint64_t pitch;
    uint64_t ls, ks;
};

struct DigitsArgs {
    MatView v;
    const GemmLine* lines;
    uint32_t nlines, K, Kp, rows_p; // rows_p: extended rows (padded) of this side
    uint32_t L;
    const uint32_t* moduli;   // [L]
    const uint32_t* reduce;   // [L][2]: the bits of 1.0f / m (correctly rounded), balanced 2^32 mod m
    const uint32_t* chunk_pow; // [L][64]: 2^(16h) mod m for the 16-bit mantissa chunks
    const uint32_t* shift_pow; // [L][G+1]: 2^s mod m
    uint32_t shift_max;        // G
    int8_t* d1;                // [L][rows_p][Kp]
    int8_t* d0;
};

struct ProductArgs {
    const int8_t* a1;
    const int8_t* a0;  // left digits [L][Mp][Kp]
    const int8_t* b1;
    const int8_t* b0;  // right digits [L][Np][Kp]
    uint16_t* res;     // [L][Mp][Np]
    uint32_t Mp, Np, Kp, L;
    const uint32_t* moduli;
    const int32_t* fold; // per modulus k11, kmid, k00: 65536 y, 256 y, y mod m (balanced), y = (M/m)^-1 mod m
    uint32_t lower;    // SYRK: skip tiles strictly above the diagonal
};

// Output-centric reconstruction: one lane per output C[i][j].
struct RebuildArgs {
    const uint16_t* res;
    uint32_t Mp, Np, L;       // res holds y_q = r_q (M/m_q)^-1 mod m_q (the product kernel's folded residues)
    const uint32_t* crt;      // [L][WX]: (M/m_q) * 2^t, words
    const uint32_t* bound;    // [2][WX]: M' = M * 2^t, then floor(M'/2)
    int32_t t;                // normalising shift of M' (the exact integer is X' / 2^t)
    const GemmLine* left;
    const GemmLine* right;
    uint32_t m, n;            // outputs m x n
    uint32_t* out;            // output planes (element i*out_stride + j)
    uint64_t out_pitch, out_stride;
    uint32_t syrk, lower_only, subtract;
    // two-stage rebuild over output rows [row0, row0 + rows): the CRT of band pair p = b * pj + c of output
    // (i, j) goes to stage word w at (w * pairs + p) * (rows * n) + (i - row0) * n + j (w = WX: the sign)
    uint32_t* stage;
    uint32_t row0, rows, pairs, pj;
};

// Exact dot products of fallback outputs (and the plain limb GEMM): one lane per listed output.
struct DotArgs {
    MatView left, right;      // element (line, k)
    const GemmLine* lrows;
    const GemmLine* rcols;
    const uint32_t* list;     // (i, j) pairs; null = the full m x n grid
    uint32_t count;           // listed outputs (or m*n)
    uint32_t m, n, K, logk;
    uint32_t* out;
    uint64_t out_pitch, out_stride;
    uint32_t syrk, lower_only, subtract, mirror;
    int32_t* error;           // set when an output needs more than the compiled window
};

} // namespace lf

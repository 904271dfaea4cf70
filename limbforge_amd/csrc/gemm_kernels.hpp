// gemm_kernels.hpp — kernels of the exact dense products C = RN(op(A) B) and RN(C - op(A) B).
//
// Residue path (the reference's algorithm, src/linalg.metal:24-152, rebuilt for CDNA4):
//   digits_kernel   each entry's exact band integer X = m * 2^(e-lo) reduced modulo every prime (16-bit
//                   chunk x (2^16h mod m) dot products, then the band shift), balanced residue
//                   v in (-m/2, m/2], split into int8 digits v = 256 d1 + d0 (linalg.metal:24-39)
//   product_kernel  per modulus and 128x128 tile: G11 = D1 D1', G00 = D0 D0', Gmid = D1 D0' + D0 D1' on
//                   v_mfma_i32_32x32x32_i8 with int32 accumulation from LDS-staged tiles, then
//                   y = (65536 G11 + 256 Gmid + G00) (M/m)^-1 mod m written as uint16 (linalg.metal:66-96
//                   writes the residue itself; the CRT factor rides on the digit weights here)
//   crt_pairs_kernel, rebuild_kernel  one lane per output: for each band pair, the exact integer by direct
//                   CRT X' = sum_q y_q (M/m_q) 2^t (independent 64-bit column sums, no carry chains),
//                   reduced mod M' and mapped to the symmetric range,
//                   then one rounding, or exact accumulation over the band pairs then one rounding
//                   (the reference's Garner `reconstruct` + `finish`, linalg.metal:109-152; the same integer)
// Exact dot kernels (plain limb GEMM and the lines the residue path does not take, linalg.hpp:268-318):
//   dot_window_kernel   every exact 2N-word product added into one two's-complement window
//   dot_cluster_kernel  exponent clusters with a sticky sign for outputs whose terms spread further
#pragma once
#include "gemm_common.hpp"
#include "limb_core.hpp"
#include "residue_reduce.hpp"

namespace lf {

// a global table read through the constant address space (uniform loads become scalar loads)
typedef __attribute__((address_space(4))) const u32 CU32;

// ---------------------------------------------------------------- digits
// One workgroup per (line, 256 consecutive k of [0, Kp)): thread t finds the band of entry k0 + t and its
// 16-bit mantissa chunks once, then for 16 moduli at a time forms the residue of the entry's extended
// value (linalg.metal:24-65: chunks times 2^(16h) mod m, times 2^shift mod m, sign, balanced, split into
// d1 * 256 + d0) and leaves the two digit bytes in LDS; the workgroup then writes every band row of the
// line for those 256 columns as 16-byte chunks, zero where the entry is not a member of that band (or is
// zero, carries a status, or lies past K).  Every byte of a line's rows is written, so only the padding
// rows past the last line need clearing on the host side.
constexpr u32 DIG_QB = 16; // moduli per LDS batch

__device__ __forceinline__ u32 byte_eq_mask(u32 x, u32 b4)
{
    // 0xff in each byte of x equal to the byte of b4, 0 elsewhere (the exact zero-byte test)
    const u32 t = x ^ b4;
    const u32 z = ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t | 0x7f7f7f7fu); // 0x80 where t's byte is zero
    return (z >> 7) * 0xffu;
}

template <int N> __global__ void __launch_bounds__(256) digits_kernel(DigitsArgs g)
{
    __shared__ __attribute__((aligned(16))) int8_t stage[DIG_QB][2][256];
    __shared__ __attribute__((aligned(16))) uint8_t bandof[256];
    const u32 kblocks = (g.Kp + 255u) / 256u;
    const u32 line = blockIdx.x / kblocks;
    const u32 k0 = (blockIdx.x - line * kblocks) * 256u;
    if (line >= g.nlines) return;
    const GemmLine* L = g.lines + line;
    if (L->cls != 0) return; // no rows (uniform over the workgroup)
    const u32 t = threadIdx.x, k = k0 + t;
    const int nb = L->nb;
    u32 band = 0xffu;
    int shift = 0, s = 0;
    u32 c[2 * N];
#pragma unroll
    for (int h = 0; h < 2 * N; ++h) c[h] = 0;
    if (k < g.K) {
        const u64 idx = (u64)line * g.v.ls + (u64)k * g.v.ks;
        const u32* P = g.v.planes;
        const u64 pitch = g.v.pitch;
        const u32 st = P[(u64)(N + 2) * pitch + idx];
        s = (int)P[(u64)(N + 1) * pitch + idx];
        if (!st && s) {
            const int e = (int)P[(u64)N * pitch + idx];
            for (int b = 0; b < nb; ++b)
                if (e >= L->lo[b] && e <= L->hi[b]) {
                    band = (u32)b;
                    shift = e - L->lo[b];
                }
            if (band != 0xffu)
#pragma unroll
                for (int w = 0; w < N; ++w) {
                    const u32 x = P[(u64)w * pitch + idx];
                    c[2 * w] = x & 0xffffu;
                    c[2 * w + 1] = x >> 16;
                }
        }
    }
    bandof[t] = (uint8_t)band;
    const u64 plane = (u64)g.rows_p * g.Kp;
    // the modulus tables are read through the constant address space: uniform loads from it are scalar
    // loads (through a generic pointer the compiler must assume the digit stores below may alias them, and
    // loads every table row into every lane)
    const CU32* cpow = (const CU32*)g.chunk_pow;
    const CU32* mods = (const CU32*)g.moduli;
    const CU32* red = (const CU32*)g.reduce;
    for (u32 qb = 0; qb < g.L; qb += DIG_QB) {
        const u32 nq = g.L - qb < DIG_QB ? g.L - qb : DIG_QB;
        // every lane runs the same modulus loop (an entry outside every band has all chunks zero and
        // yields zero digits); the next modulus's table row is loaded while this one is reduced
        u32 cp[2 * N], cpn[2 * N], sp = g.shift_pow[qb * (g.shift_max + 1) + (u32)shift];
#pragma unroll
        for (int h = 0; h < 2 * N; ++h) cp[h] = cpow[qb * 64u + h];
        for (u32 qi = 0; qi < nq; ++qi) {
            const u32 q = qb + qi, qn = qi + 1 < nq ? q + 1 : q;
#pragma unroll
            for (int h = 0; h < 2 * N; ++h) cpn[h] = cpow[qn * 64u + h];
            const u32 spn = g.shift_pow[qn * (g.shift_max + 1) + (u32)shift];
            const int m = (int)mods[q], c32 = (int)red[2 * q + 1];
            const float inv = __uint_as_float(red[2 * q]);
            u64 acc = 0;
#pragma unroll
            for (int h = 0; h < 2 * N; ++h) acc += (u64)c[h] * cp[h];
            // acc < 2N 2^32 and r sp < 2^32: both within the 2^47 of residue_of (residue_reduce.hpp)
            u32 r = residue_of((int64_t)acc, m, inv, c32);
            r = residue_of((int64_t)(u32)__umul24(r, sp), m, inv, c32); // (__umul24 is declared int)
            int v = (int)r;
            if (s < 0) v = v ? m - v : 0;
            if (v > (m >> 1)) v -= m;
            const int d0 = ((v + 128) & 255) - 128;
            const int d1 = (v - d0) / 256;
            stage[qi][0][t] = (int8_t)d1;
            stage[qi][1][t] = (int8_t)d0;
#pragma unroll
            for (int h = 0; h < 2 * N; ++h) cp[h] = cpn[h];
            sp = spn;
        }
        __syncthreads();
        // chunk j: 16 columns (c16) of band row b of plane pl of modulus qb + qi
        const u32 chunks = nq * 2u * (u32)nb * 16u;
        for (u32 j = t; j < chunks; j += 256u) {
            const u32 c16 = j & 15u, rest = j >> 4, b = rest % (u32)nb, r2 = rest / (u32)nb, pl = r2 & 1u, qi = r2 >> 1;
            const u32 kk = k0 + 16u * c16;
            if (kk >= g.Kp) continue;
            const uint4 dv = *reinterpret_cast<const uint4*>(&stage[qi][pl][16u * c16]);
            const uint4 bv = *reinterpret_cast<const uint4*>(&bandof[16u * c16]);
            const u32 b4 = b * 0x01010101u;
            const uint4 out = make_uint4(dv.x & byte_eq_mask(bv.x, b4), dv.y & byte_eq_mask(bv.y, b4), dv.z & byte_eq_mask(bv.z, b4),
                                         dv.w & byte_eq_mask(bv.w, b4));
            int8_t* dst = (pl ? g.d0 : g.d1) + (u64)(qb + qi) * plane + (u64)L->member[b] * g.Kp + kk;
            *reinterpret_cast<uint4*>(dst) = out;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- CRT rebuild
// X' = |sum| * 2^t from the L values y_q = r_q (M/m_q)^-1 mod m_q at (row, col) (the product kernel writes
// the residues already multiplied by that factor); returns the sign (+1/-1) of the symmetric value.  The
// CRT words come from LDS (staged by the kernel, rows padded to whole 16-byte groups, read 16 bytes at a
// time) for wide integers, from scalar loads otherwise.  Measured on the 512-bit 512^3 products (with the
// factor still applied here): 67-word rows 300 us scalar against 173 us staged; 39-word rows 500 us scalar
// against 531 us staged.
constexpr int CRT_LDS_MIN_WX = 48;
template <int WX>
__device__ __forceinline__ int crt_rebuild(const RebuildArgs& g, const u32* tab, u32 row, u32 col, u32 (&X)[WX])
{
    // tab (LDS, when WX >= CRT_LDS_MIN_WX): the L rows of WX CRT words
    constexpr bool STAGED = WX >= CRT_LDS_MIN_WX;
    constexpr int WXP = (WX + 3) & ~3;
    u64 colsum[WX];
#pragma unroll
    for (int w = 0; w < WX; ++w) colsum[w] = 0;
    const u64 plane = (u64)g.Mp * g.Np, at = (u64)row * g.Np + col;
    // residues in batches of 8 moduli, the next batch's loads in flight while this one accumulates
    constexpr u32 RB = 8;
    u32 rr[RB], rn[RB];
#pragma unroll
    for (u32 u = 0; u < RB; ++u) rr[u] = u < g.L ? (u32)g.res[(u64)u * plane + at] : 0u;
    for (u32 q0 = 0; q0 < g.L; q0 += RB) {
#pragma unroll
        for (u32 u = 0; u < RB; ++u) rn[u] = q0 + RB + u < g.L ? (u32)g.res[(u64)(q0 + RB + u) * plane + at] : 0u;
#pragma unroll
        for (u32 u = 0; u < RB; ++u) {
            const u32 q = q0 + u;
            if (q >= g.L) break;
            u32 y = rr[u];
            // y < 2^16 is known to the compiler from the 16-bit load; knowing it, the compiler splits the
            // 64-bit column sums over unaligned register pairs and copies every one around its
            // multiply-add (one extra move per v_mad_u64_u32).  The empty asm hides the range.
            asm("" : "+v"(y));
            if constexpr (STAGED) {
                const uint4* C4 = reinterpret_cast<const uint4*>(tab + q * WXP);
#pragma unroll
                for (int w4 = 0; w4 < WXP / 4; ++w4) {
                    const uint4 c = C4[w4];
                    const u32 cw[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (4 * w4 + k < WX) colsum[4 * w4 + k] += (u64)y * cw[k];
                }
            } else {
                const u32* C = g.crt + q * WX;
#pragma unroll
                for (int w = 0; w < WX; ++w) colsum[w] += (u64)y * C[w];
            }
        }
#pragma unroll
        for (u32 u = 0; u < RB; ++u) rr[u] = rn[u];
    }
    u64 carry = 0;
#pragma unroll
    for (int w = 0; w < WX; ++w) {
        const u64 v = colsum[w] + carry;
        X[w] = (u32)v;
        carry = v >> 32;
    }
    // X' < L M' with M' normalised to bit 32(WX-1)-1: estimate q = floor(X'/M') from the top words
    const u32* Mb = g.bound;
    const u64 xt = ((u64)X[WX - 1] << 32) | X[WX - 2];
    u64 qe = xt / ((u64)Mb[WX - 2] + 1u);
    // X' -= qe * M'
    u64 bw = 0;
#pragma unroll
    for (int w = 0; w < WX; ++w) {
        const u64 p = qe * Mb[w] + bw;
        const u32 lo = (u32)p, old = X[w];
        X[w] = old - lo;
        bw = (p >> 32) + (old < lo ? 1u : 0u);
    }
    for (int rep = 0; rep < 3; ++rep) {
        int cmp = 0;
#pragma unroll
        for (int w = WX - 1; w >= 0; --w)
            if (cmp == 0 && X[w] != Mb[w]) cmp = X[w] > Mb[w] ? 1 : -1;
        if (cmp < 0) break;
        u32 b2 = 0;
#pragma unroll
        for (int w = 0; w < WX; ++w) {
            const u64 v = (u64)X[w] - Mb[w] - b2;
            X[w] = (u32)v;
            b2 = (u32)(v >> 63);
        }
    }
    // symmetric range: X' > floor(M'/2) means the negative value X' - M'
    const u32* H = g.bound + WX;
    int cmp = 0;
#pragma unroll
    for (int w = WX - 1; w >= 0; --w)
        if (cmp == 0 && X[w] != H[w]) cmp = X[w] > H[w] ? 1 : -1;
    if (cmp <= 0) return 1;
    u32 b3 = 0;
#pragma unroll
    for (int w = 0; w < WX; ++w) {
        const u64 v = (u64)Mb[w] - X[w] - b3;
        X[w] = (u32)v;
        b3 = (u32)(v >> 63);
    }
    return -1;
}

template <int N> __device__ __forceinline__ void store_out(u32* base, u64 pitch, u64 stride, u32 i, u32 j, bool mirror, const Num<N>& r)
{
    const u64 at = (u64)i * stride + j;
#pragma unroll
    for (int k = 0; k < N; ++k) base[k * pitch + at] = r.w[k];
    base[N * pitch + at] = (u32)r.e;
    base[(N + 1) * pitch + at] = (u32)r.s;
    base[(N + 2) * pitch + at] = r.st;
    if (mirror && i != j) {
        const u64 bt = (u64)j * stride + i;
#pragma unroll
        for (int k = 0; k < N; ++k) base[k * pitch + bt] = r.w[k];
        base[N * pitch + bt] = (u32)r.e;
        base[(N + 1) * pitch + bt] = (u32)r.s;
        base[(N + 2) * pitch + bt] = r.st;
    }
}

template <int N> __device__ __forceinline__ Num<N> load_out(const u32* base, u64 pitch, u64 at)
{
    Num<N> r;
#pragma unroll
    for (int k = 0; k < N; ++k) r.w[k] = base[k * pitch + at];
    r.e = (int)base[N * pitch + at];
    r.s = (int)base[(N + 1) * pitch + at];
    r.st = base[(N + 2) * pitch + at];
    return r;
}

// Stage 1 of the rebuild: one lane per (band pair, output) of rows [row0, row0 + rows) reconstructs the
// pair's exact integer by CRT, reduced to the symmetric range, into the stage (WX words and a sign word).
// Only this stage touches the residues; the register budget holds the column sums and nothing else.
// Wide CRT integers stage their words in LDS once per workgroup (dynamic shared memory of
// crt_table_bytes(L, WX)).
constexpr u64 crt_table_bytes(u32 L, int WX) { return WX >= CRT_LDS_MIN_WX ? (u64)L * ((WX + 3) & ~3) * 4 : 0; }
template <int WX> __global__ void __launch_bounds__(256) crt_pairs_kernel(RebuildArgs g)
{
    constexpr int WXP = (WX + 3) & ~3;
    extern __shared__ __attribute__((aligned(16))) u32 tab[];
    if constexpr (WX >= CRT_LDS_MIN_WX) {
        for (u32 k = threadIdx.x; k < g.L * WXP; k += 256u) {
            const u32 q = k / WXP, w = k % WXP;
            tab[k] = w < (u32)WX ? g.crt[q * WX + w] : 0u;
        }
        __syncthreads();
    }
    const u64 span = (u64)g.rows * g.n;
    const u64 idx = (u64)blockIdx.x * 256u + threadIdx.x;
    if (idx >= span) return;
    const u32 i = g.row0 + (u32)(idx / g.n), j = (u32)(idx % g.n);
    if (g.syrk && j > i) return;
    const GemmLine* Li = g.left + i;
    const GemmLine* Lj = g.right + j;
    if (Li->cls != 0 || Lj->cls != 0) return;
    const u32 p = blockIdx.y, b = p / g.pj, c = p % g.pj;
    if ((int)b >= Li->nb || (int)c >= Lj->nb) return;
    u32 rr = (u32)Li->member[b], rc = (u32)Lj->member[c];
    if (g.syrk && rc > rr) { // the extended SYRK product is symmetric: read its lower triangle
        const u32 t0 = rr;
        rr = rc;
        rc = t0;
    }
    u32 X[WX];
    const int sg = crt_rebuild<WX>(g, tab, rr, rc, X);
    u32* dst = g.stage + (u64)p * span + idx;
    const u64 wstride = (u64)g.pairs * span;
#pragma unroll
    for (int w = 0; w < WX; ++w) dst[w * wstride] = X[w];
    dst[WX * wstride] = (u32)sg;
}

template <int WX> __device__ __forceinline__ int read_stage(const RebuildArgs& g, u32 b, u32 c, u32 i, u32 j, u32 (&X)[WX])
{
    const u64 span = (u64)g.rows * g.n;
    const u32* src = g.stage + (u64)(b * g.pj + c) * span + (u64)(i - g.row0) * g.n + j;
    const u64 wstride = (u64)g.pairs * span;
#pragma unroll
    for (int w = 0; w < WX; ++w) X[w] = src[w * wstride];
    return (int)src[WX * wstride];
}

// WX: CRT integer words (M' normalised to bit 32(WX-1)-1); ACC: multi-band accumulator words (kept in LDS,
// one row of ACC+1 words per lane; 1 = no multi-band outputs in this instantiation).
constexpr u32 REBUILD_BLOCK = 128;
template <int N, int WX, int ACC> __global__ void __launch_bounds__(REBUILD_BLOCK) rebuild_kernel(RebuildArgs g)
{
    __shared__ u32 rebuild_acc[ACC > 1 ? REBUILD_BLOCK * (ACC + 1) : 1];
    const u32 nblocks = (g.n + REBUILD_BLOCK - 1) / REBUILD_BLOCK;
    const u32 i = g.row0 + blockIdx.x / nblocks;
    const u32 j = (blockIdx.x % nblocks) * REBUILD_BLOCK + threadIdx.x;
    if (i >= g.row0 + g.rows || i >= g.m || j >= g.n) return;
    if (g.syrk && j > i) return; // SYRK computes the lower triangle (mirrored when full)
    const bool mirror = g.syrk && !g.lower_only;
    const GemmLine* Li = g.left + i;
    const GemmLine* Lj = g.right + j;
    const u64 at = (u64)i * g.out_stride + j;
    Num<N> old = zero<N>(ST_OK);
    if (g.subtract) old = load_out<N>(g.out, g.out_pitch, at);
    if (Li->cls == 1 || Lj->cls == 1) {
        // zero or status line: zero with the statuses; a zero line leaves an update unchanged
        u32 st = Li->status | Lj->status | (g.subtract ? old.st : 0u);
        if (!g.subtract || st) store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(st));
        return;
    }
    if (Li->cls != 0 || Lj->cls != 0) return; // exact dot kernels
    if (g.subtract && old.st) {
        store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(old.st));
        return;
    }
    const int nbi = Li->nb, nbj = Lj->nb;
    const i64 base = (i64)Li->low + Lj->low - 2 * (32 * N - 1);
    const i64 old_scale = (i64)old.e - (32 * N - 1);
    Num<N> result;
    if (nbi == 1 && nbj == 1) {
        u32 X[WX];
        const int sg = read_stage<WX>(g, 0, 0, i, j, X);
        if (!g.subtract) result = pack<N, WX>(X, base - g.t, sg, ST_OK);
        else result = round_sum2<N, WX, N>(X, base - g.t, -sg, old.w, old_scale, old.s);
    } else if constexpr (ACC <= WX) {
        result = zero<N>(ST_INVALID); // not reached: the planner sends multi-band lines of this class elsewhere
    } else {
        // exact sum over the band pairs in this lane's LDS accumulator (two's complement, ACC words): each
        // pair's integer is added at its word offset with a runtime address, so no register barrel shifter
        const int tb = (g.t + 31) & ~31;
        u32* acc = rebuild_acc + threadIdx.x * (ACC + 1);
        for (int w = 0; w < ACC; ++w) acc[w] = 0;
        for (int b = 0; b < nbi; ++b)
            for (int c = 0; c < nbj; ++c) {
                u32 X[WX];
                const int sg = read_stage<WX>(g, (u32)b, (u32)c, i, j, X);
                const i64 shift = (i64)(Li->lo[b] - Li->low) + (Lj->lo[c] - Lj->low) + tb - g.t;
                const int o = (int)(shift >> 5);
                const u32 bs = (u32)(shift & 31);
                if (sg > 0) {
                    u32 cy = 0;
#pragma unroll
                    for (int t = 0; t <= WX; ++t) {
                        const u32 hi = t < WX ? X[t] : 0u, lo = t ? X[t - 1] : 0u;
                        const u32 v = bs ? __builtin_amdgcn_alignbit(hi, lo, 32u - bs) : hi;
                        const u64 z = (u64)acc[o + t] + v + cy;
                        acc[o + t] = (u32)z;
                        cy = (u32)(z >> 32);
                    }
                    for (int w = o + WX + 1; cy && w < ACC; ++w) {
                        const u64 z = (u64)acc[w] + cy;
                        acc[w] = (u32)z;
                        cy = (u32)(z >> 32);
                    }
                } else {
                    u32 bw = 0;
#pragma unroll
                    for (int t = 0; t <= WX; ++t) {
                        const u32 hi = t < WX ? X[t] : 0u, lo = t ? X[t - 1] : 0u;
                        const u32 v = bs ? __builtin_amdgcn_alignbit(hi, lo, 32u - bs) : hi;
                        const u64 z = (u64)acc[o + t] - v - bw;
                        acc[o + t] = (u32)z;
                        bw = (u32)(z >> 63);
                    }
                    for (int w = o + WX + 1; bw && w < ACC; ++w) {
                        const u64 z = (u64)acc[w] - bw;
                        acc[w] = (u32)z;
                        bw = (u32)(z >> 63);
                    }
                }
            }
        u32 A[ACC];
#pragma unroll
        for (int w = 0; w < ACC; ++w) A[w] = acc[w];
        int sign = 1;
        if (A[ACC - 1] >> 31) {
            sign = -1;
            u32 cy = 1;
#pragma unroll
            for (int w = 0; w < ACC; ++w) {
                const u64 v = (u64)(u32)~A[w] + cy;
                A[w] = (u32)v;
                cy = (u32)(v >> 32);
            }
        }
        if (!g.subtract) result = pack<N, ACC>(A, base - tb, sign, ST_OK);
        else result = round_sum2<N, ACC, N>(A, base - tb, -sign, old.w, old_scale, old.s);
    }
    store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, result);
}

// ---------------------------------------------------------------- exact dot products
__device__ __forceinline__ void output_of(const DotArgs& g, u32 idx, u32& i, u32& j)
{
    if (g.list) {
        i = g.list[2 * idx];
        j = g.list[2 * idx + 1];
    } else {
        i = idx / g.n;
        j = idx - i * g.n;
    }
}

template <int N> __device__ __forceinline__ Num<N> load_view(const MatView& v, u32 line, u32 k)
{
    return load_out<N>(v.planes, v.pitch, (u64)line * v.ls + (u64)k * v.ks);
}

// The plain exact limb GEMM: every term a_k b_k is formed exactly (2N words) and added into one W-word
// two's-complement window whose lsb is the smallest possible term exponent of the output; one rounding.
template <int N, int W> __global__ void __launch_bounds__(256) dot_window_kernel(DotArgs g)
{
    const u32 idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= g.count) return;
    u32 i, j;
    output_of(g, idx, i, j);
    if (g.syrk && j > i) return;
    const GemmLine* Li = g.lrows + i;
    const GemmLine* Lj = g.rcols + j;
    const u64 at = (u64)i * g.out_stride + j;
    const bool mirror = g.mirror != 0;
    Num<N> old = zero<N>(ST_OK);
    if (g.subtract) old = load_out<N>(g.out, g.out_pitch, at);
    if (Li->cls == 1 || Lj->cls == 1) {
        u32 st = Li->status | Lj->status | (g.subtract ? old.st : 0u);
        if (!g.subtract || st) store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(st));
        return;
    }
    if (g.subtract && old.st) {
        store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(old.st));
        return;
    }
    const i64 lo = (i64)Li->bottom + Lj->bottom;
    u32 acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = 0;
    for (u32 k = 0; k < g.K; ++k) {
        const Num<N> a = load_view<N>(g.left, i, k);
        const Num<N> b = load_view<N>(g.right, j, k);
        if (!a.s || !b.s) continue;
        u32 p[2 * N], w[W];
        mantissa_product<N>(a.w, b.w, p);
        place<W, 2 * N>(w, p, (i64)a.e + b.e - lo);
        accumulate_signed<W>(acc, w, a.s * b.s);
    }
    int sign = 1;
    if (acc[W - 1] >> 31) {
        sign = -1;
        u32 cy = 1;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const u64 v = (u64)(u32)~acc[w] + cy;
            acc[w] = (u32)v;
            cy = (u32)(v >> 32);
        }
    }
    const i64 scale = lo - 2 * (32 * N - 1);
    Num<N> r;
    if (!g.subtract) r = pack<N, W>(acc, scale, sign, ST_OK);
    else r = round_sum2<N, W, N>(acc, scale, -sign, old.w, (i64)old.e - (32 * N - 1), old.s);
    store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, r);
}

// Exact dot product for terms spread beyond any fixed window: the cluster algorithm of
// linalg.hpp:298-317.  Terms in descending exponent order form clusters (a term joins while its top bit
// lies within gap = 32N + ceil(log2(K+1)) + 4 bits of the cluster's lowest bit); the first nonzero cluster
// sum S is exact and only the sign of the next nonzero cluster matters, so the result is
// RN(S * 2^D +/- 1) * 2^-D with D = 32N + 64.  Clusters are found by repeated passes over the term
// exponents (no sort).  cluster_size_kernel measures the widest cluster of every listed output;
// cluster_sum_kernel then sums each cluster exactly in that output's own global scratch (separate
// positive and negative accumulators, so carries stay short) and keeps only the top N+3 words of S and a
// sticky unit for the rounding.  The update's old value is one more term.
struct ClusterTerms {
    const DotArgs& g;
    u32 i, j;
    Num<1> cinfo; // sign and exponent of the old value (limbs not used)
    bool has_c;
    __device__ i64 scale_of(u32 k, int& sg, int N) const
    {
        if (k == g.K) {
            sg = cinfo.s;
            return (i64)cinfo.e - (32 * N - 1);
        }
        const u64 ia = (u64)i * g.left.ls + (u64)k * g.left.ks, ib = (u64)j * g.right.ls + (u64)k * g.right.ks;
        const int sa = (int)g.left.planes[(u64)(N + 1) * g.left.pitch + ia];
        const int sb = (int)g.right.planes[(u64)(N + 1) * g.right.pitch + ib];
        sg = g.subtract ? -sa * sb : sa * sb;
        if (!sg) return -((i64)1 << 62);
        const int ea = (int)g.left.planes[(u64)N * g.left.pitch + ia];
        const int eb = (int)g.right.planes[(u64)N * g.right.pitch + ib];
        return (i64)ea + eb - 2 * (32 * N - 1);
    }
    __device__ u32 terms() const { return g.K + (has_c ? 1u : 0u); }
    // next cluster below `cutoff`: its highest and lowest term exponents; false when none is left
    __device__ bool next(i64 cutoff, i64& hi, i64& low, int N) const
    {
        const i64 width = 64 * N, gap = 32 * N + (i64)g.logk + 4, NONE = -((i64)1 << 62);
        hi = NONE;
        for (u32 k = 0; k < terms(); ++k) {
            int sg;
            const i64 sc = scale_of(k, sg, N);
            if (sg && sc < cutoff && sc > hi) hi = sc;
        }
        if (hi == NONE) return false;
        low = hi;
        for (bool changed = true; changed;) {
            changed = false;
            for (u32 k = 0; k < terms(); ++k) {
                int sg;
                const i64 sc = scale_of(k, sg, N);
                if (sg && sc < low && sc + width - 1 + gap >= low) {
                    low = sc;
                    changed = true;
                }
            }
        }
        return true;
    }
};

// Trivial outputs and updates of status entries; true when the output is finished.
template <int N> __device__ __forceinline__ bool dot_prologue(const DotArgs& g, u32 i, u32 j, Num<N>& old)
{
    const GemmLine* Li = g.lrows + i;
    const GemmLine* Lj = g.rcols + j;
    const bool mirror = g.mirror != 0;
    old = zero<N>(ST_OK);
    if (g.subtract) old = load_out<N>(g.out, g.out_pitch, (u64)i * g.out_stride + j);
    if (Li->cls == 1 || Lj->cls == 1) {
        u32 st = Li->status | Lj->status | (g.subtract ? old.st : 0u);
        if (!g.subtract || st) store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(st));
        return true;
    }
    if (g.subtract && old.st) {
        store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, mirror, zero<N>(old.st));
        return true;
    }
    return false;
}

// words[idx] = accumulator words of the widest cluster of listed output idx (0 if trivial)
template <int N> __global__ void __launch_bounds__(256) cluster_size_kernel(DotArgs g, u32* words)
{
    const u32 idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= g.count) return;
    u32 i, j;
    output_of(g, idx, i, j);
    words[idx] = 0;
    if (g.syrk && j > i) return;
    const GemmLine* Li = g.lrows + i;
    const GemmLine* Lj = g.rcols + j;
    if (Li->cls == 1 || Lj->cls == 1) return;
    Num<1> c = zero<1>(ST_OK);
    bool has_c = false;
    if (g.subtract) {
        const u64 at = (u64)i * g.out_stride + j;
        c.e = (int)g.out[(u64)N * g.out_pitch + at];
        c.s = (int)g.out[(u64)(N + 1) * g.out_pitch + at];
        has_c = c.s != 0;
    }
    const ClusterTerms T{g, i, j, c, has_c};
    i64 cutoff = ((i64)1 << 62), hi, low;
    i64 widest = 0;
    while (T.next(cutoff, hi, low, N)) {
        const i64 w = (hi + 64 * N - low) / 32 + 3;
        widest = w > widest ? w : widest;
        cutoff = low;
    }
    words[idx] = (u32)(widest < 0xffffffffll ? widest : 0xffffffffll);
}

// acc[0..W) += x (S words) * 2^off, off >= 0, in global memory; carries propagate as far as needed
__device__ __forceinline__ void global_add(u32* acc, u64 W, const u32* x, int S, i64 off)
{
    const u64 ws = (u64)(off >> 5);
    const u32 bs = (u32)(off & 31);
    u64 carry = 0;
    for (int t = 0; t <= S; ++t) {
        const u32 hi = t < S ? x[t] : 0u, lo = t ? x[t - 1] : 0u;
        const u32 v = bs ? __builtin_amdgcn_alignbit(hi, lo, 32u - bs) : hi;
        const u64 at = ws + (u64)t;
        if (at >= W) break;
        const u64 s = (u64)acc[at] + v + carry;
        acc[at] = (u32)s;
        carry = s >> 32;
    }
    for (u64 at = ws + (u64)S + 1; carry && at < W; ++at) {
        const u64 s = (u64)acc[at] + carry;
        acc[at] = (u32)s;
        carry = s >> 32;
    }
}

template <int N> __global__ void __launch_bounds__(256) cluster_sum_kernel(DotArgs g, const u32* words, const u64* offsets, u32* scratch)
{
    const u32 idx = blockIdx.x * 256u + threadIdx.x;
    if (idx >= g.count) return;
    u32 i, j;
    output_of(g, idx, i, j);
    if (g.syrk && j > i) return;
    Num<N> old;
    if (dot_prologue<N>(g, i, j, old)) return;
    Num<1> c = zero<1>(ST_OK);
    c.e = old.e;
    c.s = old.s;
    const ClusterTerms T{g, i, j, c, g.subtract && old.s != 0};
    const u64 W = words[idx];
    u32* pos = scratch + offsets[idx];
    u32* neg = pos + W;
    constexpr int KW = N + 4; // kept words of S: a sticky word, then the top N+3 words
    u32 S[KW];
#pragma unroll
    for (int w = 0; w < KW; ++w) S[w] = 0;
    i64 scale1 = 0, cutoff = ((i64)1 << 62), hi, low;
    int nsums = 0, sign1 = 0, sign2 = 0;
    while (nsums < 2 && T.next(cutoff, hi, low, N)) {
        for (u64 w = 0; w < 2 * W; ++w) pos[w] = 0;
        for (u32 k = 0; k < T.terms(); ++k) {
            int sg;
            const i64 sc = T.scale_of(k, sg, N);
            if (!sg || sc < low || sc >= cutoff) continue;
            u32* dst = sg > 0 ? pos : neg;
            if (k == g.K) global_add(dst, W, old.w, N, sc - low);
            else {
                const Num<N> a = load_view<N>(g.left, i, k);
                const Num<N> b = load_view<N>(g.right, j, k);
                u32 p[2 * N];
                mantissa_product<N>(a.w, b.w, p);
                global_add(dst, W, p, 2 * N, sc - low);
            }
        }
        // sign and magnitude of pos - neg
        int cmp = 0;
        for (u64 w = W; w-- > 0;)
            if (pos[w] != neg[w]) {
                cmp = pos[w] > neg[w] ? 1 : -1;
                break;
            }
        if (cmp) {
            if (!nsums) {
                u32* big = cmp > 0 ? pos : neg;
                const u32* small = cmp > 0 ? neg : pos;
                u32 bw = 0;
                u64 top = 0;
                for (u64 w = 0; w < W; ++w) {
                    const u64 v = (u64)big[w] - small[w] - bw;
                    big[w] = (u32)v;
                    bw = (u32)(v >> 63);
                    if (big[w]) top = w;
                }
                // keep words [top-(N+2), top] and a sticky word for everything below
                const i64 first = (i64)top - (N + 2);
                bool sticky = false;
                for (i64 w = 0; w < first; ++w) sticky = sticky || big[w] != 0;
                S[0] = sticky ? 1u : 0u;
#pragma unroll
                for (int t = 0; t < N + 3; ++t) {
                    const i64 w = first + t;
                    S[t + 1] = (w >= 0 && w < (i64)W) ? big[w] : 0u;
                }
                scale1 = low + 32 * (first - 1);
                sign1 = cmp;
            } else
                sign2 = cmp;
            ++nsums;
        }
        cutoff = low;
    }
    Num<N> r;
    if (!nsums) r = zero<N>(ST_OK);
    else if (nsums == 1) r = pack<N, KW>(S, scale1, sign1, ST_OK);
    else {
        // S * 2^D -/+ 1 puts the sticky unit below the rounding position of S
        constexpr int DW = N + 2, EW = KW + DW;
        u32 ext[EW];
#pragma unroll
        for (int w = 0; w < EW; ++w) ext[w] = w >= DW ? S[w - DW] : 0u;
        if (sign2 == sign1) ext[0] = 1;
        else {
            u32 bw = 1;
#pragma unroll
            for (int w = 0; w < EW; ++w) {
                const u64 v = (u64)ext[w] - bw;
                ext[w] = (u32)v;
                bw = (u32)(v >> 63);
            }
        }
        r = pack<N, EW>(ext, scale1 - 32 * DW, sign1, ST_OK);
    }
    store_out<N>(g.out, g.out_pitch, g.out_stride, i, j, g.mirror != 0, r);
}

} // namespace lf

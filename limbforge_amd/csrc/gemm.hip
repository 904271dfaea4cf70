// gemm.hip — exact dense products on MI355X: the int8-MFMA residue product kernel and the host planner
// behind lf_gemm_host / lf_syrk_host / lf_gemm_planes (the reference's Linalg::gemm / syrk and
// Linalg::Impl::run, src/linalg.mm:295-407).
//
// Per call: band analysis of every line on the device (linalg.mm:82-111, same classes and bands; only the
// line table comes back to the host), modulus count from the widest band (linalg.mm:341-345), CRT constants
// in exact integer arithmetic, then on one
// stream: digits of both extended operands, the residue GEMM per modulus on int8 MFMA, one output-centric
// CRT rebuild pass that also writes the trivial outputs, and the exact-dot kernels for the outputs of lines
// outside the residue path.  Nothing is computed on the host besides the plan.
#include "limbforge_hip.h"
#include "gemm_kernels.hpp"
#include "gemm_launch.hpp"
#include "gemm_plan.hpp"
#include "residue_reduce.hpp"
#include "host_copy.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

namespace lf {

// ------------------------------------------------------------------ int8 MFMA residue product
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int PT = 128;       // output tile per workgroup (rows and columns)
constexpr int PK = 32;        // K per MFMA step
constexpr u32 XCDS = 8;       // MI355X: 8 XCDs, workgroups dealt to them round-robin by linear id

// One 32x32 block of residues times the CRT factor y = (M/m)^-1 mod m: (65536 G11 + 256 Gmid + G00) y mod m,
// written as uint16, in 32-bit integer and FP32 arithmetic (residue_fold; the factor rides on the digit
// weights, so the CRT kernel reads y_q itself).
struct FoldWeights {
    int k11, kmid, k00;
};
__device__ __forceinline__ void store_block(uint16_t* out, u32 Np, u32 row, u32 col, u32 lane, const v16i& c11, const v16i& c00,
                                            const v16i& cmid, const FoldWeights& k, int m, float inv, int c32)
{
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        // 32x32 accumulator layout: lane l, element i -> row 8(i/4) + 4(l/32) + i%4, column l%32
        const u32 rr = row + 8u * (u32)(i / 4) + 4u * (lane >> 5) + (u32)(i % 4);
        const u32 cc = col + (lane & 31u);
        out[(u64)rr * Np + cc] = (uint16_t)residue_fold(c11[i], cmid[i], c00[i], k.k11, k.kmid, k.k00, m, inv, c32);
    }
}

// One 128x128 output tile of one modulus: four waves, each a 64x64 sub-tile of 2x2 MFMA blocks with three
// int32 accumulators each (G11, G00, Gmid).  The digit tiles of a 32-deep K step (both planes of 128 A
// rows and 128 B columns, 16 KB) travel from global memory straight into LDS (global_load_lds_dwordx4: lane
// i's 16 bytes land at 16 i of the instruction's LDS base) through a ring of four stages, so three steps
// are in flight while one is computed: each wave issues four loads per step, waits for its own loads of
// the step it is about to use (vmcnt), and one barrier per step both publishes them and frees the oldest
// stage for refilling.  In a stage, row r's two 16-byte halves sit at chunks 2r + (h xor bit 3 of r), so
// the 16 lanes of an MFMA operand read hit distinct banks.  Each wave then reads 8 fragments and issues
// 16 v_mfma_i32_32x32x32_i8.  The 1-D grid is dealt to the XCDs so that all tiles of a modulus run on
// one XCD (modulus q on XCD q mod 8): a modulus's digit planes are fetched into one L2 instead of eight.
constexpr u32 STAGES = 4;
__device__ __forceinline__ u32 chunk_of(u32 r, u32 h) { return 2u * r + (h ^ ((r >> 3) & 1u)); }

template <u32 V> __device__ __forceinline__ void wait_vmcnt()
{
    static_assert(V < 64, "vmcnt field");
    __builtin_amdgcn_s_waitcnt((int)((V & 15u) | (7u << 4) | (15u << 8) | ((V >> 4) << 14)));
}

__global__ void __launch_bounds__(256, 2) product_kernel(ProductArgs g)
{
    const u32 tiles_x = g.Np / PT, tiles = tiles_x * (g.Mp / PT);
    const u32 xcd = blockIdx.x % XCDS, slot = blockIdx.x / XCDS;
    const u32 q = xcd + XCDS * (slot / tiles), tile = slot % tiles;
    if (q >= g.L) return;
    const u32 row0 = (tile / tiles_x) * PT, col0 = (tile % tiles_x) * PT;
    if (g.lower && col0 > row0 + PT - 1) return; // tile strictly above the diagonal
    __shared__ __attribute__((aligned(16))) int8_t ring[STAGES][4][PT * PK]; // [stage][a1, a0, b1, b0][chunks]
    const u64 aplane = (u64)g.Mp * g.Kp, bplane = (u64)g.Np * g.Kp;
    const u32 t = threadIdx.x;
    const u32 lane = t & 63u, wave = t >> 6;
    const u32 wr = (wave >> 1) * 64u, wc = (wave & 1u) * 64u;
    // loads: this wave fills chunks [64 wave, 64 wave + 64) of each of the four regions; lane i takes the
    // chunk at position 64 wave + i, i.e. row p/2, half (p & 1) xor bit 3 of the row
    const u32 p = 64u * wave + lane, pr = p >> 1, ph = (p & 1u) ^ ((pr >> 3) & 1u);
    const int8_t* src[4] = {g.a1 + q * aplane + (u64)(row0 + pr) * g.Kp + 16u * ph, g.a0 + q * aplane + (u64)(row0 + pr) * g.Kp + 16u * ph,
                            g.b1 + q * bplane + (u64)(col0 + pr) * g.Kp + 16u * ph, g.b0 + q * bplane + (u64)(col0 + pr) * g.Kp + 16u * ph};
    auto issue = [&](u32 st) {
        const u32 buf = st % STAGES, k0 = st * PK;
#pragma unroll
        for (int rg = 0; rg < 4; ++rg)
            __builtin_amdgcn_global_load_lds((const void*)(src[rg] + k0), (__attribute__((address_space(3))) void*)&ring[buf][rg][1024u * wave], 16, 0,
                                             0);
    };
    v16i c11[2][2] = {}, c00[2][2] = {}, cmid[2][2] = {};
    const u32 r = lane & 31u, kh = lane >> 5;
    const u32 steps = g.Kp / PK;
#pragma unroll
    for (u32 st = 0; st + 1 < STAGES; ++st)
        if (st < steps) issue(st);
    for (u32 s = 0; s < steps; ++s) {
        const u32 ahead = steps - 1 - s; // stages issued after s, capped at STAGES - 2
        if (ahead >= STAGES - 2) wait_vmcnt<4 * (STAGES - 2)>();
        else if (ahead == 1) wait_vmcnt<4>();
        else wait_vmcnt<0>();
        // a bare barrier: __syncthreads' fence would wait for every load in flight (vmcnt 0); the loads
        // this step reads were waited for above, and the reads of the stage refilled next have completed
        // (their values fed the previous step's MFMAs)
        __builtin_amdgcn_s_barrier();
        if (s + STAGES - 1 < steps) issue(s + STAGES - 1);
        const u32 cur = s % STAGES;
        v4i a1[2], a0[2], b1[2], b0[2];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const u32 ca = 16u * chunk_of(wr + 32u * hh + r, kh), cb = 16u * chunk_of(wc + 32u * hh + r, kh);
            a1[hh] = *(const v4i*)(&ring[cur][0][ca]);
            a0[hh] = *(const v4i*)(&ring[cur][1][ca]);
            b1[hh] = *(const v4i*)(&ring[cur][2][cb]);
            b0[hh] = *(const v4i*)(&ring[cur][3][cb]);
        }
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                c11[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1[mi], b1[ni], c11[mi][ni], 0, 0, 0);
                c00[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0[mi], b0[ni], c00[mi][ni], 0, 0, 0);
                cmid[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1[mi], b0[ni], cmid[mi][ni], 0, 0, 0);
                cmid[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0[mi], b1[ni], cmid[mi][ni], 0, 0, 0);
            }
    }
    const int m = (int)g.moduli[q];
    const float inv = 1.0f / (float)m;
    const int c32u = (int)(((u64)1 << 32) % (u64)m), c32 = c32u > m / 2 ? c32u - m : c32u; // balanced 2^32 mod m
    const FoldWeights k{g.fold[3 * q], g.fold[3 * q + 1], g.fold[3 * q + 2]};
    uint16_t* out = g.res + (u64)q * g.Mp * g.Np;
    // one block at a time (the scheduling barriers keep the accumulators in AGPRs until their turn)
    store_block(out, g.Np, row0 + wr, col0 + wc, lane, c11[0][0], c00[0][0], cmid[0][0], k, m, inv, c32);
    __builtin_amdgcn_sched_barrier(0);
    store_block(out, g.Np, row0 + wr, col0 + wc + 32u, lane, c11[0][1], c00[0][1], cmid[0][1], k, m, inv, c32);
    __builtin_amdgcn_sched_barrier(0);
    store_block(out, g.Np, row0 + wr + 32u, col0 + wc, lane, c11[1][0], c00[1][0], cmid[1][0], k, m, inv, c32);
    __builtin_amdgcn_sched_barrier(0);
    store_block(out, g.Np, row0 + wr + 32u, col0 + wc + 32u, lane, c11[1][1], c00[1][1], cmid[1][1], k, m, inv, c32);
}

hipError_t launch_product(const ProductArgs& g, hipStream_t s)
{
    if (!g.L || !g.Mp || !g.Np) return hipSuccess;
    const u32 tiles = (g.Np / PT) * (g.Mp / PT);
    const u32 per_xcd = (g.L + XCDS - 1) / XCDS; // moduli per XCD (the last ones may idle)
    hipLaunchKernelGGL(product_kernel, dim3(XCDS * tiles * per_xcd), dim3(256), 0, s, g);
    return hipGetLastError();
}

} // namespace lf

// the C ABI's per-thread error message (capi.hip)
void lf_set_last_error(const std::string& message);

namespace {
using lf::GemmLine;
using lf::GEMM_MAX_BANDS;
typedef unsigned __int128 u128;

constexpr std::size_t MAX_K = 65472; // 2 Kp 2^14 < 2^31 keeps every int8 product sum exact in int32

int fail(int code, const std::string& m)
{
    lf_set_last_error(m);
    return code;
}

// ---- exact multiword helpers (host) ----
struct Big {
    std::vector<uint32_t> w; // little endian
    void trim()
    {
        while (!w.empty() && !w.back()) w.pop_back();
    }
    int bits() const
    {
        for (int i = int(w.size()) - 1; i >= 0; --i)
            if (w[i]) return 32 * i + 32 - __builtin_clz(w[i]);
        return 0;
    }
};
Big big_from(uint64_t v)
{
    Big b;
    b.w = {uint32_t(v), uint32_t(v >> 32)};
    b.trim();
    return b;
}
void big_mul_small(Big& b, uint32_t m)
{
    uint64_t c = 0;
    for (auto& x : b.w) {
        uint64_t t = uint64_t(x) * m + c;
        x = uint32_t(t);
        c = t >> 32;
    }
    if (c) b.w.push_back(uint32_t(c));
}
// b / m exactly (m divides b)
Big big_div_small(const Big& b, uint32_t m)
{
    Big q;
    q.w.assign(b.w.size(), 0);
    uint64_t r = 0;
    for (int i = int(b.w.size()) - 1; i >= 0; --i) {
        uint64_t cur = (r << 32) | b.w[i];
        q.w[i] = uint32_t(cur / m);
        r = cur % m;
    }
    q.trim();
    return q;
}
uint32_t big_mod_small(const Big& b, uint32_t m)
{
    uint64_t r = 0;
    for (int i = int(b.w.size()) - 1; i >= 0; --i) r = ((r << 32) | b.w[i]) % m;
    return uint32_t(r);
}
// b << s into exactly `words` words (must fit)
std::vector<uint32_t> big_shifted(const Big& b, int s, int words)
{
    std::vector<uint32_t> out(words, 0);
    const int ws = s / 32, bs = s % 32;
    for (std::size_t i = 0; i < b.w.size(); ++i) {
        uint64_t v = uint64_t(b.w[i]) << bs;
        if (i + ws < std::size_t(words)) out[i + ws] |= uint32_t(v);
        if (i + ws + 1 < std::size_t(words)) out[i + ws + 1] |= uint32_t(v >> 32);
    }
    return out;
}
uint32_t inverse_mod(uint64_t a, uint64_t m) // a^-1 mod m (gcd 1)
{
    int64_t t = 0, nt = 1, r = int64_t(m), nr = int64_t(a % m);
    while (nr) {
        int64_t q = r / nr, x = t - q * nt;
        t = nt;
        nt = x;
        x = r - q * nr;
        r = nr;
        nr = x;
    }
    return uint32_t(t < 0 ? t + int64_t(m) : t);
}

// ------------------------------------------------------------------ band analysis on the device
// The reference's per-line analysis (linalg.mm:82-111), one wavefront per line reading the line's exponent,
// sign and status planes: status OR, largest and smallest exponent of a nonzero status-free entry, the class
// (0 bands, 1 zero or status line, 2 exact fallback) and the greedy bands [hi - G, hi] taken from the largest
// exponent down (lo: the smallest exponent inside), one pass over the line per band.  Lines with more than
// max_bands bands or a spread above max_spread are class 2; the first GEMM_MAX_BANDS bands are stored and nb
// counts them all (up to max_bands + 1).
using lf::GemmLine;
using lf::MatView;
using lf::u32;
using lf::u64;
using lf::GEMM_MAX_BANDS;

struct AnalyzeArgs {
    MatView v;
    uint32_t field_plane; // N: exponent plane (sign N + 1, status N + 2)
    uint32_t nlines, K;
    int32_t G, max_bands, max_spread;
    uint32_t force;
    GemmLine* out;
};

__device__ __forceinline__ int wave_max(int x)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}
__device__ __forceinline__ int wave_min(int x)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) x = min(x, __shfl_xor(x, o));
    return x;
}
__device__ __forceinline__ u32 wave_or(u32 x)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) x |= (u32)__shfl_xor((int)x, o);
    return x;
}

__global__ void __launch_bounds__(256) analyze_kernel(AnalyzeArgs a)
{
    const u32 lane = threadIdx.x & 63u, line = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (line >= a.nlines) return; // whole wavefronts leave together
    const u32* ep = a.v.planes + (u64)a.field_plane * a.v.pitch;
    const u32* sp = ep + a.v.pitch;
    const u32* tp = sp + a.v.pitch;
    const u64 base = (u64)line * a.v.ls;
    u32 status = 0;
    int top = INT_MIN, bottom = INT_MAX;
    for (u32 k = lane; k < a.K; k += 64u) {
        const u64 i = base + (u64)k * a.v.ks;
        const int e = (int)ep[i], sg = (int)sp[i];
        const u32 st = tp[i];
        status |= st;
        if (!st && sg) {
            top = max(top, e);
            bottom = min(bottom, e);
        }
    }
    status = wave_or(status);
    top = wave_max(top);
    bottom = wave_min(bottom);
    GemmLine L;
    L.cls = 1;
    L.nb = 0;
    L.low = 0;
    L.status = status;
    const bool any = top != INT_MIN;
    L.top = any ? top : 0;
    L.bottom = any ? bottom : 0;
#pragma unroll
    for (int b = 0; b < GEMM_MAX_BANDS; ++b) L.member[b] = L.lo[b] = L.hi[b] = 0;
    if (!status && any) {
        if (a.force) L.cls = 2;
        else if (a.max_bands >= 1 && (long)top - bottom <= a.G) {
            L.cls = 0;
            L.nb = 1;
            L.lo[0] = bottom;
            L.hi[0] = top;
            L.low = bottom;
        } else if ((long)top - bottom > a.max_spread) {
            L.cls = 2;
        } else {
            int hi = top, nb = 0;
            bool fallback = false;
            while (true) {
                int lo = hi, next = INT_MIN;
                for (u32 k = lane; k < a.K; k += 64u) {
                    const u64 i = base + (u64)k * a.v.ks;
                    const int e = (int)ep[i], sg = (int)sp[i];
                    const u32 st = tp[i];
                    if (st || !sg || e > hi) continue;
                    if ((long)e >= (long)hi - a.G) lo = min(lo, e);
                    else next = max(next, e);
                }
                lo = wave_min(lo);
                next = wave_max(next);
                if (nb < GEMM_MAX_BANDS) {
                    L.lo[nb] = lo;
                    L.hi[nb] = hi;
                }
                ++nb;
                L.low = lo;
                if (nb > a.max_bands) {
                    fallback = true;
                    break;
                }
                if (next == INT_MIN) break;
                hi = next;
            }
            L.cls = fallback ? 2 : 0;
            L.nb = fallback ? 0 : nb;
        }
    }
    if (lane == 0) a.out[line] = L;
}

// Primes <= 65279, descending: balanced residues then split into two int8 digits (linalg.mm:38-40).
const std::vector<uint32_t>& primes()
{
    static const std::vector<uint32_t> p = [] {
        std::vector<uint32_t> r;
        for (uint32_t q = 65279; r.size() < 1200; --q) {
            bool prime = true;
            for (uint32_t d = 2; d * d <= q; ++d)
                if (q % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) r.push_back(q);
        }
        return r;
    }();
    return p;
}

using lf::Product;

struct Side {
    std::vector<GemmLine> lines;
    std::vector<std::vector<std::pair<long, long>>> bands;
    std::vector<uint64_t> histogram;
    uint64_t fallback = 0;
    std::vector<std::vector<int>> band_members; // band b -> lines having a band b (in line order)
};

// The analysis of one side on the device: one analyze_kernel launch over the side's lines, the line table
// back to the host (the only data the planner reads), then the band histogram and fallback count.
int analyze_side(const MatView& v, int N, std::size_t lines, std::size_t K, const lf_linalg_options& o, bool force, GemmLine* dev,
                 hipStream_t stream, Side& s)
{
    s.lines.assign(lines, GemmLine{});
    s.bands.assign(lines, {});
    if (lines) {
        AnalyzeArgs a{v, uint32_t(N), uint32_t(lines), uint32_t(K), o.band_bits, o.max_bands, o.max_spread, force ? 1u : 0u, dev};
        hipLaunchKernelGGL(analyze_kernel, dim3(unsigned((lines + 3) / 4)), dim3(256), 0, stream, a);
        if (hipError_t e = hipGetLastError()) return fail(LF_RUNTIME_ERROR, std::string("analysis launch: ") + hipGetErrorString(e));
        if (hipError_t e = hipMemcpyAsync(s.lines.data(), dev, sizeof(GemmLine) * lines, hipMemcpyDeviceToHost, stream))
            return fail(LF_RUNTIME_ERROR, std::string("download line table: ") + hipGetErrorString(e));
        if (hipError_t e = hipStreamSynchronize(stream))
            return fail(LF_RUNTIME_ERROR, std::string("device execution (analysis): ") + hipGetErrorString(e));
    }
    s.histogram.assign(std::size_t(std::max(o.max_bands, 0)) + 1, 0);
    for (std::size_t i = 0; i < lines; ++i) {
        GemmLine& L = s.lines[i];
        if (L.cls == 1) {
            ++s.histogram[0];
            continue;
        }
        if (L.cls == 2) {
            ++s.fallback;
            continue;
        }
        ++s.histogram[L.nb];
        for (int b = 0; b < L.nb && b < GEMM_MAX_BANDS; ++b) s.bands[i].push_back({L.lo[b], L.hi[b]});
    }
    return LF_OK;
}

// GPU-path capacity: lines whose bands this build's rebuild kernels cannot hold take the exact-dot path
// (same results; the report counts them as fallback lines).
void route_capacity(Side& s, int wx_class)
{
    for (std::size_t i = 0; i < s.lines.size(); ++i) {
        GemmLine& L = s.lines[i];
        if (L.cls != 0) continue;
        const bool too_many = L.nb > GEMM_MAX_BANDS;
        const bool too_wide = L.nb > 1 && (wx_class != 0 || long(L.hi[0]) - L.low > 512);
        if (too_many || too_wide) {
            L.cls = 2;
            ++s.fallback;
            --s.histogram[L.nb];
            s.bands[i].clear();
        }
    }
}

// Five timing events per host thread and device, created once and reused by every product.
hipError_t stage_events(hipEvent_t** out)
{
    struct Set {
        hipEvent_t e[5] = {};
        ~Set()
        {
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    };
    thread_local Set sets[16];
    int device = 0;
    if (hipError_t err = hipGetDevice(&device)) return err;
    Set& s = sets[device & 15];
    for (auto& x : s.e)
        if (!x)
            if (hipError_t err = hipEventCreate(&x)) return err;
    *out = s.e;
    return hipSuccess;
}

int ceil_log2(std::size_t x)
{
    int l = 0;
    while ((std::size_t(1) << l) < x) ++l;
    return l;
}

struct DeviceBuf {
    void* p = nullptr;
    ~DeviceBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(std::size_t bytes)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc(&p, std::max<std::size_t>(bytes, 16));
    }
};

// Device work buffers of the planner, kept per host thread and device and grown on demand, so that a
// sequence of products (the updates of a factorization) does not allocate and free on every call.  A
// product synchronizes its stream before returning, so the next call on the thread may reuse them.
struct Workspace {
    std::vector<std::pair<void*, std::size_t>> slots;
    ~Workspace()
    {
        for (auto& s : slots)
            if (s.first) (void)hipFree(s.first);
    }
};
struct PoolBuf {
    int slot;
    void* p = nullptr;
    hipError_t alloc(std::size_t bytes)
    {
        thread_local std::vector<std::pair<int, Workspace>> pools; // (device, buffers)
        int dev = 0;
        if (hipError_t e = hipGetDevice(&dev)) return e;
        Workspace* ws = nullptr;
        for (auto& d : pools)
            if (d.first == dev) ws = &d.second;
        if (!ws) {
            pools.emplace_back(dev, Workspace{});
            ws = &pools.back().second;
        }
        if (ws->slots.size() <= std::size_t(slot)) ws->slots.resize(slot + 1, {nullptr, 0});
        auto& s = ws->slots[slot];
        bytes = std::max<std::size_t>(bytes, 16);
        if (s.second < bytes) {
            if (s.first) (void)hipFree(s.first);
            s = {nullptr, 0};
            if (hipError_t e = hipMalloc(&s.first, bytes)) return e;
            s.second = bytes;
        }
        p = s.first;
        return hipSuccess;
    }
};

#define GEMM_HIP(call, where)                                                                                          \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(LF_RUNTIME_ERROR, std::string(where) + ": " + hipGetErrorString(e_));         \
    } while (0)

typedef hipError_t (*digits_fn)(const lf::DigitsArgs&, hipStream_t);
typedef hipError_t (*rebuild_fn)(int, const lf::RebuildArgs&, hipStream_t);
typedef hipError_t (*dot_fn)(int, const lf::DotArgs&, hipStream_t);
typedef hipError_t (*csize_fn)(const lf::DotArgs&, uint32_t*, hipStream_t);
typedef hipError_t (*csum_fn)(const lf::DotArgs&, const uint32_t*, const uint64_t*, uint32_t*, hipStream_t);
struct GemmTables;
template <int N> void fill_gemm(GemmTables& t);
struct GemmTables {
    digits_fn digits[33] = {};
    rebuild_fn rebuild[33] = {};
    dot_fn dot[33] = {};
    csize_fn cluster_size[33] = {};
    csum_fn cluster_sum[33] = {};
    GemmTables() { fill_gemm<2>(*this); }
};
template <int N> void fill_gemm(GemmTables& t)
{
    t.digits[N] = lf::launch_digits<N>;
    t.rebuild[N] = lf::launch_rebuild<N>;
    t.dot[N] = lf::launch_dot<N>;
    t.cluster_size[N] = lf::launch_cluster_size<N>;
    t.cluster_sum[N] = lf::launch_cluster_sum<N>;
    if constexpr (N < 32) fill_gemm<N + 1>(t);
}
const GemmTables& gemm_tables()
{
    static const GemmTables t;
    return t;
}

} // namespace

int lf::run_product(const Product& P, lf_linalg_report* rep, float* device_ms)
{
    using Clock = std::chrono::steady_clock;
    const auto since = [](Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); };
    const int N = P.bits / 32;
    if (rep) {
        std::memset(rep, 0, sizeof *rep);
        rep->m = P.m;
        rep->n = P.n;
        rep->k = P.K;
    }
    auto t0 = Clock::now();
    const bool force = P.opt.force_fallback || P.K > MAX_K || P.opt.method == 1;
    // line tables: written by the analysis on the device, completed with the member lists on the host
    PoolBuf dl{0}, dr{1};
    GEMM_HIP(dl.alloc(sizeof(GemmLine) * std::max<std::size_t>(P.m, 1)), "hipMalloc (lines)");
    if (!P.syrk) GEMM_HIP(dr.alloc(sizeof(GemmLine) * std::max<std::size_t>(P.n, 1)), "hipMalloc (lines)");
    Side ls, rs_own;
    if (int r = analyze_side(P.left, N, P.m, P.K, P.opt, force, (GemmLine*)dl.p, P.stream, ls)) return r;
    if (!P.syrk)
        if (int r = analyze_side(P.right, N, P.n, P.K, P.opt, force, (GemmLine*)dr.p, P.stream, rs_own)) return r;
    Side& rs = P.syrk ? ls : rs_own;
    // modulus count from the widest band of both sides (linalg.mm:341-345)
    int pmax = 0;
    for (const Side* s : {&ls, &rs})
        for (std::size_t i = 0; i < s->lines.size(); ++i)
            if (s->lines[i].cls == 0)
                for (int b = 0; b < s->lines[i].nb && b < GEMM_MAX_BANDS; ++b) pmax = std::max(pmax, s->lines[i].hi[b] - s->lines[i].lo[b]);
    const int logk = ceil_log2(P.K);
    const double need = 2.0 * P.bits + 2 * pmax + logk + 2;
    double have = 0;
    unsigned L = 0;
    while (have <= need) have += std::log2(double(primes()[L++]));
    // the same bound checked exactly: M > 2^need, i.e. M has more than `need` bits (M is odd and `need` an
    // integer); the double sum above decides L as the reference does, this only guards against its rounding
    Big M = big_from(1);
    for (unsigned q = 0; q < L; ++q) big_mul_small(M, primes()[q]);
    while (M.bits() <= int(need)) {
        big_mul_small(M, primes()[L]);
        ++L;
    }
    const int mbits = M.bits();
    int wx_class = mbits + 1 <= 32 * (lf::rebuild_words(N, 0) - 1) ? 0 : mbits + 1 <= 32 * (lf::rebuild_words(N, 1) - 1) ? 1 : 2;
    if (wx_class == 2) { // beyond the compiled CRT classes: every band line takes the exact-dot path
        for (Side* s : {&ls, &rs})
            for (auto& Ln : s->lines)
                if (Ln.cls == 0) {
                    --s->histogram[Ln.nb];
                    Ln.cls = 2;
                    ++s->fallback;
                }
        wx_class = 0;
    }
    route_capacity(ls, wx_class);
    if (!P.syrk) route_capacity(rs, wx_class);
    // extended operands: band-major member lists (linalg.mm:326-330)
    std::size_t Me = 0, Ne = 0;
    auto assign_members = [](Side& s) {
        std::size_t next = 0;
        for (int b = 0; b < GEMM_MAX_BANDS; ++b)
            for (auto& Ln : s.lines)
                if (Ln.cls == 0 && Ln.nb > b) Ln.member[b] = int(next++);
        return next;
    };
    Me = assign_members(ls);
    Ne = P.syrk ? Me : assign_members(rs);
    bool any_pair = false;
    for (auto& a : ls.lines)
        if (a.cls == 0) {
            for (auto& b : rs.lines)
                if (b.cls == 0) {
                    any_pair = true;
                    break;
                }
            break;
        }
    if (rep) {
        rep->band_slots = uint32_t(std::min<std::size_t>(ls.histogram.size(), 65));
        for (std::size_t b = 0; b < ls.histogram.size() && b < 65; ++b) rep->left_bands[b] = ls.histogram[b];
        for (std::size_t b = 0; b < rs.histogram.size() && b < 65; ++b) rep->right_bands[b] = rs.histogram[b];
        rep->left_fallback = ls.fallback;
        rep->right_fallback = rs.fallback;
        rep->analysis_seconds = since(t0);
    }
    t0 = Clock::now();
    // line tables with the member lists
    GEMM_HIP(hipMemcpyAsync(dl.p, ls.lines.data(), sizeof(GemmLine) * P.m, hipMemcpyHostToDevice, P.stream), "upload lines");
    const GemmLine* dright = (const GemmLine*)dl.p;
    if (!P.syrk) {
        GEMM_HIP(hipMemcpyAsync(dr.p, rs.lines.data(), sizeof(GemmLine) * P.n, hipMemcpyHostToDevice, P.stream), "upload lines");
        dright = (const GemmLine*)dr.p;
    }
    // outputs for the exact-dot kernels, by window class
    std::vector<uint32_t> lists[3];
    uint64_t trivial = 0, gpu_out = 0, multi = 0;
    for (std::size_t i = 0; i < P.m; ++i)
        for (std::size_t j = 0; j < (P.syrk ? i + 1 : P.n); ++j) {
            const GemmLine &a = ls.lines[i], &b = rs.lines[j];
            if (a.cls == 1 || b.cls == 1) {
                ++trivial;
                continue;
            }
            if (a.cls == 0 && b.cls == 0) {
                ++gpu_out;
                multi += a.nb > 1 || b.nb > 1;
                continue;
            }
            const long span = (long(a.top) + b.top) - (long(a.bottom) + b.bottom) + 64L * N + logk + 3;
            const int cls = span <= 32L * lf::dot_window_words(0) ? 0 : span <= 32L * lf::dot_window_words(1) ? 1 : 2;
            lists[cls].push_back(uint32_t(i));
            lists[cls].push_back(uint32_t(j));
        }
    const uint64_t fallback_outputs = (lists[0].size() + lists[1].size() + lists[2].size()) / 2;
    // residue path
    const std::size_t Kp = std::max<std::size_t>((std::max<std::size_t>(P.K, 1) + 63) / 64 * 64, 64);
    const std::size_t Mp = (Me + PT - 1) / PT * PT, Np = (Ne + PT - 1) / PT * PT; // product tiles
    PoolBuf dconst{2}, dig_l1{3}, dig_l0{4}, dig_r1{5}, dig_r0{6}, dres{7}, dlist[3] = {{8}, {9}, {10}}, derr{11};
    // events bounding the stages: start, digits done, products done, rebuild done, exact dots done
    hipEvent_t* ev = nullptr;
    GEMM_HIP(stage_events(&ev), "hipEventCreate");
    hipEvent_t e0 = ev[0], e1 = ev[4];
    const int WX = lf::rebuild_words(N, wx_class);
    const int G = P.opt.band_bits;
    // constants: moduli | reduce[L][2] | chunk_pow[L][64] | shift_pow[L][G+1] | fold[L][3] | crt[L][WX] | bound[2][WX]
    std::vector<uint32_t> cst;
    std::size_t o_mod = 0, o_rec = 0, o_chunk = 0, o_shift = 0, o_fold = 0, o_crt = 0, o_bound = 0;
    int tshift = 0;
    if (any_pair) {
        o_mod = cst.size();
        for (unsigned q = 0; q < L; ++q) cst.push_back(primes()[q]);
        o_rec = cst.size();
        for (unsigned q = 0; q < L; ++q) {
            const uint32_t m = primes()[q];
            const float inv = 1.0f / float(m);
            uint32_t bits;
            std::memcpy(&bits, &inv, 4);
            const int64_t c32 = int64_t((uint64_t(1) << 32) % m);
            cst.push_back(bits);
            cst.push_back(uint32_t(int32_t(c32 > int64_t(m / 2) ? c32 - int64_t(m) : c32)));
        }
        o_chunk = cst.size();
        for (unsigned q = 0; q < L; ++q) {
            const uint32_t m = primes()[q];
            uint64_t p = 1;
            for (int h = 0; h < 64; ++h) {
                cst.push_back(uint32_t(p));
                p = (p << 16) % m;
            }
        }
        o_shift = cst.size();
        for (unsigned q = 0; q < L; ++q) {
            const uint32_t m = primes()[q];
            uint64_t p = 1;
            for (int s = 0; s <= G; ++s) {
                cst.push_back(uint32_t(p));
                p = (p << 1) % m;
            }
        }
        o_fold = cst.size();
        tshift = 32 * (WX - 1) - mbits;
        std::vector<std::vector<uint32_t>> crt(L);
        for (unsigned q = 0; q < L; ++q) {
            const uint32_t m = primes()[q];
            Big Mq = big_div_small(M, m);
            // the product kernel's digit weights 65536 y, 256 y, y (mod m, balanced), y = (M/m)^-1 mod m
            const uint64_t y = inverse_mod(big_mod_small(Mq, m), m);
            for (uint64_t w : {uint64_t(65536), uint64_t(256), uint64_t(1)}) {
                const int64_t k = int64_t(w * y % m);
                cst.push_back(uint32_t(int32_t(k > int64_t(m / 2) ? k - int64_t(m) : k)));
            }
            crt[q] = big_shifted(Mq, tshift, WX);
        }
        o_crt = cst.size();
        for (unsigned q = 0; q < L; ++q) cst.insert(cst.end(), crt[q].begin(), crt[q].end());
        o_bound = cst.size();
        std::vector<uint32_t> Ms = big_shifted(M, tshift, WX), H(WX, 0);
        for (int w = 0; w < WX; ++w) H[w] = (Ms[w] >> 1) | (w + 1 < WX ? Ms[w + 1] << 31 : 0u);
        cst.insert(cst.end(), Ms.begin(), Ms.end());
        cst.insert(cst.end(), H.begin(), H.end());
        GEMM_HIP(dconst.alloc(cst.size() * 4), "hipMalloc (constants)");
        GEMM_HIP(hipMemcpyAsync(dconst.p, cst.data(), cst.size() * 4, hipMemcpyHostToDevice, P.stream), "upload constants");
        const std::size_t lplane = Mp * Kp, rplane = Np * Kp;
        GEMM_HIP(dig_l1.alloc(L * lplane), "hipMalloc (digits)");
        GEMM_HIP(dig_l0.alloc(L * lplane), "hipMalloc (digits)");
        if (!P.syrk) {
            GEMM_HIP(dig_r1.alloc(L * rplane), "hipMalloc (digits)");
            GEMM_HIP(dig_r0.alloc(L * rplane), "hipMalloc (digits)");
        }
        GEMM_HIP(dres.alloc(std::size_t(L) * Mp * Np * 2), "hipMalloc (residues)");
    }
    for (int c = 0; c < 3; ++c)
        if (!lists[c].empty()) {
            GEMM_HIP(dlist[c].alloc(lists[c].size() * 4), "hipMalloc (outputs)");
            GEMM_HIP(hipMemcpyAsync(dlist[c].p, lists[c].data(), lists[c].size() * 4, hipMemcpyHostToDevice, P.stream), "upload outputs");
        }
    GEMM_HIP(derr.alloc(4), "hipMalloc");
    GEMM_HIP(hipMemsetAsync(derr.p, 0, 4, P.stream), "hipMemsetAsync");
    if (rep) rep->upload_seconds = since(t0);
    t0 = Clock::now();
    GEMM_HIP(hipEventRecord(e0, P.stream), "hipEventRecord");
    const uint32_t* C32 = (const uint32_t*)dconst.p;
    if (any_pair) {
        // the digits kernel writes every byte of the lines' rows; the padding rows [Me, Mp) of each
        // modulus plane are cleared here
        auto clear_tail = [&](void* base, std::size_t rows, std::size_t rows_p) -> hipError_t {
            if (rows_p == rows) return hipSuccess;
            return hipMemset2DAsync((int8_t*)base + rows * Kp, rows_p * Kp, 0, (rows_p - rows) * Kp, L, P.stream);
        };
        GEMM_HIP(clear_tail(dig_l1.p, Me, Mp), "hipMemset2DAsync");
        GEMM_HIP(clear_tail(dig_l0.p, Me, Mp), "hipMemset2DAsync");
        if (!P.syrk) {
            GEMM_HIP(clear_tail(dig_r1.p, Ne, Np), "hipMemset2DAsync");
            GEMM_HIP(clear_tail(dig_r0.p, Ne, Np), "hipMemset2DAsync");
        }
        lf::DigitsArgs d{};
        d.L = L;
        d.Kp = uint32_t(Kp);
        d.K = uint32_t(P.K);
        d.moduli = C32 + o_mod;
        d.reduce = C32 + o_rec;
        d.chunk_pow = C32 + o_chunk;
        d.shift_pow = C32 + o_shift;
        d.shift_max = uint32_t(G);
        d.v = P.left;
        d.lines = (const GemmLine*)dl.p;
        d.nlines = uint32_t(P.m);
        d.rows_p = uint32_t(Mp);
        d.d1 = (int8_t*)dig_l1.p;
        d.d0 = (int8_t*)dig_l0.p;
        GEMM_HIP(gemm_tables().digits[N](d, P.stream), "digits launch");
        if (!P.syrk) {
            d.v = P.right;
            d.lines = dright;
            d.nlines = uint32_t(P.n);
            d.rows_p = uint32_t(Np);
            d.d1 = (int8_t*)dig_r1.p;
            d.d0 = (int8_t*)dig_r0.p;
            GEMM_HIP(gemm_tables().digits[N](d, P.stream), "digits launch");
        }
        GEMM_HIP(hipEventRecord(ev[1], P.stream), "hipEventRecord");
        lf::ProductArgs pa{};
        pa.a1 = (const int8_t*)dig_l1.p;
        pa.a0 = (const int8_t*)dig_l0.p;
        pa.b1 = (const int8_t*)(P.syrk ? dig_l1.p : dig_r1.p);
        pa.b0 = (const int8_t*)(P.syrk ? dig_l0.p : dig_r0.p);
        pa.res = (uint16_t*)dres.p;
        pa.Mp = uint32_t(Mp);
        pa.Np = uint32_t(Np);
        pa.Kp = uint32_t(Kp);
        pa.L = L;
        pa.moduli = C32 + o_mod;
        pa.fold = (const int32_t*)(C32 + o_fold);
        pa.lower = P.syrk;
        GEMM_HIP(lf::launch_product(pa, P.stream), "product launch");
    } else
        GEMM_HIP(hipEventRecord(ev[1], P.stream), "hipEventRecord");
    GEMM_HIP(hipEventRecord(ev[2], P.stream), "hipEventRecord");
    {
        lf::RebuildArgs r{};
        r.res = (const uint16_t*)dres.p;
        r.Mp = uint32_t(Mp);
        r.Np = uint32_t(Np);
        r.L = any_pair ? L : 0;
        r.crt = C32 + o_crt;
        r.bound = C32 + o_bound;
        r.t = tshift;
        r.left = (const GemmLine*)dl.p;
        r.right = dright;
        r.m = uint32_t(P.m);
        r.n = uint32_t(P.n);
        r.out = (uint32_t*)P.C.planes;
        r.out_pitch = P.C.pitch;
        r.out_stride = P.out_stride ? P.out_stride : P.n;
        r.syrk = P.syrk;
        r.lower_only = P.lower_only;
        r.subtract = P.subtract;
        // band pairs per output and the stage of the pair CRTs, in row chunks of at most 1 GiB
        uint32_t pi = 0, pj = 0;
        for (auto& Ln : ls.lines) pi = std::max<uint32_t>(pi, Ln.cls == 0 ? uint32_t(Ln.nb) : 0u);
        for (auto& Ln : rs.lines) pj = std::max<uint32_t>(pj, Ln.cls == 0 ? uint32_t(Ln.nb) : 0u);
        r.pj = std::max<uint32_t>(pj, 1);
        r.pairs = any_pair ? pi * pj : 0;
        const uint64_t per_row = uint64_t(r.pairs) * P.n * (WX + 1) * 4;
        const uint64_t chunk = per_row ? std::max<uint64_t>(1, (uint64_t(1) << 30) / per_row) : P.m;
        PoolBuf dstage{12};
        if (r.pairs) GEMM_HIP(dstage.alloc(per_row * std::min<uint64_t>(chunk, P.m)), "hipMalloc (CRT stage)");
        r.stage = (uint32_t*)dstage.p;
        for (uint64_t row0 = 0; row0 < P.m; row0 += chunk) {
            r.row0 = uint32_t(row0);
            r.rows = uint32_t(std::min<uint64_t>(chunk, P.m - row0));
            GEMM_HIP(gemm_tables().rebuild[N](wx_class, r, P.stream), "rebuild launch");
        }
    }
    GEMM_HIP(hipEventRecord(ev[3], P.stream), "hipEventRecord");
    for (int c = 0; c < 3; ++c) {
        if (lists[c].empty()) continue;
        lf::DotArgs x{};
        x.left = P.left;
        x.right = P.right;
        x.lrows = (const GemmLine*)dl.p;
        x.rcols = dright;
        x.list = (const uint32_t*)dlist[c].p;
        x.count = uint32_t(lists[c].size() / 2);
        x.m = uint32_t(P.m);
        x.n = uint32_t(P.n);
        x.K = uint32_t(P.K);
        x.logk = uint32_t(ceil_log2(P.K + 1));
        x.out = (uint32_t*)P.C.planes;
        x.out_pitch = P.C.pitch;
        x.out_stride = P.out_stride ? P.out_stride : P.n;
        x.syrk = P.syrk;
        x.lower_only = P.lower_only;
        x.subtract = P.subtract;
        x.mirror = P.syrk && !P.lower_only;
        x.error = (int32_t*)derr.p;
        if (c < 2) {
            GEMM_HIP(gemm_tables().dot[N](c, x, P.stream), "exact dot launch");
            continue;
        }
        // exponent clusters: size every output's widest cluster, then sum in exactly sized scratch
        const std::size_t cnt = x.count;
        DeviceBuf dwords, doffs, dscratch;
        GEMM_HIP(dwords.alloc(cnt * 4), "hipMalloc (cluster sizes)");
        GEMM_HIP(gemm_tables().cluster_size[N](x, (uint32_t*)dwords.p, P.stream), "cluster size launch");
        std::vector<uint32_t> words(cnt);
        GEMM_HIP(hipMemcpyAsync(words.data(), dwords.p, cnt * 4, hipMemcpyDeviceToHost, P.stream), "download cluster sizes");
        GEMM_HIP(hipStreamSynchronize(P.stream), "device execution (cluster sizes)");
        // scratch in chunks of outputs within a budget (two accumulators per output)
        const uint64_t budget = uint64_t(4) << 30; // bytes
        for (std::size_t first = 0; first < cnt;) {
            std::vector<uint64_t> offs;
            uint64_t total = 0;
            std::size_t last = first;
            while (last < cnt && (last == first || (total + 2ull * words[last]) * 4 <= budget)) {
                offs.push_back(total);
                total += 2ull * words[last];
                ++last;
            }
            if (total * 4 > (uint64_t(64) << 30))
                return fail(LF_RUNTIME_ERROR, "exact dot: an exponent cluster needs more than 64 GB of accumulator");
            GEMM_HIP(doffs.alloc(offs.size() * 8), "hipMalloc (cluster offsets)");
            GEMM_HIP(dscratch.alloc(std::max<uint64_t>(total, 1) * 4), "hipMalloc (cluster scratch)");
            GEMM_HIP(hipMemcpyAsync(doffs.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, P.stream), "upload offsets");
            lf::DotArgs y = x;
            y.list = x.list + 2 * first;
            y.count = uint32_t(last - first);
            GEMM_HIP(gemm_tables().cluster_sum[N](y, (const uint32_t*)dwords.p + first, (const uint64_t*)doffs.p, (uint32_t*)dscratch.p, P.stream),
                     "cluster sum launch");
            GEMM_HIP(hipStreamSynchronize(P.stream), "device execution (clusters)");
            first = last;
        }
    }
    GEMM_HIP(hipEventRecord(e1, P.stream), "hipEventRecord");
    int32_t err = 0;
    GEMM_HIP(hipMemcpyAsync(&err, derr.p, 4, hipMemcpyDeviceToHost, P.stream), "download status");
    GEMM_HIP(hipStreamSynchronize(P.stream), "device execution (linalg)");
    float ms = 0, stage_ms[4] = {};
    GEMM_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    for (int q = 0; q < 4; ++q) GEMM_HIP(hipEventElapsedTime(&stage_ms[q], ev[q], ev[q + 1]), "hipEventElapsedTime");
    if (device_ms) *device_ms = ms;
    if (err) return fail(LF_RUNTIME_ERROR, "exact dot: internal window overflow");
    if (rep) {
        rep->gpu_seconds = ms * 1e-3;
        rep->digits_seconds = stage_ms[0] * 1e-3;
        rep->product_seconds = stage_ms[1] * 1e-3;
        rep->rebuild_seconds = stage_ms[2] * 1e-3;
        rep->dot_seconds = stage_ms[3] * 1e-3;
        rep->gemm_rows = Me;
        rep->gemm_cols = Ne;
        rep->moduli = any_pair ? L : 0;
        rep->int8_gemms = any_pair ? 3ull * L : 0;
        rep->gpu_outputs = gpu_out;
        rep->multi_band_outputs = multi;
        rep->fallback_outputs = fallback_outputs;
        rep->trivial_outputs = trivial;
        rep->crt_words = any_pair ? uint32_t(WX) : 0;
        uint32_t np = 0;
        if (any_pair) {
            int lb = 0, rb = 0;
            for (auto& Ln : ls.lines) lb = std::max(lb, Ln.cls == 0 ? Ln.nb : 0);
            for (auto& Ln : rs.lines) rb = std::max(rb, Ln.cls == 0 ? Ln.nb : 0);
            for (int b = 0; b < lb; ++b)
                for (int c = 0; c < (P.syrk ? b + 1 : rb); ++c) {
                    if (np >= 64) break;
                    uint32_t rows = 0, cols = 0;
                    for (auto& Ln : ls.lines) rows += Ln.cls == 0 && Ln.nb > b;
                    for (auto& Ln : rs.lines) cols += Ln.cls == 0 && Ln.nb > c;
                    rep->pairs[np++] = {uint32_t(b), uint32_t(c), rows, cols, L};
                }
        }
        rep->npairs = np;
        rep->assembly_seconds = since(t0);
    }
    return LF_OK;
}

int lf::check_linalg_options(const lf_linalg_options* o)
{
    if (o->band_bits < 0 || o->band_bits > 1024 || o->max_bands < 0 || o->max_spread < 0 || o->max_spread > 8192)
        return fail(LF_INVALID_ARGUMENT, "linalg options: band_bits in [0,1024], max_bands >= 0, max_spread in [0,8192]");
    if (o->method < 0 || o->method > 1) return fail(LF_INVALID_ARGUMENT, "linalg options: method 0 (residue GEMM) or 1 (limb GEMM)");
    return LF_OK;
}

namespace {
using lf::check_linalg_options;

// Staging of the host-array products, kept per host thread and device and grown on demand, so repeated
// calls allocate nothing after the first: a stream and the device AoS / limb-plane buffers (planner pool
// slots); the host arrays pass through host_copy's pinned chunk ring.
struct HostStage {
    int device = -1;
    hipStream_t stream = nullptr;
    ~HostStage()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
    hipError_t ready()
    {
        if (!stream)
            if (hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) return e;
        return hipSuccess;
    }
};
HostStage& host_stage()
{
    thread_local HostStage stages[16];
    int device = 0;
    (void)hipGetDevice(&device);
    return stages[device & 15];
}

} // namespace

extern "C" {

void lf_linalg_default_options(lf_linalg_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->band_bits = 64;
    o->max_bands = 4;
    o->max_spread = 512;
    o->resident_bands = 2;
}

int lf_gemm_host(int bits, int transpose_a, const void* A, const void* B, uint64_t m, uint64_t n, uint64_t k, void* C, int subtract,
                 const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, double* wall_seconds)
{
    const auto start = std::chrono::steady_clock::now();
    if (gpu_seconds) *gpu_seconds = 0;
    if (wall_seconds) *wall_seconds = 0;
    lf_linalg_options o;
    if (options) o = *options;
    else lf_linalg_default_options(&o);
    if (int r = check_linalg_options(&o)) return r;
    if (bits < 64 || bits > 1024 || bits % 32) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    const uint64_t limit = uint64_t(1) << 31;
    if (m >= limit || n >= limit || k >= limit || m * n >= limit || m * k >= limit || k * n >= limit)
        return fail(LF_INVALID_ARGUMENT, "matrix dimensions exceed 32-bit indexing");
    if (m && n && !C) return fail(LF_INVALID_ARGUMENT, "null output");
    if (m * n && k && (!A || !B)) return fail(LF_INVALID_ARGUMENT, "null input");
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->m = m;
        report->n = n;
        report->k = k;
    }
    if (!m || !n) return LF_OK;
    const int N = bits / 32;
    const std::size_t E = bits / 8 + 12;
    HostStage& st = host_stage();
    GEMM_HIP(st.ready(), "hipStreamCreate");
    const std::size_t na = m * k, nb = k * n, nc = m * n;
    // AoS staging of all three operands on the device (the uploads run back to back), then their planes
    PoolBuf aos{13}, pa{14}, pb{15}, pc{16};
    GEMM_HIP(aos.alloc((na + nb + nc) * E), "hipMalloc (staging)");
    GEMM_HIP(pa.alloc(lf_planes_bytes(bits, 0, std::max<uint64_t>(na, 1))), "hipMalloc (A)");
    GEMM_HIP(pb.alloc(lf_planes_bytes(bits, 0, std::max<uint64_t>(nb, 1))), "hipMalloc (B)");
    GEMM_HIP(pc.alloc(lf_planes_bytes(bits, 0, nc)), "hipMalloc (C)");
    char* daos = static_cast<char*>(aos.p);
    lf_operand OA{pa.p, lf_plane_pitch(std::max<uint64_t>(na, 1))}, OB{pb.p, lf_plane_pitch(std::max<uint64_t>(nb, 1))},
        OC{pc.p, lf_plane_pitch(nc)};
    if (na) {
        GEMM_HIP(lf::copy_to_device(daos, A, na * E, st.stream), "upload A");
        if (int r = lf_aos_to_planes(bits, 0, daos, &OA, na, st.stream)) return r;
    }
    if (nb) {
        GEMM_HIP(lf::copy_to_device(daos + na * E, B, nb * E, st.stream), "upload B");
        if (int r = lf_aos_to_planes(bits, 0, daos + na * E, &OB, nb, st.stream)) return r;
    }
    if (subtract) {
        GEMM_HIP(lf::copy_to_device(daos + (na + nb) * E, C, nc * E, st.stream), "upload C");
        if (int r = lf_aos_to_planes(bits, 0, daos + (na + nb) * E, &OC, nc, st.stream)) return r;
    }
    Product P{};
    P.bits = bits;
    P.syrk = false;
    P.lower_only = false;
    P.subtract = subtract != 0;
    // op(A) row i, entry k: A[i*k + kk] or (transpose) A[kk*m + i]; B column j, entry kk: B[kk*n + j]
    P.left = {(const uint32_t*)OA.planes, OA.pitch, transpose_a ? 1 : k, transpose_a ? m : 1};
    P.right = {(const uint32_t*)OB.planes, OB.pitch, 1, n};
    P.m = m;
    P.n = n;
    P.K = k;
    P.C = OC;
    P.opt = o;
    P.stream = st.stream;
    float ms = 0;
    if (int r = run_product(P, report, &ms)) return r;
    if (int r = lf_planes_to_aos(bits, 0, &OC, daos, nc, st.stream)) return r;
    GEMM_HIP(lf::copy_to_host(C, daos, nc * E, st.stream), "download C");
    if (gpu_seconds) *gpu_seconds = ms * 1e-3;
    if (wall_seconds) *wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return LF_OK;
}

// ---- products on device operands (resident products) ----
namespace {
int check_product_planes(const lf_operand* o, uint64_t elements, const char* name)
{
    if (!elements) return LF_OK;
    if (!o || !o->planes) return fail(LF_INVALID_ARGUMENT, std::string("null operand ") + name);
    if (o->pitch < elements) return fail(LF_INVALID_ARGUMENT, std::string("operand ") + name + " pitch is smaller than its element count");
    return LF_OK;
}
bool overlaps(int N, const lf_operand& a, const lf_operand& b)
{
    const uintptr_t a0 = (uintptr_t)a.planes, a1 = a0 + (uintptr_t)(N + 3) * a.pitch * 4;
    const uintptr_t b0 = (uintptr_t)b.planes, b1 = b0 + (uintptr_t)(N + 3) * b.pitch * 4;
    return a0 < b1 && b0 < a1;
}
} // namespace

int lf_gemm(int bits, int transpose_a, const lf_operand* A, const lf_operand* B, uint64_t m, uint64_t n, uint64_t k, const lf_operand* C,
            int subtract, const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, void* stream)
{
    if (gpu_seconds) *gpu_seconds = 0;
    lf_linalg_options o;
    if (options) o = *options;
    else lf_linalg_default_options(&o);
    if (int r = check_linalg_options(&o)) return r;
    if (bits < 64 || bits > 1024 || bits % 32) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    const uint64_t limit = uint64_t(1) << 31;
    if (m >= limit || n >= limit || k >= limit || m * n >= limit || m * k >= limit || k * n >= limit)
        return fail(LF_INVALID_ARGUMENT, "matrix dimensions exceed 32-bit indexing");
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->m = m;
        report->n = n;
        report->k = k;
    }
    if (!m || !n) return LF_OK;
    if (int r = check_product_planes(C, m * n, "C")) return r;
    if (k) {
        if (int r = check_product_planes(A, m * k, "A")) return r;
        if (int r = check_product_planes(B, k * n, "B")) return r;
        if (overlaps(bits / 32, *C, *A) || overlaps(bits / 32, *C, *B)) return fail(LF_INVALID_ARGUMENT, "gemm: C must not overlap A or B");
    }
    Product P{};
    P.bits = bits;
    P.subtract = subtract != 0;
    const uint32_t* a = k ? (const uint32_t*)A->planes : (const uint32_t*)C->planes;
    const uint32_t* b = k ? (const uint32_t*)B->planes : (const uint32_t*)C->planes;
    P.left = {a, k ? A->pitch : C->pitch, transpose_a ? 1 : k, transpose_a ? m : 1};
    P.right = {b, k ? B->pitch : C->pitch, 1, n};
    P.m = m;
    P.n = n;
    P.K = k;
    P.C = *C;
    P.opt = o;
    P.stream = (hipStream_t)stream;
    float ms = 0;
    if (int r = run_product(P, report, &ms)) return r;
    if (gpu_seconds) *gpu_seconds = ms * 1e-3;
    return LF_OK;
}

int lf_syrk(int bits, const lf_operand* A, uint64_t rows, uint64_t cols, const lf_operand* C, int lower_only, int subtract,
            const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, void* stream)
{
    if (gpu_seconds) *gpu_seconds = 0;
    lf_linalg_options o;
    if (options) o = *options;
    else lf_linalg_default_options(&o);
    if (int r = check_linalg_options(&o)) return r;
    if (bits < 64 || bits > 1024 || bits % 32) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (subtract && !lower_only) return fail(LF_INVALID_ARGUMENT, "syrk: subtract updates the lower triangle only (lower_only)");
    const uint64_t limit = uint64_t(1) << 31;
    if (rows >= limit || cols >= limit || cols * cols >= limit || rows * cols >= limit)
        return fail(LF_INVALID_ARGUMENT, "matrix dimensions exceed 32-bit indexing");
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->m = cols;
        report->n = cols;
        report->k = rows;
    }
    if (!cols) return LF_OK;
    if (int r = check_product_planes(C, cols * cols, "C")) return r;
    if (rows) {
        if (int r = check_product_planes(A, rows * cols, "A")) return r;
        if (overlaps(bits / 32, *C, *A)) return fail(LF_INVALID_ARGUMENT, "syrk: C must not overlap A");
    }
    Product P{};
    P.bits = bits;
    P.syrk = true;
    P.lower_only = lower_only != 0;
    P.subtract = subtract != 0;
    P.left = {rows ? (const uint32_t*)A->planes : (const uint32_t*)C->planes, rows ? A->pitch : C->pitch, 1, cols};
    P.right = P.left;
    P.m = cols;
    P.n = cols;
    P.K = rows;
    P.C = *C;
    P.opt = o;
    P.stream = (hipStream_t)stream;
    float ms = 0;
    if (int r = run_product(P, report, &ms)) return r;
    if (gpu_seconds) *gpu_seconds = ms * 1e-3;
    return LF_OK;
}

int lf_syrk_host(int bits, const void* A, uint64_t rows, uint64_t cols, void* C, int lower_only, int subtract,
                 const lf_linalg_options* options, lf_linalg_report* report, double* gpu_seconds, double* wall_seconds)
{
    const auto start = std::chrono::steady_clock::now();
    if (gpu_seconds) *gpu_seconds = 0;
    if (wall_seconds) *wall_seconds = 0;
    lf_linalg_options o;
    if (options) o = *options;
    else lf_linalg_default_options(&o);
    if (int r = check_linalg_options(&o)) return r;
    if (bits < 64 || bits > 1024 || bits % 32) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (subtract && !lower_only) return fail(LF_INVALID_ARGUMENT, "syrk: subtract updates the lower triangle only (lower_only)");
    const uint64_t limit = uint64_t(1) << 31;
    if (rows >= limit || cols >= limit || cols * cols >= limit || rows * cols >= limit)
        return fail(LF_INVALID_ARGUMENT, "matrix dimensions exceed 32-bit indexing");
    if (cols && !C) return fail(LF_INVALID_ARGUMENT, "null output");
    if (cols && rows && !A) return fail(LF_INVALID_ARGUMENT, "null input");
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->m = cols;
        report->n = cols;
        report->k = rows;
    }
    if (!cols) return LF_OK;
    const int N = bits / 32;
    const std::size_t E = bits / 8 + 12;
    HostStage& st = host_stage();
    GEMM_HIP(st.ready(), "hipStreamCreate");
    const std::size_t na = rows * cols, nc = cols * cols;
    PoolBuf aos{13}, pa{14}, pc{16};
    GEMM_HIP(aos.alloc((na + nc) * E), "hipMalloc (staging)");
    GEMM_HIP(pa.alloc(lf_planes_bytes(bits, 0, std::max<uint64_t>(na, 1))), "hipMalloc (A)");
    GEMM_HIP(pc.alloc(lf_planes_bytes(bits, 0, nc)), "hipMalloc (C)");
    char* daos = static_cast<char*>(aos.p);
    lf_operand OA{pa.p, lf_plane_pitch(std::max<uint64_t>(na, 1))}, OC{pc.p, lf_plane_pitch(nc)};
    if (na) {
        GEMM_HIP(lf::copy_to_device(daos, A, na * E, st.stream), "upload A");
        if (int r = lf_aos_to_planes(bits, 0, daos, &OA, na, st.stream)) return r;
    }
    if (subtract || lower_only) { // untouched entries keep their old values
        GEMM_HIP(lf::copy_to_device(daos + na * E, C, nc * E, st.stream), "upload C");
        if (int r = lf_aos_to_planes(bits, 0, daos + na * E, &OC, nc, st.stream)) return r;
    }
    Product P{};
    P.bits = bits;
    P.syrk = true;
    P.lower_only = lower_only != 0;
    P.subtract = subtract != 0;
    // line i = column i of A: element (i, kk) = A[kk*cols + i]
    P.left = {(const uint32_t*)OA.planes, OA.pitch, 1, cols};
    P.right = P.left;
    P.m = cols;
    P.n = cols;
    P.K = rows;
    P.C = OC;
    P.opt = o;
    P.stream = st.stream;
    float ms = 0;
    if (int r = run_product(P, report, &ms)) return r;
    if (int r = lf_planes_to_aos(bits, 0, &OC, daos, nc, st.stream)) return r;
    GEMM_HIP(lf::copy_to_host(C, daos, nc * E, st.stream), "download C");
    if (gpu_seconds) *gpu_seconds = ms * 1e-3;
    if (wall_seconds) *wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return LF_OK;
}

} // extern "C"

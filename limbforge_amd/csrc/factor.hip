// factor.hip — the blocked Cholesky factorization and triangular solves on MI355X behind lf_cholesky /
// lf_solve and their host-array forms (the reference's Linalg::cholesky / trsm / cholesky_solve,
// src/linalg.mm:411-486, docs/numerics.md:674-748).
//
// The work matrix stays in limb planes on the device for the whole call.  Per column block [k0, k1):
//   1. the diagonal block, column by column in one workgroup (pivot, sqrt, then div of the entries below);
//   2. the rows below it, one wavefront per row (chol_panel_kernel);
//   3. the trailing update a_ij = D(a_ij; l_i, l_j; [k0, k1)) for i >= j >= k1, in two parts as the
//      reference splits it (linalg.mm:440-452): the next panel's columns [k1, k2) by the look-ahead kernel
//      on the caller's stream, followed there by the next panel, while the rest (i >= j >= k2) runs as one
//      SYRK product of the update GEMM (band analysis on the device, residue GEMM on int8 MFMA plus
//      exact-dot lines, gemm.hip) on a second stream.
// Every entry is one rounding of the same exact value in either part, so the bits are those of the serial
// order.  The solves run block by block: the in-block recurrence on one wavefront per right-hand side, the
// off-block updates as products of the update GEMM.
#include "factor_launch.hpp"
#include "gemm_plan.hpp"
#include "host_copy.hpp"
#include "limbforge_hip.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

void lf_set_last_error(const std::string& message);

namespace {
using Clock = std::chrono::steady_clock;

int fail(int code, const std::string& m)
{
    lf_set_last_error(m);
    return code;
}

#define FACTOR_HIP(call, where)                                                                                        \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(LF_RUNTIME_ERROR, std::string(where) + ": " + hipGetErrorString(e_));         \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
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

struct OwnedStream {
    hipStream_t s = nullptr;
    ~OwnedStream()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

typedef hipError_t (*diag_fn)(const lf::FactorArgs&, uint32_t, hipStream_t);
typedef hipError_t (*panel_fn)(const lf::FactorArgs&, hipStream_t);
typedef hipError_t (*lookahead_fn)(const lf::FactorArgs&, uint32_t, uint32_t, hipStream_t);
typedef hipError_t (*invalid_fn)(uint32_t*, uint64_t, uint32_t, uint32_t, hipStream_t);
typedef hipError_t (*solve_fn)(int, const lf::SolveArgs&, hipStream_t);
typedef hipError_t (*solve_update_fn)(int, const lf::SolveArgs&, uint32_t, hipStream_t);
typedef hipError_t (*recip_fn)(const uint32_t*, uint64_t, uint32_t, uint32_t*, uint64_t, hipStream_t);
typedef hipError_t (*steps_fn)(const uint32_t*, uint64_t, uint32_t*, uint64_t, int32_t*, int, uint64_t*, uint32_t*, hipStream_t);
struct FactorTables;
template <int N> void fill_factor(FactorTables& t);
struct FactorTables {
    diag_fn diag[33] = {};
    panel_fn panel[33] = {};
    lookahead_fn lookahead[33] = {};
    invalid_fn invalid[33] = {};
    hipError_t (*zero_upper[33])(uint32_t*, uint64_t, uint32_t, hipStream_t) = {};
    solve_fn solve[33] = {};
    solve_update_fn solve_update[33] = {};
    recip_fn reciprocals[33] = {};
    steps_fn steps[33] = {};
    FactorTables() { fill_factor<2>(*this); }
};
template <int N> void fill_factor(FactorTables& t)
{
    t.diag[N] = lf::launch_chol_diag<N>;
    t.panel[N] = lf::launch_chol_panel<N>;
    t.lookahead[N] = lf::launch_chol_lookahead<N>;
    t.invalid[N] = lf::launch_chol_invalid<N>;
    t.zero_upper[N] = lf::launch_zero_upper<N>;
    t.solve[N] = lf::launch_solve<N>;
    t.solve_update[N] = lf::launch_solve_update<N>;
    t.reciprocals[N] = lf::launch_reciprocals<N>;
    t.steps[N] = lf::launch_factor_steps<N>;
    if constexpr (N < 32) fill_factor<N + 1>(t);
}
const FactorTables& tables()
{
    static const FactorTables t;
    return t;
}

bool valid_bits(int bits) { return bits >= 64 && bits <= 1024 && bits % 32 == 0; }

// Copies the N+3 planes of `count` elements between operands of possibly different pitch.
int copy_planes(int N, const lf_operand& from, const lf_operand& to, uint64_t count, hipStream_t s)
{
    if (from.planes == to.planes && from.pitch == to.pitch) return LF_OK;
    if (from.pitch == to.pitch)
        FACTOR_HIP(hipMemcpyAsync(to.planes, from.planes, (uint64_t)(N + 3) * from.pitch * 4, hipMemcpyDeviceToDevice, s), "copy planes");
    else
        FACTOR_HIP(hipMemcpy2DAsync(to.planes, to.pitch * 4, from.planes, from.pitch * 4, count * 4, N + 3, hipMemcpyDeviceToDevice, s),
                   "copy planes");
    return LF_OK;
}

// A matrix operand of `elements` values: planes present and a pitch that holds them.
int check_matrix(const lf_operand* o, uint64_t elements, const char* name)
{
    if (!o || !o->planes) return fail(LF_INVALID_ARGUMENT, std::string("null operand ") + name);
    if (o->pitch < elements) return fail(LF_INVALID_ARGUMENT, std::string("operand ") + name + " pitch is smaller than its element count");
    return LF_OK;
}

// Two operands of N-word values: the same planes at the same pitch (when `same_ok`), or disjoint memory.
int check_overlap(int N, const lf_operand& a, const lf_operand& b, bool same_ok, const char* what)
{
    if (same_ok && a.planes == b.planes && a.pitch == b.pitch) return LF_OK;
    const uintptr_t a0 = (uintptr_t)a.planes, a1 = a0 + (N + 3) * a.pitch * 4;
    const uintptr_t b0 = (uintptr_t)b.planes, b1 = b0 + (N + 3) * b.pitch * 4;
    if (a0 < b1 && b0 < a1)
        return fail(LF_INVALID_ARGUMENT, std::string(what) + (same_ok ? ": operands overlap without being the same planes at the same pitch"
                                                                       : ": the output overlaps an input it does not replace"));
    return LF_OK;
}

int check_factor(int bits, uint64_t n, uint64_t nrhs, const lf_factor_options* fo, lf_factor_options& o, const lf_linalg_options* lo,
                 lf_linalg_options& l)
{
    if (fo) o = *fo;
    else lf_factor_default_options(&o);
    if (lo) l = *lo;
    else lf_linalg_default_options(&l);
    if (int r = lf::check_linalg_options(&l)) return r;
    if (!valid_bits(bits)) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    const uint64_t limit = uint64_t(1) << 31;
    if (n >= limit || n * n >= limit || nrhs >= limit || n * nrhs >= limit)
        return fail(LF_INVALID_ARGUMENT, "matrix dimensions exceed 32-bit indexing");
    return LF_OK;
}

// One update product on the work matrices: C (m x nout at cbase, row stride cstride) -= left * right^T.
int update(int bits, const lf_linalg_options& o, bool syrk, const lf::MatView& left, const lf::MatView& right, std::size_t m,
           std::size_t nout, std::size_t K, uint32_t* cbase, uint64_t cpitch, uint64_t cstride, hipStream_t s, float& ms)
{
    lf::Product P{};
    P.bits = bits;
    P.syrk = syrk;
    P.lower_only = syrk;
    P.subtract = true;
    P.left = left;
    P.right = syrk ? left : right;
    P.m = m;
    P.n = nout;
    P.K = K;
    P.C = {cbase, cpitch};
    P.opt = o;
    P.stream = s;
    P.out_stride = cstride;
    return lf::run_product(P, nullptr, &ms);
}

double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// wavefronts (and scratch slots) of a direct off-block solve update
constexpr uint64_t SOLVE_UPDATE_WAVES = 2048;
// Off-block solve updates of up to this many multiply-add terms (rows x right-hand sides x block) also run
// as direct exact dots: on 2048 wavefronts 2^22 terms of 256 bits take about 50 us, less than one update
// GEMM's planning round trip (band analysis, line-table download, constants, about 0.4 ms of host time per
// update at n = 512 with 64 right-hand sides).  Both forms round the same exact value once.
constexpr double SOLVE_DIRECT_TERMS = 4194304.0;

} // namespace

extern "C" {

void lf_factor_default_options(lf_factor_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->block = 32;
    o->host_macs = 5e4;
    o->solve_host_macs = 4e5;
    o->gpu = 1;
}

int lf_cholesky(int bits, const lf_operand* A, uint64_t n, const lf_operand* L, const lf_factor_options* options,
                const lf_linalg_options* linalg, lf_cholesky_info* info, void* stream)
{
    const auto start = Clock::now();
    lf_factor_options fo;
    lf_linalg_options lo;
    if (int r = check_factor(bits, n, 0, options, fo, linalg, lo)) return r;
    if (n) {
        if (int r = check_matrix(A, n * n, "A")) return r;
        if (int r = check_matrix(L, n * n, "L")) return r;
        if (int r = check_overlap(bits / 32, *A, *L, true, "cholesky")) return r;
    }
    lf_cholesky_info ci;
    std::memset(&ci, 0, sizeof ci);
    ci.pivot = n;
    if (!n) {
        if (info) *info = ci;
        return LF_OK;
    }
    const int N = bits / 32;
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t n32 = uint32_t(n);
    const uint64_t nb = fo.block ? std::min<uint64_t>(fo.block, n) : n;
    uint32_t* W = (uint32_t*)L->planes;
    const uint64_t pitch = L->pitch;
    if (int r = copy_planes(N, *A, *L, n * n, s)) return r;
    FACTOR_HIP(tables().zero_upper[N](W, pitch, n32, s), "zero upper launch");
    // pivot record and per-thread cluster scratch (rows of a panel, or the diagonal block's threads)
    DevBuf dinfo, dscratch;
    const uint64_t swords = lf::factor_scratch_words(N, nb);
    const uint64_t threads = std::max<uint64_t>(n, 256);
    if (2 * swords * threads * 4 > (uint64_t(32) << 30))
        return fail(LF_INVALID_ARGUMENT, "cholesky: the block is too large for the exact-dot scratch (use a smaller block)");
    FACTOR_HIP(dinfo.alloc(sizeof(lf::FactorInfo)), "hipMalloc (pivot)");
    FACTOR_HIP(dscratch.alloc(2 * swords * threads * 4), "hipMalloc (scratch)");
    // reciprocals of the pivots (N+1 planes of n entries), for the divisions of the rows below each block
    DevBuf drecip;
    const uint64_t rpitch = (n + 63) / 64 * 64;
    FACTOR_HIP(drecip.alloc((N + 1) * rpitch * 4), "hipMalloc (reciprocals)");
    // the pivot record comes back through pinned memory, so that reading it does not wait on the stream
    struct PinnedInfo {
        lf::FactorInfo* p = nullptr;
        ~PinnedInfo()
        {
            if (p) (void)hipHostFree(p);
        }
    } hinfo;
    FACTOR_HIP(hipHostMalloc((void**)&hinfo.p, sizeof(lf::FactorInfo), hipHostMallocDefault), "hipHostMalloc (pivot)");
    *hinfo.p = lf::FactorInfo{-1, 0, 0, 0};
    FACTOR_HIP(hipMemcpyAsync(dinfo.p, hinfo.p, sizeof(lf::FactorInfo), hipMemcpyHostToDevice, s), "upload pivot");
    // Look-ahead (the reference overlaps the next panel with the rest of the update on a second host thread,
    // linalg.mm:440-452): after the panel of block [k0, k1), the entries of the next panel's columns
    // [k1, k2) are updated by a look-ahead kernel on stream s, the next panel follows on s, and the rest of
    // the trailing update (rows and columns from k2) runs on a second stream at the same time.  Every entry
    // still gets one rounding of the same exact value per block, so the bits are those of the serial order.
    OwnedStream upd;
    FACTOR_HIP(hipStreamCreateWithFlags(&upd.s, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t ev[6];
    for (auto& e : ev) e = nullptr;
    struct EventsGuard {
        hipEvent_t* e;
        ~EventsGuard()
        {
            for (int i = 0; i < 6; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    for (auto& e : ev) FACTOR_HIP(hipEventCreate(&e), "hipEventCreate");
    hipEvent_t &p0 = ev[0], &p1 = ev[1], &panel_done = ev[2], &rest_done = ev[3], &l0 = ev[4], &l1 = ev[5];
    const uint32_t waves = uint32_t(threads);
    bool rest_pending = false;
    auto args = [&](uint64_t a, uint64_t b) {
        return lf::FactorArgs{W, pitch, n32, uint32_t(a), uint32_t(b), (lf::FactorInfo*)dinfo.p, (uint32_t*)dscratch.p, swords,
                              (uint32_t*)drecip.p, rpitch};
    };
    // the panel of block [a, b) on s, its pivot record on its way to the host, panel_done recorded
    auto panel = [&](uint64_t a, uint64_t b) -> int {
        const lf::FactorArgs f = args(a, b);
        FACTOR_HIP(hipEventRecord(p0, s), "hipEventRecord");
        FACTOR_HIP(tables().diag[N](f, uint32_t(std::min<uint64_t>(256, (b - a + 63) / 64 * 64)), s), "panel launch");
        FACTOR_HIP(tables().panel[N](f, s), "panel launch");
        FACTOR_HIP(hipEventRecord(p1, s), "hipEventRecord");
        FACTOR_HIP(hipMemcpyAsync(hinfo.p, dinfo.p, sizeof(lf::FactorInfo), hipMemcpyDeviceToHost, s), "download pivot");
        FACTOR_HIP(hipEventRecord(panel_done, s), "hipEventRecord");
        return LF_OK;
    };
    // waits for the panel just enqueued; LF_OK and `failed` set when its pivot failed
    auto finish_panel = [&](Clock::time_point t, bool& failed) -> int {
        FACTOR_HIP(hipEventSynchronize(panel_done), "device execution (panel)");
        float ms = 0;
        FACTOR_HIP(hipEventElapsedTime(&ms, p0, p1), "hipEventElapsedTime");
        ci.timing_gpu_seconds += ms * 1e-3;
        ci.panel_seconds += since(t);
        ++ci.blocks;
        if (hinfo.p->error) return fail(LF_LOGIC_ERROR, "cholesky: exact-dot scratch overflow");
        failed = hinfo.p->pivot >= 0;
        if (failed) {
            ci.pivot = uint64_t(hinfo.p->pivot);
            ci.pivot_sign = hinfo.p->sign;
            ci.pivot_status = hinfo.p->status;
            if (rest_pending) FACTOR_HIP(hipStreamWaitEvent(s, rest_done, 0), "hipStreamWaitEvent");
            FACTOR_HIP(tables().invalid[N](W, pitch, n32, uint32_t(hinfo.p->pivot), s), "invalid launch");
            FACTOR_HIP(hipStreamSynchronize(s), "device execution");
        }
        return LF_OK;
    };
    {
        const auto t = Clock::now();
        if (int r = panel(0, std::min(n, nb))) return r;
        bool failed = false;
        if (int r = finish_panel(t, failed)) return r;
        if (failed) {
            ci.timing_wall_seconds = since(start);
            if (info) *info = ci;
            return LF_OK;
        }
    }
    for (uint64_t k0 = 0; k0 < n;) {
        const uint64_t k1 = std::min(n, k0 + nb), w = k1 - k0;
        if (k1 >= n) break;
        const uint64_t k2 = std::min(n, k1 + nb);
        // (1) the next panel's columns, after the previous block's rest of the update
        auto t = Clock::now();
        if (rest_pending) FACTOR_HIP(hipStreamWaitEvent(s, rest_done, 0), "hipStreamWaitEvent");
        FACTOR_HIP(hipEventRecord(l0, s), "hipEventRecord");
        FACTOR_HIP(tables().lookahead[N](args(k0, k1), uint32_t(k2), waves, s), "look-ahead launch");
        FACTOR_HIP(hipEventRecord(l1, s), "hipEventRecord");
        // (2) the next panel on s
        const auto tp = Clock::now();
        if (int r = panel(k1, k2)) return r;
        // (3) the rest of this block's update on the second stream, after this block's panel (the look-ahead
        // and the next panel write only columns [k1, k2), which this part neither reads nor writes)
        rest_pending = false;
        if (k2 < n) {
            FACTOR_HIP(hipStreamWaitEvent(upd.s, l0, 0), "hipStreamWaitEvent");
            const lf::MatView v{W + (k2 * n + k0), pitch, n, 1};
            float ms = 0;
            if (int r = update(bits, lo, true, v, v, n - k2, n - k2, w, W + (k2 * n + k2), pitch, n, upd.s, ms)) return r;
            FACTOR_HIP(hipEventRecord(rest_done, upd.s), "hipEventRecord");
            rest_pending = true;
            ci.timing_gpu_seconds += ms * 1e-3;
            ++ci.gpu_updates;
        }
        ci.update_seconds += since(t);
        {
            float ms = 0;
            FACTOR_HIP(hipEventSynchronize(l1), "device execution (look-ahead)");
            FACTOR_HIP(hipEventElapsedTime(&ms, l0, l1), "hipEventElapsedTime");
            ci.timing_gpu_seconds += ms * 1e-3;
        }
        // (4) the next panel's pivot
        bool failed = false;
        if (int r = finish_panel(tp, failed)) return r;
        if (failed) break;
        k0 = k1;
    }
    if (rest_pending) FACTOR_HIP(hipStreamSynchronize(upd.s), "device execution (update)");
    ci.timing_wall_seconds = since(start);
    if (info) *info = ci;
    return LF_OK;
}

int lf_solve(int bits, int passes, const lf_operand* L, uint64_t n, const lf_operand* B, uint64_t nrhs, const lf_operand* X,
             const lf_factor_options* options, const lf_linalg_options* linalg, double* gpu_seconds, double* wall_seconds, void* stream)
{
    const auto start = Clock::now();
    if (gpu_seconds) *gpu_seconds = 0;
    if (wall_seconds) *wall_seconds = 0;
    lf_factor_options fo;
    lf_linalg_options lo;
    if (int r = check_factor(bits, n, nrhs, options, fo, linalg, lo)) return r;
    if (passes < 1 || passes > 3) return fail(LF_INVALID_ARGUMENT, "solve: passes 1 (L X = B), 2 (L^T X = B) or 3 (both)");
    if (!n || !nrhs) return LF_OK;
    if (int r = check_matrix(L, n * n, "L")) return r;
    if (int r = check_matrix(B, n * nrhs, "B")) return r;
    if (int r = check_matrix(X, n * nrhs, "X")) return r;
    if (int r = check_overlap(bits / 32, *B, *X, true, "solve")) return r;
    if (int r = check_overlap(bits / 32, *L, *X, false, "solve")) return r;
    const int N = bits / 32;
    const hipStream_t s = (hipStream_t)stream;
    const uint64_t nb = fo.block ? std::min<uint64_t>(fo.block, n) : n;
    const uint32_t* F = (const uint32_t*)L->planes;
    const uint64_t fpitch = L->pitch;
    uint32_t* Xp = (uint32_t*)X->planes;
    const uint64_t xpitch = X->pitch;
    if (int r = copy_planes(N, *B, *X, n * nrhs, s)) return r;
    DevBuf dscratch, derr;
    const uint64_t swords = lf::factor_scratch_words(N, nb);
    // scratch slots: one per right-hand side for the in-block kernels, and the wavefronts of the direct
    // off-block updates
    const uint64_t slots = std::max<uint64_t>(nrhs, SOLVE_UPDATE_WAVES);
    if (2 * swords * slots * 4 > (uint64_t(32) << 30))
        return fail(LF_INVALID_ARGUMENT, "solve: the block is too large for the exact-dot scratch (use a smaller block)");
    FACTOR_HIP(dscratch.alloc(2 * swords * slots * 4), "hipMalloc (scratch)");
    FACTOR_HIP(derr.alloc(4), "hipMalloc");
    FACTOR_HIP(hipMemsetAsync(derr.p, 0, 4, s), "hipMemsetAsync");
    // reciprocals of the n diagonal entries, for every division of both passes
    DevBuf drecip;
    const uint64_t rpitch = (n + 63) / 64 * 64;
    FACTOR_HIP(drecip.alloc((N + 1) * rpitch * 4), "hipMalloc (reciprocals)");
    FACTOR_HIP(tables().reciprocals[N](F, fpitch, uint32_t(n), (uint32_t*)drecip.p, rpitch, s), "reciprocals launch");
    double gpu = 0;
    // device time: an event pair around every in-block kernel and direct update, read after the last
    // synchronization (the update GEMM reports its own kernels' time)
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events()
        {
            for (auto x : e) (void)hipEventDestroy(x);
        }
    } events;
    auto timed = [&](auto&& launch) -> hipError_t {
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t e = hipEventCreate(&a);
        if (e != hipSuccess) return e;
        events.e.push_back(a);
        if ((e = hipEventCreate(&b)) != hipSuccess) return e;
        events.e.push_back(b);
        if ((e = hipEventRecord(a, s)) != hipSuccess) return e;
        if ((e = launch()) != hipSuccess) return e;
        return hipEventRecord(b, s);
    };
    auto block_solve = [&](int backward, uint64_t r0, uint64_t r1) {
        lf::SolveArgs a{F, fpitch, Xp, xpitch, uint32_t(n), uint32_t(nrhs), uint32_t(r0), uint32_t(r1), (uint32_t*)dscratch.p, swords,
                        (int32_t*)derr.p, (const uint32_t*)drecip.p, rpitch};
        return timed([&] { return tables().solve[N](backward, a, s); });
    };
    float ms = 0;
    // An off-block update of fewer multiply-adds than solve_host_macs (the reference's threshold for taking
    // such an update off the GPU, linalg.hpp:73-78), or of at most SOLVE_DIRECT_TERMS, runs as direct exact
    // dots on the stream, without the update GEMM's planning and its synchronisation; larger ones are
    // products of the update GEMM.  A negative solve_host_macs sends every update to the update GEMM.
    auto direct = [&](uint64_t rows, uint64_t K) {
        const double terms = double(rows) * double(nrhs) * double(K);
        return fo.solve_host_macs >= 0 && (terms < fo.solve_host_macs || terms <= SOLVE_DIRECT_TERMS);
    };
    auto dot_update = [&](int backward, uint64_t r0, uint64_t r1) {
        lf::SolveArgs a{F, fpitch, Xp, xpitch, uint32_t(n), uint32_t(nrhs), uint32_t(r0), uint32_t(r1), (uint32_t*)dscratch.p, swords,
                        (int32_t*)derr.p, (const uint32_t*)drecip.p, rpitch};
        return timed([&] { return tables().solve_update[N](backward, a, uint32_t(slots), s); });
    };
    if (passes & 1)
        for (uint64_t r0 = 0; r0 < n; r0 += nb) {
            const uint64_t r1 = std::min(n, r0 + nb), K = r1 - r0;
            FACTOR_HIP(block_solve(0, r0, r1), "solve launch");
            if (r1 == n) break;
            // rows [r1, n) of X -= L[r1:, r0:r1] X[r0:r1, :] (the update waits for the block on the stream)
            if (direct(n - r1, K)) {
                FACTOR_HIP(dot_update(0, r0, r1), "solve update launch");
                continue;
            }
            const lf::MatView left{F + (r1 * n + r0), fpitch, n, 1}, right{Xp + r0 * nrhs, xpitch, 1, nrhs};
            if (int r = update(bits, lo, false, left, right, n - r1, nrhs, K, Xp + r1 * nrhs, xpitch, nrhs, s, ms)) return r;
            gpu += ms * 1e-3;
        }
    if (passes & 2)
        for (uint64_t r0 = (n - 1) / nb * nb;; r0 -= nb) {
            const uint64_t r1 = std::min(n, r0 + nb), K = r1 - r0;
            FACTOR_HIP(block_solve(1, r0, r1), "solve launch");
            if (!r0) break;
            // rows [0, r0) of X -= L[r0:r1, :r0]^T X[r0:r1, :]
            if (direct(r0, K)) {
                FACTOR_HIP(dot_update(1, r0, r1), "solve update launch");
                continue;
            }
            const lf::MatView left{F + r0 * n, fpitch, 1, n}, right{Xp + r0 * nrhs, xpitch, 1, nrhs};
            if (int r = update(bits, lo, false, left, right, r0, nrhs, K, Xp, xpitch, nrhs, s, ms)) return r;
            gpu += ms * 1e-3;
        }
    int32_t err = 0;
    FACTOR_HIP(hipMemcpyAsync(&err, derr.p, 4, hipMemcpyDeviceToHost, s), "download status");
    FACTOR_HIP(hipStreamSynchronize(s), "device execution (solve)");
    for (std::size_t k = 0; k + 1 < events.e.size(); k += 2) {
        float t = 0;
        if (hipEventElapsedTime(&t, events.e[k], events.e[k + 1]) == hipSuccess) gpu += t * 1e-3;
    }
    if (err) return fail(LF_LOGIC_ERROR, "solve: exact-dot scratch overflow");
    if (gpu_seconds) *gpu_seconds = gpu;
    if (wall_seconds) *wall_seconds = since(start);
    return LF_OK;
}

// ---- host-array forms: upload, the device call on one stream, download ----
int lf_cholesky_host(int bits, const void* A, uint64_t n, void* L, const lf_factor_options* options, const lf_linalg_options* linalg,
                     lf_cholesky_info* info)
{
    const auto start = Clock::now();
    lf_factor_options fo;
    lf_linalg_options lo;
    if (int r = check_factor(bits, n, 0, options, fo, linalg, lo)) return r;
    if (n && (!A || !L)) return fail(LF_INVALID_ARGUMENT, "null input");
    if (!n) {
        if (info) {
            std::memset(info, 0, sizeof *info);
        }
        return LF_OK;
    }
    const std::size_t E = bits / 8 + 12, count = n * n;
    OwnedStream st;
    FACTOR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking), "hipStreamCreate");
    DevBuf aos, planes;
    FACTOR_HIP(aos.alloc(count * E), "hipMalloc (staging)");
    FACTOR_HIP(planes.alloc(lf_planes_bytes(bits, 0, count)), "hipMalloc (work matrix)");
    lf_operand W{planes.p, lf_plane_pitch(count)};
    FACTOR_HIP(lf::copy_to_device(aos.p, A, count * E, st.s), "upload A");
    if (int r = lf_aos_to_planes(bits, 0, aos.p, &W, count, st.s)) return r;
    lf_cholesky_info ci;
    if (int r = lf_cholesky(bits, &W, n, &W, &fo, &lo, &ci, st.s)) return r;
    if (int r = lf_planes_to_aos(bits, 0, &W, aos.p, count, st.s)) return r;
    FACTOR_HIP(lf::copy_to_host(L, aos.p, count * E, st.s), "download L");
    FACTOR_HIP(hipStreamSynchronize(st.s), "device execution (cholesky)");
    ci.timing_wall_seconds = since(start);
    if (info) *info = ci;
    return LF_OK;
}

int lf_solve_host(int bits, int passes, const void* L, uint64_t n, const void* B, uint64_t nrhs, void* X, const lf_factor_options* options,
                  const lf_linalg_options* linalg, double* gpu_seconds, double* wall_seconds)
{
    const auto start = Clock::now();
    if (gpu_seconds) *gpu_seconds = 0;
    if (wall_seconds) *wall_seconds = 0;
    lf_factor_options fo;
    lf_linalg_options lo;
    if (int r = check_factor(bits, n, nrhs, options, fo, linalg, lo)) return r;
    if (passes < 1 || passes > 3) return fail(LF_INVALID_ARGUMENT, "solve: passes 1 (L X = B), 2 (L^T X = B) or 3 (both)");
    if (!n || !nrhs) return LF_OK;
    if (!L || !B || !X) return fail(LF_INVALID_ARGUMENT, "null input");
    const std::size_t E = bits / 8 + 12, fc = n * n, xc = n * nrhs;
    OwnedStream st;
    FACTOR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking), "hipStreamCreate");
    DevBuf aos, fplanes, xplanes;
    FACTOR_HIP(aos.alloc(std::max(fc, xc) * E), "hipMalloc (staging)");
    FACTOR_HIP(fplanes.alloc(lf_planes_bytes(bits, 0, fc)), "hipMalloc (factor)");
    FACTOR_HIP(xplanes.alloc(lf_planes_bytes(bits, 0, xc)), "hipMalloc (right-hand sides)");
    lf_operand F{fplanes.p, lf_plane_pitch(fc)}, Xo{xplanes.p, lf_plane_pitch(xc)};
    FACTOR_HIP(lf::copy_to_device(aos.p, L, fc * E, st.s), "upload L");
    if (int r = lf_aos_to_planes(bits, 0, aos.p, &F, fc, st.s)) return r;
    FACTOR_HIP(hipStreamSynchronize(st.s), "upload L");
    FACTOR_HIP(lf::copy_to_device(aos.p, B, xc * E, st.s), "upload B");
    if (int r = lf_aos_to_planes(bits, 0, aos.p, &Xo, xc, st.s)) return r;
    double g = 0;
    if (int r = lf_solve(bits, passes, &F, n, &Xo, nrhs, &Xo, &fo, &lo, &g, nullptr, st.s)) return r;
    if (int r = lf_planes_to_aos(bits, 0, &Xo, aos.p, xc, st.s)) return r;
    FACTOR_HIP(lf::copy_to_host(X, aos.p, xc * E, st.s), "download X");
    FACTOR_HIP(hipStreamSynchronize(st.s), "device execution (solve)");
    if (gpu_seconds) *gpu_seconds = g;
    if (wall_seconds) *wall_seconds = since(start);
    return LF_OK;
}

int lf_factor_step_cycles(int bits, int reps, uint64_t* cycles)
{
    if (bits < 64 || bits > 1024 || bits % 32) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (reps < 1 || reps > 1024) return fail(LF_INVALID_ARGUMENT, "reps must be in [1, 1024]");
    if (!cycles) return fail(LF_INVALID_ARGUMENT, "null pointer");
    const int N = bits / 32;
    constexpr uint64_t pitch = 128; // 67 values: terms 0-30 and 32-62, addend 64, divisor 65, dividend 66
    std::vector<uint32_t> h((N + 3) * pitch, 0);
    uint64_t state = 0x9E3779B97F4A7C15ull + (uint64_t)bits;
    auto next = [&] {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return state;
    };
    for (uint64_t i = 0; i < 67; ++i) {
        for (int k = 0; k < N; ++k) h[k * pitch + i] = (uint32_t)next();
        h[(N - 1) * pitch + i] |= 0x80000000u;
        h[N * pitch + i] = (uint32_t)((int)(next() % 41) - 20);
        h[(N + 1) * pitch + i] = (i >= 64 || (next() & 1)) ? 1u : (uint32_t)-1;
    }
    const uint64_t swords = lf::factor_scratch_words(N, 32);
    DevBuf dv, ds, derr, dout, dsink;
    FACTOR_HIP(dv.alloc(h.size() * 4), "hipMalloc");
    FACTOR_HIP(ds.alloc(2 * swords * 4), "hipMalloc");
    FACTOR_HIP(derr.alloc(4), "hipMalloc");
    FACTOR_HIP(dout.alloc(6 * 8), "hipMalloc");
    FACTOR_HIP(dsink.alloc(64 * 4), "hipMalloc");
    OwnedStream st;
    FACTOR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking), "hipStreamCreate");
    FACTOR_HIP(hipMemcpyAsync(dv.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, st.s), "upload");
    FACTOR_HIP(hipMemsetAsync(derr.p, 0, 4, st.s), "hipMemsetAsync");
    FACTOR_HIP(tables().steps[N]((const uint32_t*)dv.p, pitch, (uint32_t*)ds.p, swords, (int32_t*)derr.p, reps, (uint64_t*)dout.p,
                                 (uint32_t*)dsink.p, st.s),
               "step timing launch");
    int32_t err = 0;
    FACTOR_HIP(hipMemcpyAsync(cycles, dout.p, 6 * 8, hipMemcpyDeviceToHost, st.s), "download");
    FACTOR_HIP(hipMemcpyAsync(&err, derr.p, 4, hipMemcpyDeviceToHost, st.s), "download status");
    FACTOR_HIP(hipStreamSynchronize(st.s), "step timing");
    if (err) return fail(LF_LOGIC_ERROR, "step timing: exact-dot scratch overflow");
    return LF_OK;
}

} // extern "C"

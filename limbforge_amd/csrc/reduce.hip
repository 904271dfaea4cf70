// reduce.hip — C ABI of the reductions: segmented dot products (Engine::segmented_dot /
// CommandBatch::segmented_dot) and the adjacent-pair tree sum (CommandBatch::tree_sum); argument checks with
// the reference's error classes and the width dispatch to reduce_w.hip.
#include "host_run.hpp"
#include "host_copy.hpp"
#include "limbforge_hip.h"
#include "reduce_launch.hpp"

#include <chrono>
#include <limits>
#include <string>

namespace {
typedef hipError_t (*segdot_fn)(const lf::SegDotArgs&, bool, hipStream_t);
typedef hipError_t (*tree_fn)(const lf::TreeArgs&, bool, hipStream_t);

template <int N> void fill(segdot_fn* s, tree_fn* t)
{
    s[N] = lf::launch_segmented_dot<N>;
    t[N] = lf::launch_tree_pass<N>;
    if constexpr (N < 32) fill<N + 1>(s, t);
}
struct Tables {
    segdot_fn segdot[33] = {};
    tree_fn tree[33] = {};
    Tables() { fill<2>(segdot, tree); }
};
const Tables& tables()
{
    static const Tables t;
    return t;
}

int fail(int code, const std::string& m) { return lf::set_error(code, m.c_str()); }
int device_error(hipError_t e, const char* where) { return fail(LF_RUNTIME_ERROR, std::string(where) + ": " + hipGetErrorString(e)); }
#define RED_HIP(call, where)                                                                                           \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return device_error(e_, where);                                                          \
    } while (0)

bool valid_bits(int bits) { return bits >= 64 && bits <= 1024 && bits % 32 == 0; }

// element 0 of `planes` planes set to zero: the canonical zero of an empty sum (a kernel rather than a
// memset node, so the call also behaves inside captured graphs)
__global__ void zero_first_kernel(uint32_t* p, uint64_t pitch, int planes)
{
    if ((int)threadIdx.x < planes) p[threadIdx.x * pitch] = 0;
}

int check_planes(const lf_operand* o, uint64_t elements, const char* name)
{
    if (!elements) return LF_OK;
    if (!o || !o->planes) return fail(LF_INVALID_ARGUMENT, std::string("null operand ") + name);
    if (o->pitch < elements) return fail(LF_INVALID_ARGUMENT, std::string("operand ") + name + " pitch is smaller than its element count");
    return LF_OK;
}

int check_segdot(int bits, uint64_t segments, uint64_t length)
{
    if (!valid_bits(bits)) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (length && segments > std::numeric_limits<uint64_t>::max() / length)
        return fail(LF_INVALID_ARGUMENT, "segmented dot: segments x length overflows");
    if (segments * length > (uint64_t(1) << 40) || segments > std::numeric_limits<uint32_t>::max())
        return fail(LF_INVALID_ARGUMENT, "segmented dot: shape exceeds the supported indexing");
    return LF_OK;
}
} // namespace

extern "C" {

int lf_segmented_dot(int bits, int complex_values, uint64_t segments, uint64_t length, int shared_right, const lf_operand* a,
                     const lf_operand* b, const lf_operand* out, void* stream)
{
    if (int r = check_segdot(bits, segments, length)) return r;
    if (!segments) return LF_OK;
    if (int r = check_planes(out, segments, "out")) return r;
    if (length) {
        if (int r = check_planes(a, segments * length, "a")) return r;
        if (int r = check_planes(b, shared_right ? length : segments * length, "b")) return r;
    }
    lf::SegDotArgs g{};
    g.a = length ? (const uint32_t*)a->planes : nullptr;
    g.b = length ? (const uint32_t*)b->planes : nullptr;
    g.pa = length ? a->pitch : 0;
    g.pb = length ? b->pitch : 0;
    g.out = (uint32_t*)out->planes;
    g.po = out->pitch;
    g.segments = segments;
    g.length = length;
    g.shared_right = shared_right ? 1 : 0;
    RED_HIP(tables().segdot[bits / 32](g, complex_values != 0, (hipStream_t)stream), "segmented dot launch");
    return LF_OK;
}

int lf_segmented_dot_host(int bits, int complex_values, uint64_t segments, uint64_t length, int shared_right, const void* a,
                          const void* b, void* out, double* gpu_seconds, double* wall_seconds)
{
    const auto start = std::chrono::steady_clock::now();
    if (gpu_seconds) *gpu_seconds = 0;
    if (wall_seconds) *wall_seconds = 0;
    if (int r = check_segdot(bits, segments, length)) return r;
    if (!segments) return LF_OK;
    if (!out || (length && (!a || !b))) return fail(LF_INVALID_ARGUMENT, "null host array");
    const int cplx = complex_values ? 1 : 0;
    const uint64_t na = segments * length, nb = shared_right ? length : na;
    const uint64_t E = (uint64_t)(bits / 8 + 12) * (cplx ? 2 : 1);
    struct Dev {
        void* p = nullptr;
        ~Dev()
        {
            if (p) (void)hipFree(p);
        }
    } stage, pa, pb, po;
    hipStream_t s = nullptr;
    RED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const uint64_t big = na > segments ? na : segments;
    RED_HIP(hipMalloc(&stage.p, (big ? big : 1) * E), "hipMalloc (staging)");
    RED_HIP(hipMalloc(&pa.p, lf_planes_bytes(bits, cplx, na ? na : 1)), "hipMalloc (a)");
    RED_HIP(hipMalloc(&pb.p, lf_planes_bytes(bits, cplx, nb ? nb : 1)), "hipMalloc (b)");
    RED_HIP(hipMalloc(&po.p, lf_planes_bytes(bits, cplx, segments)), "hipMalloc (out)");
    lf_operand A{pa.p, lf_plane_pitch(na ? na : 1)}, B{pb.p, lf_plane_pitch(nb ? nb : 1)}, O{po.p, lf_plane_pitch(segments)};
    if (na) {
        RED_HIP(lf::copy_to_device(stage.p, a, na * E, s), "upload a");
        if (int r = lf_aos_to_planes(bits, cplx, stage.p, &A, na, s)) return r;
        RED_HIP(lf::copy_to_device(stage.p, b, nb * E, s), "upload b");
        if (int r = lf_aos_to_planes(bits, cplx, stage.p, &B, nb, s)) return r;
    }
    hipEvent_t e0, e1;
    RED_HIP(hipEventCreate(&e0), "hipEventCreate");
    RED_HIP(hipEventCreate(&e1), "hipEventCreate");
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    RED_HIP(hipEventRecord(e0, s), "hipEventRecord");
    if (int r = lf_segmented_dot(bits, cplx, segments, length, shared_right, &A, &B, &O, s)) return r;
    RED_HIP(hipEventRecord(e1, s), "hipEventRecord");
    if (int r = lf_planes_to_aos(bits, cplx, &O, stage.p, segments, s)) return r;
    RED_HIP(lf::copy_to_host(out, stage.p, segments * E, s), "download out");
    RED_HIP(hipStreamSynchronize(s), "device execution (segmented dot)");
    float ms = 0;
    RED_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (gpu_seconds) *gpu_seconds = ms * 1e-3;
    if (wall_seconds) *wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return LF_OK;
}

int lf_tree_sum(int bits, int complex_values, const lf_operand* in, uint64_t count, const lf_operand* out, void* stream)
{
    if (!valid_bits(bits)) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (count > (uint64_t(1) << 40)) return fail(LF_INVALID_ARGUMENT, "tree sum: input exceeds the supported indexing");
    if (int r = check_planes(out, 1, "out")) return r;
    const int cplx = complex_values ? 1 : 0;
    const hipStream_t s = (hipStream_t)stream;
    if (!count) { // canonical zero
        const int planes = (bits / 32 + 3) * (cplx ? 2 : 1);
        hipLaunchKernelGGL(zero_first_kernel, dim3(1), dim3(128), 0, s, (uint32_t*)out->planes, out->pitch, planes);
        RED_HIP(hipGetLastError(), "tree sum launch");
        return LF_OK;
    }
    if (int r = check_planes(in, count, "in")) return r;
    // passes of 512 values each, through two stream-ordered partial buffers
    uint64_t first = (count + lf::TREE_SPAN - 1) / lf::TREE_SPAN;
    void* part[2] = {nullptr, nullptr};
    uint64_t ppitch[2] = {0, 0};
    if (first > 1) {
        ppitch[0] = lf_plane_pitch(first);
        RED_HIP(hipMallocAsync(&part[0], lf_planes_bytes(bits, cplx, first), s), "hipMallocAsync (tree partials)");
        const uint64_t second = (first + lf::TREE_SPAN - 1) / lf::TREE_SPAN;
        if (second > 1) {
            ppitch[1] = lf_plane_pitch(second);
            RED_HIP(hipMallocAsync(&part[1], lf_planes_bytes(bits, cplx, second), s), "hipMallocAsync (tree partials)");
        }
    }
    lf::TreeArgs g{(const uint32_t*)in->planes, nullptr, in->pitch, 0, count};
    int level = 0, rc = LF_OK;
    while (true) {
        const uint64_t next = (g.count + lf::TREE_SPAN - 1) / lf::TREE_SPAN;
        if (next == 1) {
            g.out = (uint32_t*)out->planes;
            g.pout = out->pitch;
        } else {
            g.out = (uint32_t*)part[level % 2];
            g.pout = ppitch[level % 2];
        }
        const hipError_t e = tables().tree[bits / 32](g, cplx != 0, s);
        if (e != hipSuccess) {
            rc = device_error(e, "tree sum launch");
            break;
        }
        if (next == 1) break;
        g.in = g.out;
        g.pin = g.pout;
        g.count = next;
        ++level;
    }
    for (void* p : part)
        if (p) (void)hipFreeAsync(p, s);
    return rc;
}

} // extern "C"

// mixed.hip — mixed-precision chain batches behind lf_mixed_chains / lf_mixed_graph_create, and the
// launch-shape autotuner lf_autotune_mixed.
//
// A batch is a list of precision buckets (all values of one width, already resident as limb planes).
// lf_mixed_chains issues one launch per bucket in order on a stream: the plain "one launch per
// precision" form.  lf_mixed_graph_create builds one hipGraph of at most four persistent kernel nodes,
// one per group of eight widths (mixed_group.hip): each group kernel runs all of its buckets in one launch
// (workgroups find their bucket in a device table), so one graph launch runs every precision concurrently
// within the runtime's hardware queues.  The autotuner times the bucket
// kernel of a width on a synthetic sample for each workgroup size and keeps the fastest per device and
// width; the lanes tuner times a batch's graph with 1, 2 or 4 lanes per value in each group of widths.
#include "lf_internal.hpp"
#include "limbforge_hip.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

void lf_set_last_error(const std::string& message);

namespace {

int fail(int code, const std::string& m)
{
    lf_set_last_error(m);
    return code;
}

#define MIXED_HIP(call, where)                                                                                         \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(LF_RUNTIME_ERROR, std::string(where) + ": " + hipGetErrorString(e_));         \
    } while (0)

typedef hipError_t (*mixed_fn)(const lf::MixedArgs&, unsigned, hipStream_t);
typedef const void* (*entry_fn)();
struct MixedTables;
template <int N> void fill_mixed(MixedTables& t);
struct MixedTables {
    mixed_fn launch[33] = {};
    entry_fn entry[33] = {};
    MixedTables() { fill_mixed<2>(*this); }
};
template <int N> void fill_mixed(MixedTables& t)
{
    t.launch[N] = lf::launch_mixed_chain<N>;
    t.entry[N] = lf::mixed_chain_entry<N>;
    if constexpr (N < 32) fill_mixed<N + 1>(t);
}
const MixedTables& tables()
{
    static const MixedTables t;
    return t;
}

bool valid_bits(int bits) { return bits >= 64 && bits <= 1024 && bits % 32 == 0; }
bool valid_block(unsigned b) { return b == 64 || b == 128 || b == 256; }

int check_bucket(const lf_chain_bucket& b)
{
    if (!valid_bits(b.bits)) return fail(LF_INVALID_ARGUMENT, "bucket bits must be a multiple of 32 in [64,1024]");
    if (b.count > 0xffffffffull) return fail(LF_INVALID_ARGUMENT, "bucket exceeds 32-bit indexing");
    if (b.count && (!b.x.planes || !b.f.planes || !b.g.planes || !b.out.planes || !b.len || !b.mask))
        return fail(LF_INVALID_ARGUMENT, "null bucket operand");
    if (b.count && (b.x.pitch < b.count || b.x.pitch != b.f.pitch || b.x.pitch != b.g.pitch || b.x.pitch != b.out.pitch))
        return fail(LF_INVALID_ARGUMENT, "bucket operands need one pitch of at least count");
    return LF_OK;
}

lf::MixedArgs args_of(const lf_chain_bucket& b)
{
    return {(const uint32_t*)b.x.planes, (const uint32_t*)b.f.planes, (const uint32_t*)b.g.planes, (uint32_t*)b.out.planes, b.x.pitch,
            b.len, b.mask, uint32_t(b.count)};
}

std::mutex tuned_mutex;
std::map<std::pair<int, int>, unsigned> tuned; // (device, bits) -> workgroup size

unsigned block_for(const unsigned* block, int i, int bits)
{
    if (block && block[i]) return block[i];
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(tuned_mutex);
    auto it = tuned.find({dev, bits});
    return it == tuned.end() ? 256u : it->second;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

typedef const void* (*group_fn)(int);
const group_fn group_entries[4] = {lf::mixed_group_entry<0>, lf::mixed_group_entry<1>, lf::mixed_group_entry<2>, lf::mixed_group_entry<3>};

// Lanes per value of each group chosen by lf_autotune_mixed_lanes, per device; untuned: one lane.
std::map<int, std::vector<unsigned>> tuned_lanes;

void lanes_for(const unsigned* group_lanes, unsigned (&out)[4])
{
    for (int k = 0; k < 4; ++k) out[k] = 1;
    if (group_lanes) {
        for (int k = 0; k < 4; ++k) out[k] = group_lanes[k] ? group_lanes[k] : 1;
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(tuned_mutex);
    auto it = tuned_lanes.find(dev);
    if (it != tuned_lanes.end())
        for (int k = 0; k < 4; ++k) out[k] = it->second[k];
}

int check_lanes(const unsigned* group_lanes)
{
    if (!group_lanes) return LF_OK;
    for (int k = 0; k < 4; ++k) {
        const unsigned l = group_lanes[k];
        if (l > 1 && !group_entries[k](int(l)))
            return fail(LF_INVALID_ARGUMENT, "lanes per value must be 0 or 1 for the groups of 2-8 and 9-16 limbs, and 0, 1, 2 or 4 for the others");
        if (l != 0 && l != 1 && l != 2 && l != 4)
            return fail(LF_INVALID_ARGUMENT, "lanes per value must be 0, 1, 2 or 4");
    }
    return LF_OK;
}

// device tables of the graphs made by lf_mixed_graph_create, freed by lf_graph_destroy
std::mutex tables_mutex;
std::map<void*, void*> graph_tables;

} // namespace

// called by lf_graph_destroy (capi.hip) after the executable graph is destroyed
void lf_mixed_graph_release(void* graph_exec)
{
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(tables_mutex);
        auto it = graph_tables.find(graph_exec);
        if (it == graph_tables.end()) return;
        p = it->second;
        graph_tables.erase(it);
    }
    (void)hipFree(p);
}

extern "C" {

int lf_mixed_chains(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, void* stream)
{
    if (nbuckets < 0 || (nbuckets && !buckets)) return fail(LF_INVALID_ARGUMENT, "bad bucket list");
    for (int i = 0; i < nbuckets; ++i) {
        if (int r = check_bucket(buckets[i])) return r;
        if (block && block[i] && !valid_block(block[i])) return fail(LF_INVALID_ARGUMENT, "workgroup size must be 64, 128 or 256");
    }
    for (int i = 0; i < nbuckets; ++i) {
        const lf_chain_bucket& b = buckets[i];
        MIXED_HIP(tables().launch[b.bits / 32](args_of(b), block_for(block, i, b.bits), (hipStream_t)stream), "mixed chain launch");
    }
    return LF_OK;
}

int lf_mixed_graph_create(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, void** graph_exec)
{
    return lf_mixed_graph_create_lanes(buckets, nbuckets, block, nullptr, graph_exec);
}

int lf_mixed_graph_create_lanes(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, const unsigned* group_lanes,
                                void** graph_exec)
{
    if (!graph_exec) return fail(LF_INVALID_ARGUMENT, "null pointer");
    if (int r = check_lanes(group_lanes)) return r;
    unsigned lanes[4];
    lanes_for(group_lanes, lanes);
    *graph_exec = nullptr;
    if (nbuckets < 0 || (nbuckets && !buckets)) return fail(LF_INVALID_ARGUMENT, "bad bucket list");
    for (int i = 0; i < nbuckets; ++i) {
        if (int r = check_bucket(buckets[i])) return r;
        if (block && block[i] && !valid_block(block[i])) return fail(LF_INVALID_ARGUMENT, "workgroup size must be 64, 128 or 256");
    }
    // per group: the buckets' arguments, limb counts and first workgroups, in one device table
    struct Group {
        std::vector<lf::MixedArgs> args;
        std::vector<int32_t> words;
        std::vector<uint32_t> first{0};
        unsigned block = 0;
        int widest = 0;
    } grp[4];
    for (int i = 0; i < nbuckets; ++i) {
        const lf_chain_bucket& b = buckets[i];
        if (!b.count) continue;
        const int N = b.bits / 32, k = lf::mixed_group_of(N);
        if (N > grp[k].widest) {
            grp[k].widest = N;
            grp[k].block = block_for(block, i, b.bits);
        }
    }
    for (int i = 0; i < nbuckets; ++i) {
        const lf_chain_bucket& b = buckets[i];
        if (!b.count) continue;
        Group& G = grp[lf::mixed_group_of(b.bits / 32)];
        G.args.push_back(args_of(b));
        G.words.push_back(b.bits / 32);
        const uint64_t blocks = (uint64_t)G.first.back() + lf::grid_blocks((uint64_t)b.count * lanes[lf::mixed_group_of(b.bits / 32)], G.block);
        if (blocks > 0x7fffffffull) return fail(LF_INVALID_ARGUMENT, "a precision group of the batch needs more than 2^31 - 1 workgroups");
        G.first.push_back(uint32_t(blocks));
    }
    std::size_t bytes = 0, off[4][3] = {};
    for (int k = 0; k < 4; ++k) {
        auto align = [&](std::size_t n) { bytes = (bytes + 255) / 256 * 256; const std::size_t o = bytes; bytes += n; return o; };
        off[k][0] = align(grp[k].args.size() * sizeof(lf::MixedArgs));
        off[k][1] = align(grp[k].words.size() * sizeof(int32_t));
        off[k][2] = align(grp[k].first.size() * sizeof(uint32_t));
    }
    std::vector<unsigned char> host(std::max<std::size_t>(bytes, 1), 0);
    for (int k = 0; k < 4; ++k) {
        std::memcpy(host.data() + off[k][0], grp[k].args.data(), grp[k].args.size() * sizeof(lf::MixedArgs));
        std::memcpy(host.data() + off[k][1], grp[k].words.data(), grp[k].words.size() * sizeof(int32_t));
        std::memcpy(host.data() + off[k][2], grp[k].first.data(), grp[k].first.size() * sizeof(uint32_t));
    }
    DevBuf table;
    MIXED_HIP(hipMalloc(&table.p, host.size()), "hipMalloc (bucket tables)");
    MIXED_HIP(hipMemcpy(table.p, host.data(), host.size(), hipMemcpyHostToDevice), "upload bucket tables");
    hipGraph_t g = nullptr;
    MIXED_HIP(hipGraphCreate(&g, 0), "hipGraphCreate");
    struct GraphGuard {
        hipGraph_t g;
        ~GraphGuard() { (void)hipGraphDestroy(g); }
    } guard{g};
    // The runtime deals independent nodes to its hardware queues in creation order, and with four nodes
    // the fourth shares the second's queue (observed in rocprof traces of this GPU: three queues).  The
    // groups are created widest-but-one first, so that the two narrowest groups, whose values finish
    // soonest, are the pair that shares: 10^6 values with groups 1 and 3 sharing took 2.9-3.4 ms.
    static const int order[4] = {2, 0, 3, 1};
    hipGraphNode_t nodes[4] = {};
    for (int oi = 0; oi < 4; ++oi) {
        const int k = order[oi];
        if (grp[k].args.empty()) continue;
        const char* base = (const char*)table.p;
        const lf::MixedArgs* a = (const lf::MixedArgs*)(base + off[k][0]);
        const int32_t* w = (const int32_t*)(base + off[k][1]);
        const uint32_t* f = (const uint32_t*)(base + off[k][2]);
        uint32_t nb = uint32_t(grp[k].args.size());
        void* params[] = {&a, &w, &f, &nb};
        hipKernelNodeParams p;
        std::memset(&p, 0, sizeof p);
        p.func = const_cast<void*>(group_entries[k](int(lanes[k])));
        p.gridDim = dim3(grp[k].first.back());
        p.blockDim = dim3(grp[k].block);
        p.sharedMemBytes = 0;
        p.kernelParams = params;
        p.extra = nullptr;
        MIXED_HIP(hipGraphAddKernelNode(&nodes[k], g, nullptr, 0, &p), "hipGraphAddKernelNode");
    }
    hipGraphExec_t x = nullptr;
    MIXED_HIP(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    {
        std::lock_guard<std::mutex> lock(tables_mutex);
        graph_tables[(void*)x] = table.p;
    }
    table.p = nullptr; // owned by the graph now
    *graph_exec = x;
    return LF_OK;
}

int lf_autotune_mixed(int bits, uint64_t sample, unsigned* block, double* best_ms)
{
    if (!valid_bits(bits)) return fail(LF_INVALID_ARGUMENT, "bits must be a multiple of 32 in [64,1024]");
    if (!block) return fail(LF_INVALID_ARGUMENT, "null pointer");
    if (sample < 64 || sample > (uint64_t(1) << 24)) return fail(LF_INVALID_ARGUMENT, "autotune sample must be in [64, 2^24]");
    const int N = bits / 32;
    const uint64_t pitch = lf_plane_pitch(sample);
    // synthetic sample: random normalized operands (exponents in [-64, 64]), 32-step chains, half fma
    std::mt19937_64 rng(20261007u + unsigned(bits));
    std::vector<uint32_t> planes(3 * (N + 3) * pitch, 0);
    for (int o = 0; o < 3; ++o) {
        uint32_t* p = planes.data() + o * (N + 3) * pitch;
        for (uint64_t i = 0; i < sample; ++i) {
            for (int k = 0; k < N; ++k) p[k * pitch + i] = uint32_t(rng());
            p[(N - 1) * pitch + i] |= 0x80000000u;
            p[N * pitch + i] = uint32_t(int(rng() % 129) - 64);
            p[(N + 1) * pitch + i] = rng() & 1 ? 1u : uint32_t(-1);
        }
    }
    std::vector<uint32_t> len(sample, 32);
    std::vector<uint64_t> mask(sample);
    for (auto& m : mask) m = rng();
    DevBuf dx, dout, dlen, dmask;
    const size_t pbytes = (N + 3) * pitch * 4;
    MIXED_HIP(hipMalloc(&dx.p, 3 * pbytes), "hipMalloc");
    MIXED_HIP(hipMalloc(&dout.p, pbytes), "hipMalloc");
    MIXED_HIP(hipMalloc(&dlen.p, sample * 4), "hipMalloc");
    MIXED_HIP(hipMalloc(&dmask.p, sample * 8), "hipMalloc");
    MIXED_HIP(hipMemcpy(dx.p, planes.data(), 3 * pbytes, hipMemcpyHostToDevice), "upload");
    MIXED_HIP(hipMemcpy(dlen.p, len.data(), sample * 4, hipMemcpyHostToDevice), "upload");
    MIXED_HIP(hipMemcpy(dmask.p, mask.data(), sample * 8, hipMemcpyHostToDevice), "upload");
    const char* base = (const char*)dx.p;
    lf::MixedArgs a{(const uint32_t*)base, (const uint32_t*)(base + pbytes), (const uint32_t*)(base + 2 * pbytes), (uint32_t*)dout.p, pitch,
                    (const uint32_t*)dlen.p, (const uint64_t*)dmask.p, uint32_t(sample)};
    hipStream_t s = nullptr;
    MIXED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    unsigned best = 256;
    float best_t = 1e30f;
    for (unsigned blk : {64u, 128u, 256u}) {
        float t_min = 1e30f;
        for (int rep = 0; rep < 4 && err == hipSuccess; ++rep) { // the first run warms
            err = hipEventRecord(e0, s);
            if (err == hipSuccess) err = tables().launch[N](a, blk, s);
            if (err == hipSuccess) err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float t = 0;
            if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
            if (rep && t < t_min) t_min = t;
        }
        if (err != hipSuccess) break;
        if (t_min < best_t) {
            best_t = t_min;
            best = blk;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (err != hipSuccess) return fail(LF_RUNTIME_ERROR, std::string("autotune: ") + hipGetErrorString(err));
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        std::lock_guard<std::mutex> lock(tuned_mutex);
        tuned[{dev, bits}] = best;
    }
    *block = best;
    if (best_ms) *best_ms = best_t;
    return LF_OK;
}

int lf_autotune_mixed_lanes(const lf_chain_bucket* buckets, int nbuckets, const unsigned* block, unsigned* group_lanes, float* best_ms)
{
    if (!group_lanes) return fail(LF_INVALID_ARGUMENT, "null pointer");
    if (nbuckets < 1 || !buckets) return fail(LF_INVALID_ARGUMENT, "bad bucket list");
    bool present[4] = {};
    for (int i = 0; i < nbuckets; ++i) {
        if (int r = check_bucket(buckets[i])) return r;
        if (buckets[i].count) present[lf::mixed_group_of(buckets[i].bits / 32)] = true;
    }
    hipStream_t s = nullptr;
    MIXED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    // the batch's graph with the given lanes: best of three replays after a warm-up
    auto time_graph = [&](const unsigned (&lanes)[4], float& ms) -> int {
        void* x = nullptr;
        if (int r = lf_mixed_graph_create_lanes(buckets, nbuckets, block, lanes, &x)) return r;
        ms = 1e30f;
        for (int rep = 0; rep < 4 && err == hipSuccess; ++rep) {
            err = hipEventRecord(e0, s);
            if (err == hipSuccess) err = hipGraphLaunch((hipGraphExec_t)x, s);
            if (err == hipSuccess) err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float t = 0;
            if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
            if (rep && t < ms) ms = t;
        }
        (void)hipGraphExecDestroy((hipGraphExec_t)x);
        lf_mixed_graph_release(x);
        return LF_OK;
    };
    // coordinate descent from one lane everywhere, widest group first (it bounds the batch)
    unsigned best[4] = {1, 1, 1, 1};
    float best_t = 1e30f;
    int rc = err == hipSuccess ? time_graph(best, best_t) : LF_OK;
    for (int k = 3; k >= 2 && rc == LF_OK && err == hipSuccess; --k) {
        if (!present[k]) continue;
        unsigned base = best[k];
        for (unsigned l : {1u, 2u, 4u}) {
            if (l == base) continue;
            unsigned trial[4] = {best[0], best[1], best[2], best[3]};
            trial[k] = l;
            float t = 0;
            rc = time_graph(trial, t);
            if (rc != LF_OK || err != hipSuccess) break;
            if (t < best_t) {
                best_t = t;
                best[k] = l;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc != LF_OK) return rc;
    if (err != hipSuccess) return fail(LF_RUNTIME_ERROR, std::string("autotune: ") + hipGetErrorString(err));
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        std::lock_guard<std::mutex> lock(tuned_mutex);
        tuned_lanes[dev] = std::vector<unsigned>(best, best + 4);
    }
    for (int k = 0; k < 4; ++k) group_lanes[k] = best[k];
    if (best_ms) *best_ms = best_t;
    return LF_OK;
}

} // extern "C"

// noalloc.hip — lf_run_host_noalloc: the host-array element-wise call (lf_run_host) for callers that may
// not allocate device memory (serving loops, harnesses that forbid hipMalloc).  Its device staging is one
// statically sized array per GPU, held until the process ends; host code never uses the array's name as
// an address: the array's device address on the current GPU comes from hipGetSymbolAddress, looked up
// once per device and kept.  Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU race.  Nothing in the array needs clearing: each call writes every staging word it later reads.
#include "host_run.hpp"

#include <mutex>
#include <string>

constexpr uint64_t SCRATCH_BYTES = uint64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

int scratch_address(char** base)
{
    static std::mutex lock;
    static char* address[64] = {};
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return lf::set_error(LF_RUNTIME_ERROR, "run_host_noalloc: hipGetDevice failed");
    if (device < 0 || device >= 64) return lf::set_error(LF_RUNTIME_ERROR, "run_host_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> hold(lock);
    if (!address[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return lf::set_error(LF_RUNTIME_ERROR, (std::string("run_host_noalloc: hipGetSymbolAddress (scratch): ") + hipGetErrorString(e)).c_str());
        address[device] = (char*)ptr;
    }
    *base = address[device];
    return LF_OK;
}

uint64_t rounded(uint64_t bytes) { return (bytes + 255) / 256 * 256; }
} // namespace

extern "C" {

int lf_run_host_noalloc(int bits, int op, const void* a, const void* b, const void* c, void* out, uint64_t count,
                        const lf_broadcast* b_index, const lf_broadcast* c_index, void* stream)
{
    lf::HostRun R;
    if (int r = lf::plan_host_run(bits, op, a, b, c, out, count, b_index, c_index, R)) return r;
    if (!count) return LF_OK;
    uint64_t total = 0;
    for (int k = 0; k < 5; ++k) total += rounded(R.bytes[k]);
    if (total > SCRATCH_BYTES)
        return lf::set_error(LF_INVALID_ARGUMENT, ("run_host_noalloc: the batch needs " + std::to_string(total) +
                                                   " bytes of device staging, more than the " + std::to_string(SCRATCH_BYTES) +
                                                   "-byte scratch array")
                                                      .c_str());
    char* base = nullptr;
    if (int r = scratch_address(&base)) return r;
    void* dev[5];
    for (int k = 0; k < 5; ++k) {
        dev[k] = base;
        base += rounded(R.bytes[k]);
    }
    return lf::stage_and_run(R, a, b, c, out, dev, (hipStream_t)stream, nullptr, false);
}

int lf_noalloc_scratch_bytes(uint64_t* bytes)
{
    if (!bytes) return lf::set_error(LF_INVALID_ARGUMENT, "null pointer");
    *bytes = SCRATCH_BYTES;
    return LF_OK;
}

} // extern "C"

// host_run.hpp — the host-array element-wise call (lf_run_host and its variants): argument checks and
// staging sizes, then the staged run through device limb planes.
#pragma once
#include "limbforge_hip.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lf {
// The checked shape of one host-array element-wise call and the device staging it needs: [0] AoS bytes of
// the largest operand, [1] a's planes, [2] b's, [3] c's, [4] the result's.
struct HostRun {
    int bits = 0, op = 0, cplx = 0;
    uint64_t count = 0, nb = 0, nc = 0, E = 0;
    const lf_broadcast* b_index = nullptr;
    const lf_broadcast* c_index = nullptr;
    uint64_t bytes[5] = {};
};

// Checks the arguments (the reference's error classes, as a failing lf_* call) and fills R; count 0 is valid
// and leaves R.count 0.
int plan_host_run(int bits, int op, const void* a, const void* b, const void* c, void* out, uint64_t count, const lf_broadcast* b_index,
                  const lf_broadcast* c_index, HostRun& R);
// Copies the host operands into device staging dev[0..4] (sized by plan_host_run), transposes them into
// limb planes, runs the operation and brings the result back; returns after stream s has finished.
// ev (may be null): two events recorded around the operation's kernel.  ring: the host arrays pass through
// the calling thread's pinned chunk ring (host_copy.hpp, allocated on its first use); false: plain
// hipMemcpyAsync from the host arrays, for callers that must not allocate.
int stage_and_run(const HostRun& R, const void* a, const void* b, const void* c, void* out, void* const* dev, hipStream_t s,
                  const hipEvent_t* ev, bool ring = true);
// Records `message` as this host thread's lf_last_error and returns `code`.
int set_error(int code, const char* message);
} // namespace lf

// engine_internal.hpp — private state behind the C++ facade (Engine, Buffer storage, batches), shared by
// the facade's translation units (engine.cpp, linalg.cpp).  The role of the reference's
// src/engine_internal.hpp: units such as Linalg encode their own launches into a CommandBatch.
#pragma once
#include "limbforge/engine.hpp"
#include "limbforge_hip.h"

#include <atomic>
#include <chrono>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace limbforge {
namespace detail {

// Throws the reference's exception class for a C ABI error code (engine.mm:20-28).
[[noreturn]] void raise_error(int code, const std::string& where);
inline void check(int rc, const char* where)
{
    if (rc) raise_error(rc, where);
}

// One device allocation: limb planes for arithmetic formats, a plain array for status words.  A host
// shadow in the public AoS layout backs Buffer::mapped(); it is written back to the device before the
// buffer's next use by a submitted batch.
struct BufferStorage {
    std::shared_ptr<int> owner;        // identity of the engine that allocated it
    std::atomic<bool> busy{false};     // claimed by an unwaited submission
    void* device = nullptr;
    std::size_t count = 0, element_bytes = 0, bytes = 0;
    std::uint64_t pitch = 0;
    int bits = 0;
    bool complex = false;
    std::mutex mutex;
    std::vector<unsigned char> shadow;
    bool host_valid = false, host_dirty = false;

    ~BufferStorage();
    lf_operand operand() const { return {device, pitch}; }
    void flush();                                   // host shadow -> device, if the host may have written it
    void to_host(void* host, std::size_t elements); // device -> AoS host array
    void from_host(const void* host, std::size_t elements);
};

// One encoded launch of a batch: it enqueues itself on the stream it is given.
using Launch = std::function<void(void* stream)>;

// Hook used by units (Linalg) to encode into a batch and to reach the engine's device state.
struct BatchAccess {
    static void retain(CommandBatch& b, const std::shared_ptr<BufferStorage>& r, bool written);
    static void enqueue(CommandBatch& b, Launch launch);
    static void on_completion(CommandBatch& b, std::function<void()> f);
    // r stays provisional for the rest of the batch: a later write to it throws std::logic_error, a later
    // read sets *read_flag
    static void mark_provisional(CommandBatch& b, const std::shared_ptr<BufferStorage>& r, std::shared_ptr<std::atomic<bool>> read_flag);
    static std::shared_ptr<int> identity(const Engine& e);
    static std::shared_ptr<BufferStorage> allocate(Engine& e, std::size_t count, int bits);
};

} // namespace detail
} // namespace limbforge

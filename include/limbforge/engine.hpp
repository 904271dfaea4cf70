#pragma once
// Element-wise engine of the LimbForge/WolfNum API on AMD Instinct MI355X (gfx950).
//
// The public surface follows the reference's include/limbforge/engine.hpp for the batched limb path:
// Operation, Broadcast, Timing, EngineOptions, Prewarm (bits, operations), Buffer<T>, Submission,
// CommandBatch (run overloads, cast, submit) and Engine (run / run_unary / run_ternary / cast on host
// arrays, make_buffer, batch, prewarm, pipeline_info, device_name), with the same element layout
// (bits/8 + 12 bytes per real value), argument meaning and exceptions.  Behind it is the C ABI of
// liblimbforge_hip.so (include/limbforge_hip.h): resident buffers live in HBM as limb planes and
// batches are ordered launches on one HIP stream.  Tree sums and segmented dot products are provided;
// the reference engine's vector recurrences and batched 4x4 systems are not part of this build.
#include "core.hpp"
#include <cstddef>
#include <cstdint>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#define LIMBFORGE_RESIDENT_API 1

namespace limbforge {
// fma/fms: RN(a*b +/- c) with one rounding.
enum class Operation { add, sub, mul, div, complex_add, complex_mul, complex_div, square, sqrt, fma, fms, complex_fma, complex_fms };
inline bool operation_is_complex(Operation op)
{
    return (int(op) >= 4 && int(op) <= 6) || op == Operation::complex_fma || op == Operation::complex_fms;
}
inline bool operation_is_unary(Operation op) { return op == Operation::square || op == Operation::sqrt; }
inline bool operation_is_ternary(Operation op) { return int(op) >= 9 && int(op) <= 12; }
inline bool operation_is_valid(Operation op) { return int(op) >= 0 && int(op) <= 12; }

struct Timing {
    double gpu_seconds, wall_seconds;
};

// Segmented dot products: out[i] = the sequential chain s <- fma(a[i*length + k], b[i*length + k], s) from
// s = 0 (b[k] for every segment with shared_right; cfma per component for complex values).
struct SegmentedDot {
    std::size_t segments = 0, length = 0;
    bool shared_right = false;
};

// Broadcast read of an operand: element i uses index (i / stride) % period (period 0: no wrap).
struct Broadcast {
    std::size_t stride = 1, period = 0;
};
inline std::size_t broadcast_elements(std::size_t count, Broadcast b)
{
    if (!b.stride) throw std::invalid_argument("broadcast stride must be positive");
    std::size_t n = count / b.stride + (count % b.stride != 0);
    return b.period && b.period < n ? b.period : n;
}

// threads_per_threadgroup: 0 or a power of two in [32, 1024] (validated; the MI355X kernels use
// 256-lane workgroups).  The cooperative flags select reference kernels that are not part of this build.
struct EngineOptions {
    unsigned threads_per_threadgroup = 0;
    bool cooperative_reductions = true;
    bool cooperative_recurrence = true;
};
struct PipelineInfo {
    unsigned simd_width, max_threads, threads_per_threadgroup;
};
// Kernels are compiled ahead of time for gfx950, so prewarming only validates the request.
struct Prewarm {
    std::vector<int> bits;
    std::vector<Operation> operations;
    bool recurrence = false, reductions = false;
};

namespace detail {
struct BufferStorage;
void* mapped(const std::shared_ptr<BufferStorage>&);
void transfer(const std::shared_ptr<BufferStorage>&, void*, std::size_t, bool);
template <class T> struct Format;
template <int N> struct Format<Number<N>> {
    static_assert(N >= 2 && N <= 32, "unsupported precision");
    static constexpr int bits = 32 * N;
    static constexpr bool complex = false;
};
template <int N> struct Format<Complex<N>> {
    static_assert(N >= 2 && N <= 32, "unsupported precision");
    static constexpr int bits = 32 * N;
    static constexpr bool complex = true;
};
// Status words: not an arithmetic format.
template <> struct Format<std::uint32_t> {
    static constexpr int bits = 0;
    static constexpr bool complex = false;
};
template <> struct Format<std::int32_t> {
    static constexpr int bits = 0;
    static constexpr bool complex = false;
};
struct Access;
} // namespace detail

template <class T> class Buffer {
    std::shared_ptr<detail::BufferStorage> storage_;
    std::size_t count_ = 0;
    Buffer(std::shared_ptr<detail::BufferStorage> storage, std::size_t count) : storage_(std::move(storage)), count_(count) {}
    friend class Engine;
    friend class CommandBatch;
    friend struct detail::Access;

public:
    Buffer() = default;
    std::size_t size() const { return count_; }
    // Host view of the contents (downloaded on first access; written back before the buffer's next use by
    // a batch).  Mapping, upload and download are rejected until the owning submission is waited.
    T* mapped() { return static_cast<T*>(detail::mapped(storage_)); }
    const T* mapped() const { return static_cast<const T*>(detail::mapped(storage_)); }
    void upload(const T* source, std::size_t count)
    {
        if (count > count_) throw std::invalid_argument("upload exceeds buffer");
        detail::transfer(storage_, const_cast<T*>(source), count * sizeof(T), true);
    }
    void download(T* target, std::size_t count) const
    {
        if (count > count_) throw std::invalid_argument("download exceeds buffer");
        detail::transfer(storage_, target, count * sizeof(T), false);
    }
};

namespace detail {
// Type-erased buffer operand for units that encode into a batch (Linalg).
struct Operand {
    std::shared_ptr<BufferStorage> storage;
    std::size_t size = 0;
};
struct Access {
    template <class T> static Operand operand(const Buffer<T>& b) { return {b.storage_, b.size()}; }
    template <class T> static Buffer<T> buffer(Operand o) { return Buffer<T>(std::move(o.storage), o.size); }
};
struct BatchAccess;
} // namespace detail

class Submission {
    struct Impl;
    std::shared_ptr<Impl> impl_;
    explicit Submission(std::shared_ptr<Impl> impl) : impl_(std::move(impl)) {}
    friend class CommandBatch;

public:
    Submission() = default;
    Submission(const Submission&) = delete;
    Submission& operator=(const Submission&) = delete;
    Submission(Submission&&) = default;
    Submission& operator=(Submission&&) = default;
    bool ready() const;
    Timing wait(); // Repeated waits return the same timing. Destruction waits safely.
};

class CommandBatch {
    struct Impl;
    std::unique_ptr<Impl> impl_;
    explicit CommandBatch(std::unique_ptr<Impl> impl);
    void encode(int bits, bool complex, Operation op, const std::shared_ptr<detail::BufferStorage>& a,
                const std::shared_ptr<detail::BufferStorage>& b, const std::shared_ptr<detail::BufferStorage>& c,
                const std::shared_ptr<detail::BufferStorage>& out, std::size_t count, Broadcast b_index = {}, Broadcast c_index = {});
    void encode_cast(int from_bits, int to_bits, bool complex, const std::shared_ptr<detail::BufferStorage>& in,
                     const std::shared_ptr<detail::BufferStorage>& out, std::size_t count);
    void encode_tree_sum(int bits, bool complex, const std::shared_ptr<detail::BufferStorage>& input,
                         const std::shared_ptr<detail::BufferStorage>& out, std::size_t count);
    void encode_segmented_dot(int bits, bool complex, const SegmentedDot& shape, const std::shared_ptr<detail::BufferStorage>& a,
                              const std::shared_ptr<detail::BufferStorage>& b, const std::shared_ptr<detail::BufferStorage>& out);
    friend class Engine;
    friend struct detail::BatchAccess;

public:
    ~CommandBatch();
    CommandBatch(CommandBatch&&) noexcept;
    CommandBatch& operator=(CommandBatch&&) noexcept;
    template <class T> void run(Operation op, const Buffer<T>& a, const Buffer<T>& b, Buffer<T>& out)
    {
        if (operation_is_unary(op) || operation_is_ternary(op)) throw std::invalid_argument("operation requires a different number of inputs");
        if (a.size() != b.size() || a.size() != out.size()) throw std::invalid_argument("buffer counts must match");
        encode(detail::Format<T>::bits, detail::Format<T>::complex, op, a.storage_, b.storage_, a.storage_, out.storage_, a.size());
    }
    // Binary operation with b read through a broadcast index (b may be shorter than a and out).
    template <class T> void run(Operation op, const Buffer<T>& a, const Buffer<T>& b, Buffer<T>& out, Broadcast b_index)
    {
        if (operation_is_unary(op) || operation_is_ternary(op)) throw std::invalid_argument("operation requires a different number of inputs");
        if (a.size() != out.size() || b.size() < broadcast_elements(a.size(), b_index))
            throw std::invalid_argument("buffer counts do not match the broadcast");
        encode(detail::Format<T>::bits, detail::Format<T>::complex, op, a.storage_, b.storage_, a.storage_, out.storage_, a.size(), b_index);
    }
    // Fused operation with broadcast b and c.
    template <class T>
    void run(Operation op, const Buffer<T>& a, const Buffer<T>& b, const Buffer<T>& c, Buffer<T>& out, Broadcast b_index, Broadcast c_index)
    {
        if (!operation_is_ternary(op)) throw std::invalid_argument("operation does not take three inputs");
        if (a.size() != out.size() || b.size() < broadcast_elements(a.size(), b_index) || c.size() < broadcast_elements(a.size(), c_index))
            throw std::invalid_argument("buffer counts do not match the broadcast");
        encode(detail::Format<T>::bits, detail::Format<T>::complex, op, a.storage_, b.storage_, c.storage_, out.storage_, a.size(), b_index, c_index);
    }
    // fma/fms: out = a*b +/- c. Any operand may alias out.
    template <class T> void run(Operation op, const Buffer<T>& a, const Buffer<T>& b, const Buffer<T>& c, Buffer<T>& out)
    {
        if (!operation_is_ternary(op)) throw std::invalid_argument("operation does not take three inputs");
        if (a.size() != b.size() || a.size() != c.size() || a.size() != out.size()) throw std::invalid_argument("buffer counts must match");
        encode(detail::Format<T>::bits, detail::Format<T>::complex, op, a.storage_, b.storage_, c.storage_, out.storage_, a.size());
    }
    template <class T> void run(Operation op, const Buffer<T>& a, Buffer<T>& out)
    {
        if (!operation_is_unary(op)) throw std::invalid_argument("operation requires two inputs");
        if (a.size() != out.size()) throw std::invalid_argument("buffer counts must match");
        encode(detail::Format<T>::bits, detail::Format<T>::complex, op, a.storage_, a.storage_, a.storage_, out.storage_, a.size());
    }
    // Precision cast: exact when widening, round to nearest even when narrowing.
    template <class From, class To> void cast(const Buffer<From>& in, Buffer<To>& out)
    {
        static_assert(detail::Format<From>::bits && detail::Format<To>::bits, "casts convert arithmetic buffers");
        static_assert(detail::Format<From>::complex == detail::Format<To>::complex, "casts keep real or complex");
        if (in.size() != out.size()) throw std::invalid_argument("buffer counts must match");
        encode_cast(detail::Format<From>::bits, detail::Format<To>::bits, detail::Format<To>::complex, in.storage_, out.storage_, in.size());
    }
    // out (one element) = the fixed adjacent-pair tree of rounded additions of input (docs: pairs (0,1), (2,3),
    // ... level by level, an unpaired last value carried unchanged; empty input gives canonical zero).
    template <class T> void tree_sum(const Buffer<T>& input, Buffer<T>& out)
    {
        static_assert(detail::Format<T>::bits, "tree sums reduce arithmetic buffers");
        if (out.size() != 1) throw std::invalid_argument("tree_sum output must contain one element");
        encode_tree_sum(detail::Format<T>::bits, detail::Format<T>::complex, input.storage_, out.storage_, input.size());
    }
    template <class T> void segmented_dot(const SegmentedDot& shape, const Buffer<T>& a, const Buffer<T>& b, Buffer<T>& out)
    {
        static_assert(detail::Format<T>::bits, "segmented dots take arithmetic buffers");
        if (shape.length && shape.segments > std::size_t(-1) / shape.length) throw std::invalid_argument("segmented dot shape overflows");
        if (a.size() < shape.segments * shape.length || b.size() < (shape.shared_right ? shape.length : shape.segments * shape.length) ||
            out.size() < shape.segments)
            throw std::invalid_argument("buffer counts do not match the segmented dot shape");
        encode_segmented_dot(detail::Format<T>::bits, detail::Format<T>::complex, shape, a.storage_, b.storage_, out.storage_);
    }
    // Single use: the encoded operations run in order on the engine's stream.  Consecutive
    // run(mul, x, y, x) steps on the same buffers are fused into one resident chain launch (identical
    // results: every step is still rounded).
    Submission submit();
};

// One Engine per host thread. Buffers are reused.
class Engine {
public:
    explicit Engine(EngineOptions options = {});
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    std::string device_name() const;
    void prewarm(const Prewarm& request);
    std::future<void> prewarm_async(Prewarm request);
    PipelineInfo pipeline_info(int bits, Operation op);
    template <class T> Buffer<T> make_buffer(std::size_t count)
    {
        static_assert(!detail::Format<T>::bits ||
                          sizeof(T) == std::size_t(detail::Format<T>::bits / 8 + 12) * (detail::Format<T>::complex ? 2 : 1),
                      "buffer layout mismatch");
        if (count > std::size_t(-1) / sizeof(T)) throw std::invalid_argument("buffer size overflow");
        return Buffer<T>(allocate(count, detail::Format<T>::bits, detail::Format<T>::complex, sizeof(T)), count);
    }
    CommandBatch batch();
    // Bits must be a multiple of 32 in [64,1024].  b_index: broadcast read of b (b then holds
    // broadcast_elements(count, b_index) values).
    Timing run(int bits, Operation op, const void* a, const void* b, void* out, std::size_t count, Broadcast b_index = {});
    Timing run_unary(int bits, Operation op, const void* a, void* out, std::size_t count);
    Timing run_ternary(int bits, Operation op, const void* a, const void* b, const void* c, void* out, std::size_t count,
                       Broadcast b_index = {}, Broadcast c_index = {});
    // Host-array precision cast between any two supported widths (complex: componentwise).
    Timing cast(int from_bits, int to_bits, bool complex, const void* in, void* out, std::size_t count);
    // Host-array segmented dot products (SegmentedDot); out holds shape.segments values.
    Timing segmented_dot(int bits, bool complex, const SegmentedDot& shape, const void* a, const void* b, void* out);

private:
    friend class CommandBatch;
    friend struct detail::BatchAccess;
    std::shared_ptr<detail::BufferStorage> allocate(std::size_t count, int bits, bool complex, std::size_t element_bytes);
    struct Impl;
    std::shared_ptr<Impl> impl;
};
} // namespace limbforge

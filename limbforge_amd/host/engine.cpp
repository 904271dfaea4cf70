// engine.cpp — the C++17 Engine / CommandBatch / Submission / Buffer facade over liblimbforge_hip.so.
//
// Ownership and ordering follow the reference runtime (src/engine.mm:282-443, docs/execution.md):
// buffers belong to the engine that made them; a batch records its operations and retains its
// buffers; submit() claims every retained buffer atomically (an unwaited submission's buffers reject
// mapping, transfers and new submissions with std::logic_error) and issues the launches in order on the
// engine's HIP stream; Submission::wait() is idempotent, returns the device and wall time and releases
// the buffers; destroying a submission waits.
#include "engine_internal.hpp"

#include <algorithm>
#include <cstring>
#include <limits>

namespace limbforge {
namespace detail {

[[noreturn]] void raise_error(int code, const std::string& where)
{
    std::string msg = where + ": " + lf_last_error();
    if (code == LF_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (code == LF_LOGIC_ERROR) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

BufferStorage::~BufferStorage()
{
    if (device) (void)lf_free(device);
}

void BufferStorage::flush()
{
    std::lock_guard<std::mutex> lock(mutex);
    if (!host_dirty) return;
    host_dirty = false;
    // from_host without the lock (it is held here)
    if (!bits) {
        check(lf_memcpy_async(device, shadow.data(), bytes, 0, nullptr), "buffer write-back");
        check(lf_stream_synchronize(nullptr), "buffer write-back");
        return;
    }
    void* stage = nullptr;
    check(lf_malloc(&stage, bytes), "buffer write-back");
    int rc = lf_memcpy_async(stage, shadow.data(), bytes, 0, nullptr);
    lf_operand op = operand();
    if (!rc) rc = lf_aos_to_planes(bits, complex, stage, &op, count, nullptr);
    if (!rc) rc = lf_stream_synchronize(nullptr);
    lf_free(stage);
    check(rc, "buffer write-back");
}

void BufferStorage::to_host(void* host, std::size_t elements)
{
    const std::size_t n = elements * element_bytes;
    if (!n) return;
    if (!bits) {
        check(lf_memcpy_async(host, device, n, 1, nullptr), "download");
        check(lf_stream_synchronize(nullptr), "download");
        return;
    }
    void* stage = nullptr;
    check(lf_malloc(&stage, n), "download");
    lf_operand op = operand();
    int rc = lf_planes_to_aos(bits, complex, &op, stage, elements, nullptr);
    if (!rc) rc = lf_memcpy_async(host, stage, n, 1, nullptr);
    if (!rc) rc = lf_stream_synchronize(nullptr);
    lf_free(stage);
    check(rc, "download");
}

void BufferStorage::from_host(const void* host, std::size_t elements)
{
    const std::size_t n = elements * element_bytes;
    if (!n) return;
    if (!bits) {
        check(lf_memcpy_async(device, host, n, 0, nullptr), "upload");
        check(lf_stream_synchronize(nullptr), "upload");
        return;
    }
    void* stage = nullptr;
    check(lf_malloc(&stage, n), "upload");
    lf_operand op = operand();
    int rc = lf_memcpy_async(stage, host, n, 0, nullptr);
    if (!rc) rc = lf_aos_to_planes(bits, complex, stage, &op, elements, nullptr);
    if (!rc) rc = lf_stream_synchronize(nullptr);
    lf_free(stage);
    check(rc, "upload");
}

void* mapped(const std::shared_ptr<BufferStorage>& s)
{
    if (!s) throw std::invalid_argument("empty buffer handle");
    if (s->busy.load(std::memory_order_acquire)) throw std::logic_error("buffer belongs to an unwaited submission");
    std::lock_guard<std::mutex> lock(s->mutex);
    if (!s->host_valid) {
        s->shadow.resize(std::max<std::size_t>(s->bytes, 1));
        s->to_host(s->shadow.data(), s->count);
        s->host_valid = true;
    }
    s->host_dirty = true; // the caller may write through the pointer
    return s->shadow.data();
}

void transfer(const std::shared_ptr<BufferStorage>& s, void* host, std::size_t bytes, bool upload)
{
    if (!s) throw std::invalid_argument("empty buffer handle");
    if (s->busy.load(std::memory_order_acquire)) throw std::logic_error("buffer belongs to an unwaited submission");
    if (bytes > s->bytes || (!host && bytes)) throw std::invalid_argument("invalid transfer");
    if (!bytes) return;
    s->flush();
    std::lock_guard<std::mutex> lock(s->mutex);
    const std::size_t elements = bytes / s->element_bytes;
    if (upload) {
        s->from_host(host, elements);
        s->host_valid = false;
    } else
        s->to_host(host, elements);
}

} // namespace detail

namespace {
void validate(int bits, std::size_t count)
{
    if (bits < 64 || bits > 1024 || bits % 32) throw std::invalid_argument("bits must be a multiple of 32 in [64,1024]");
    if (count > std::numeric_limits<std::uint32_t>::max()) throw std::invalid_argument("batch exceeds 32-bit indexing");
}
lf_broadcast to_c(Broadcast b) { return {b.stride, b.period}; }
} // namespace

struct Engine::Impl {
    EngineOptions options;
    void* stream = nullptr;
    std::shared_ptr<int> identity = std::make_shared<int>(0);
    explicit Impl(EngineOptions o) : options(o)
    {
        unsigned n = options.threads_per_threadgroup;
        if (n && (n < 32 || n > 1024 || (n & (n - 1))))
            throw std::invalid_argument("threadgroup size must be zero or a power of two in [32,1024]");
        int devices = 0;
        lf_device_count(&devices);
        if (!devices) throw std::runtime_error("no AMD GPU available");
        detail::check(lf_stream_create(&stream), "Engine");
    }
    ~Impl()
    {
        if (stream) {
            lf_stream_synchronize(stream);
            lf_stream_destroy(stream);
        }
    }
};

struct Submission::Impl {
    std::vector<std::shared_ptr<detail::BufferStorage>> resources;
    std::vector<std::function<void()>> completions;
    std::shared_ptr<void> engine; // keeps the stream alive
    void* begin = nullptr;
    void* end = nullptr;
    std::chrono::steady_clock::time_point start;
    std::once_flag finished;
    Timing timing{0, 0};
    std::string error, host_error;
    void finish()
    {
        std::call_once(finished, [&] {
            if (end && lf_event_synchronize(end)) error = lf_last_error();
            if (error.empty()) {
                float ms = 0;
                if (!lf_event_elapsed_ms(&ms, begin, end)) timing.gpu_seconds = ms * 1e-3;
                for (auto& f : completions) {
                    try {
                        f();
                    } catch (const std::exception& e) {
                        if (host_error.empty()) host_error = e.what();
                    }
                }
            }
            completions.clear();
            timing.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
            for (auto& r : resources) r->busy.store(false, std::memory_order_release);
            if (begin) lf_event_destroy(begin);
            if (end) lf_event_destroy(end);
            begin = end = nullptr;
        });
    }
    ~Impl() { finish(); }
};

bool Submission::ready() const
{
    if (!impl_) throw std::logic_error("empty submission");
    if (!impl_->end) return true;
    int done = 0;
    if (lf_event_query(impl_->end, &done)) return true;
    return done != 0;
}

Timing Submission::wait()
{
    if (!impl_) throw std::logic_error("empty submission");
    impl_->finish();
    if (!impl_->error.empty()) throw std::runtime_error("device execution: " + impl_->error);
    if (!impl_->host_error.empty()) throw std::runtime_error(impl_->host_error);
    return impl_->timing;
}

// One recorded operation.  Element-wise steps keep their arguments so that submit() can fuse chains.
struct Step {
    enum Kind { elementwise, chain, cast, custom } kind = custom;
    int bits = 0, op = 0, steps = 0, from_bits = 0;
    bool complex = false;
    std::shared_ptr<detail::BufferStorage> a, b, c, out;
    std::size_t count = 0;
    Broadcast bi{}, ci{};
    detail::Launch launch;
};

struct CommandBatch::Impl {
    std::shared_ptr<Engine::Impl> engine;
    std::vector<Step> steps;
    std::vector<std::shared_ptr<detail::BufferStorage>> resources;
    std::vector<std::function<void()>> completions;
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();
    bool submitted = false;
    explicit Impl(std::shared_ptr<Engine::Impl> e) : engine(std::move(e)) {}
    // outputs of resident products encoded earlier in this batch (docs/numerics.md "Resident products")
    std::vector<std::pair<std::shared_ptr<detail::BufferStorage>, std::shared_ptr<std::atomic<bool>>>> provisional;
    void retain(const std::shared_ptr<detail::BufferStorage>& r, bool written)
    {
        if (!r || r->owner != engine->identity) throw std::invalid_argument("buffer belongs to another engine");
        if (r->busy.load(std::memory_order_acquire)) throw std::logic_error("buffer belongs to an unwaited submission");
        for (auto& p : provisional)
            if (p.first == r) {
                if (written) throw std::logic_error("the output of a resident product is provisional within its batch: split the batch");
                p.second->store(true);
            }
        if (std::find(resources.begin(), resources.end(), r) == resources.end()) resources.push_back(r);
    }
    // run(mul, x, y, x) repeated on the same buffers becomes one resident chain launch
    void fuse_chains()
    {
        std::vector<Step> out;
        for (auto& s : steps) {
            const bool link = s.kind == Step::elementwise && s.op == int(Operation::mul) && !s.complex && s.a == s.out;
            if (link && !out.empty()) {
                Step& p = out.back();
                const bool same = (p.kind == Step::chain || (p.kind == Step::elementwise && p.op == int(Operation::mul) && !p.complex && p.a == p.out)) &&
                                  p.a == s.a && p.b == s.b && p.bits == s.bits && p.count == s.count && p.bi.stride == s.bi.stride &&
                                  p.bi.period == s.bi.period && s.b != s.a;
                if (same) {
                    if (p.kind == Step::elementwise) {
                        p.kind = Step::chain;
                        p.steps = 1;
                    }
                    ++p.steps;
                    continue;
                }
            }
            out.push_back(std::move(s));
        }
        steps.swap(out);
    }
};

CommandBatch::CommandBatch(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
CommandBatch::~CommandBatch() = default;
CommandBatch::CommandBatch(CommandBatch&&) noexcept = default;
CommandBatch& CommandBatch::operator=(CommandBatch&&) noexcept = default;

void CommandBatch::encode(int bits, bool complex, Operation op, const std::shared_ptr<detail::BufferStorage>& a,
                          const std::shared_ptr<detail::BufferStorage>& b, const std::shared_ptr<detail::BufferStorage>& c,
                          const std::shared_ptr<detail::BufferStorage>& out, std::size_t count, Broadcast b_index, Broadcast c_index)
{
    if (!impl_ || impl_->submitted) throw std::logic_error("batch already submitted or moved");
    validate(bits, count);
    if (!operation_is_valid(op) || operation_is_complex(op) != complex) throw std::invalid_argument("operation and buffer format mismatch");
    if (op == Operation::complex_fma || op == Operation::complex_fms)
        throw std::invalid_argument("complex_fma / complex_fms are not run on the GPU by this library; use the host functions cfma / cfms");
    if (!b_index.stride || !c_index.stride) throw std::invalid_argument("broadcast stride must be positive");
    for (auto* r : {&a, &b, &c})
        if (*r && (*r)->bits != bits) throw std::invalid_argument("buffer precision does not match the operation");
    impl_->retain(a, false);
    impl_->retain(b, false);
    impl_->retain(c, false);
    impl_->retain(out, true);
    if (!count) return;
    Step s;
    s.kind = Step::elementwise;
    s.bits = bits;
    s.op = int(op);
    s.complex = complex;
    s.a = a;
    s.b = b;
    s.c = c;
    s.out = out;
    s.count = count;
    s.bi = b_index;
    s.ci = c_index;
    impl_->steps.push_back(std::move(s));
}

void CommandBatch::encode_cast(int from_bits, int to_bits, bool complex, const std::shared_ptr<detail::BufferStorage>& in,
                               const std::shared_ptr<detail::BufferStorage>& out, std::size_t count)
{
    if (!impl_ || impl_->submitted) throw std::logic_error("batch already submitted or moved");
    validate(from_bits, count);
    validate(to_bits, count);
    if (in == out && from_bits != to_bits) throw std::invalid_argument("a cast between widths cannot run in place");
    impl_->retain(in, false);
    impl_->retain(out, true);
    if (!count) return;
    Step s;
    s.kind = Step::cast;
    s.bits = to_bits;
    s.from_bits = from_bits;
    s.complex = complex;
    s.a = in;
    s.out = out;
    s.count = count;
    impl_->steps.push_back(std::move(s));
}

void CommandBatch::encode_tree_sum(int bits, bool complex, const std::shared_ptr<detail::BufferStorage>& input,
                                   const std::shared_ptr<detail::BufferStorage>& out, std::size_t count)
{
    if (!impl_ || impl_->submitted) throw std::logic_error("batch already submitted or moved");
    validate(bits, count);
    impl_->retain(input, false);
    impl_->retain(out, true);
    Step s;
    s.kind = Step::custom;
    s.launch = [bits, complex, input, out, count](void* stream) {
        const lf_operand in = input->operand(), o = out->operand();
        detail::check(lf_tree_sum(bits, complex, &in, count, &o, stream), "CommandBatch::tree_sum");
    };
    impl_->steps.push_back(std::move(s));
}

void CommandBatch::encode_segmented_dot(int bits, bool complex, const SegmentedDot& shape, const std::shared_ptr<detail::BufferStorage>& a,
                                        const std::shared_ptr<detail::BufferStorage>& b, const std::shared_ptr<detail::BufferStorage>& out)
{
    if (!impl_ || impl_->submitted) throw std::logic_error("batch already submitted or moved");
    validate(bits, shape.segments);
    impl_->retain(a, false);
    impl_->retain(b, false);
    impl_->retain(out, true);
    if (!shape.segments) return;
    Step s;
    s.kind = Step::custom;
    s.launch = [bits, complex, shape, a, b, out](void* stream) {
        const lf_operand oa = a->operand(), ob = b->operand(), o = out->operand();
        detail::check(lf_segmented_dot(bits, complex, shape.segments, shape.length, shape.shared_right, &oa, &ob, &o, stream),
                      "CommandBatch::segmented_dot");
    };
    impl_->steps.push_back(std::move(s));
}

Submission CommandBatch::submit()
{
    if (!impl_ || impl_->submitted) throw std::logic_error("batch already submitted or moved");
    auto& I = *impl_;
    auto sub = std::make_shared<Submission::Impl>();
    sub->engine = I.engine;
    sub->start = I.start;
    std::size_t claimed = 0;
    for (auto& r : I.resources) {
        bool expected = false;
        if (!r->busy.compare_exchange_strong(expected, true)) {
            for (std::size_t i = 0; i < claimed; ++i) I.resources[i]->busy.store(false);
            throw std::logic_error("buffer belongs to an unwaited submission");
        }
        ++claimed;
    }
    sub->resources = I.resources;
    try {
        for (auto& r : I.resources) r->flush();
        for (auto& s : I.steps)
            if (s.out) s.out->host_valid = false;
        I.fuse_chains();
        void* stream = I.engine->stream;
        detail::check(lf_event_create(&sub->begin), "submit");
        detail::check(lf_event_create(&sub->end), "submit");
        detail::check(lf_event_record(sub->begin, stream), "submit");
        for (auto& s : I.steps) {
            if (s.kind == Step::custom) {
                s.launch(stream);
                continue;
            }
            lf_operand oa = s.a->operand(), oo = s.out->operand();
            if (s.kind == Step::cast) {
                detail::check(lf_cast(s.from_bits, s.bits, s.complex, &oa, &oo, s.count, stream), "cast");
            } else if (s.kind == Step::chain) {
                lf_operand of = s.b->operand();
                lf_broadcast fb = to_c(s.bi);
                detail::check(lf_mul_chain(s.bits, &oa, &of, &oo, s.count, &fb, s.steps, stream), "mul chain");
            } else {
                lf_operand ob = s.b->operand(), oc = s.c->operand();
                lf_broadcast bb = to_c(s.bi), cb = to_c(s.ci);
                detail::check(lf_elementwise(s.bits, s.op, &oa, &ob, &oc, &oo, s.count, &bb, &cb, stream), "element-wise");
            }
        }
        detail::check(lf_event_record(sub->end, stream), "submit");
    } catch (...) {
        if (sub->begin && !sub->end) {
            lf_event_destroy(sub->begin);
            sub->begin = nullptr;
        }
        if (sub->end) lf_stream_synchronize(I.engine->stream);
        for (auto& r : I.resources) r->busy.store(false);
        sub->resources.clear();
        I.submitted = true;
        throw;
    }
    sub->completions = std::move(I.completions);
    I.submitted = true;
    return Submission(std::move(sub));
}

// ---- units encoding into a batch ----
namespace detail {
void BatchAccess::retain(CommandBatch& b, const std::shared_ptr<BufferStorage>& r, bool written)
{
    if (!b.impl_ || b.impl_->submitted) throw std::logic_error("batch already submitted or moved");
    b.impl_->retain(r, written);
}
void BatchAccess::enqueue(CommandBatch& b, Launch launch)
{
    if (!b.impl_ || b.impl_->submitted) throw std::logic_error("batch already submitted or moved");
    Step s;
    s.kind = Step::custom;
    s.launch = std::move(launch);
    b.impl_->steps.push_back(std::move(s));
}
void BatchAccess::on_completion(CommandBatch& b, std::function<void()> f)
{
    if (!b.impl_ || b.impl_->submitted) throw std::logic_error("batch already submitted or moved");
    b.impl_->completions.push_back(std::move(f));
}
void BatchAccess::mark_provisional(CommandBatch& b, const std::shared_ptr<BufferStorage>& r, std::shared_ptr<std::atomic<bool>> read_flag)
{
    if (!b.impl_ || b.impl_->submitted) throw std::logic_error("batch already submitted or moved");
    b.impl_->provisional.emplace_back(r, std::move(read_flag));
}
std::shared_ptr<int> BatchAccess::identity(const Engine& e) { return e.impl->identity; }
std::shared_ptr<BufferStorage> BatchAccess::allocate(Engine& e, std::size_t count, int bits)
{
    return e.allocate(count, bits, false, std::size_t(bits / 8 + 12));
}
} // namespace detail

// ---- Engine ----
Engine::Engine(EngineOptions options) : impl(std::make_shared<Impl>(options)) {}
Engine::~Engine() = default;

std::string Engine::device_name() const
{
    char buf[256] = {};
    detail::check(lf_device_name(buf, sizeof buf), "device_name");
    return buf;
}

void Engine::prewarm(const Prewarm& r)
{
    for (int bits : r.bits) {
        validate(bits, 0);
        for (auto op : r.operations)
            if (!operation_is_valid(op)) throw std::invalid_argument("unknown arithmetic operation");
    }
}

std::future<void> Engine::prewarm_async(Prewarm request)
{
    prewarm(request);
    std::promise<void> done;
    done.set_value();
    return done.get_future();
}

PipelineInfo Engine::pipeline_info(int bits, Operation op)
{
    validate(bits, 0);
    if (!operation_is_valid(op)) throw std::invalid_argument("unknown arithmetic operation");
    return {64u, 256u, 256u};
}

std::shared_ptr<detail::BufferStorage> Engine::allocate(std::size_t count, int bits, bool complex, std::size_t element_bytes)
{
    if (bits) validate(bits, count);
    auto s = std::make_shared<detail::BufferStorage>();
    s->owner = impl->identity;
    s->count = count;
    s->element_bytes = element_bytes;
    s->bytes = count * element_bytes;
    s->bits = bits;
    s->complex = complex;
    s->pitch = bits ? lf_plane_pitch(count) : count;
    const std::size_t device_bytes = bits ? std::size_t(lf_planes_bytes(bits, complex, count)) : s->bytes;
    detail::check(lf_malloc(&s->device, std::max<std::size_t>(device_bytes, 4)), "buffer allocation");
    return s;
}

CommandBatch Engine::batch() { return CommandBatch(std::make_unique<CommandBatch::Impl>(impl)); }

Timing Engine::run(int bits, Operation op, const void* a, const void* b, void* out, std::size_t count, Broadcast b_index)
{
    validate(bits, count);
    if (!operation_is_valid(op)) throw std::invalid_argument("unknown arithmetic operation");
    if (operation_is_ternary(op)) throw std::invalid_argument("operation requires three inputs");
    lf_broadcast bb = to_c(b_index);
    if (!b_index.stride) throw std::invalid_argument("broadcast stride must be positive");
    Timing t{0, 0};
    detail::check(lf_run_host(bits, int(op), a, operation_is_unary(op) ? nullptr : b, nullptr, out, count, &bb, nullptr, &t.gpu_seconds,
                              &t.wall_seconds),
                  "Engine::run");
    return t;
}

Timing Engine::run_unary(int bits, Operation op, const void* a, void* out, std::size_t count)
{
    validate(bits, count);
    if (!operation_is_unary(op)) throw std::invalid_argument("operation requires two inputs");
    Timing t{0, 0};
    detail::check(lf_run_host(bits, int(op), a, nullptr, nullptr, out, count, nullptr, nullptr, &t.gpu_seconds, &t.wall_seconds),
                  "Engine::run_unary");
    return t;
}

Timing Engine::run_ternary(int bits, Operation op, const void* a, const void* b, const void* c, void* out, std::size_t count,
                           Broadcast b_index, Broadcast c_index)
{
    validate(bits, count);
    if (!operation_is_ternary(op)) throw std::invalid_argument("operation does not take three inputs");
    if (!b_index.stride || !c_index.stride) throw std::invalid_argument("broadcast stride must be positive");
    lf_broadcast bb = to_c(b_index), cb = to_c(c_index);
    Timing t{0, 0};
    detail::check(lf_run_host(bits, int(op), a, b, c, out, count, &bb, &cb, &t.gpu_seconds, &t.wall_seconds), "Engine::run_ternary");
    return t;
}

Timing Engine::segmented_dot(int bits, bool complex, const SegmentedDot& shape, const void* a, const void* b, void* out)
{
    validate(bits, shape.segments);
    Timing t{0, 0};
    detail::check(lf_segmented_dot_host(bits, complex, shape.segments, shape.length, shape.shared_right, a, b, out, &t.gpu_seconds,
                                        &t.wall_seconds),
                  "Engine::segmented_dot");
    return t;
}

Timing Engine::cast(int from_bits, int to_bits, bool complex, const void* in, void* out, std::size_t count)
{
    validate(from_bits, count);
    validate(to_bits, count);
    Timing t{0, 0};
    detail::check(lf_cast_host(from_bits, to_bits, complex, in, out, count, &t.gpu_seconds, &t.wall_seconds), "Engine::cast");
    return t;
}

} // namespace limbforge

// host_copy.hip — pinned chunk rings and the host copy pool behind copy_to_device / copy_to_host.
#include "host_copy.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace lf {
namespace {

constexpr std::size_t CHUNK = std::size_t(8) << 20; // bytes per pinned chunk
constexpr int RING = 4;                              // pinned chunks per host thread
constexpr std::size_t PIECE = std::size_t(1) << 20;  // smallest share of one copy thread

// Worker threads that copy pieces of a host copy; the calling thread copies one piece itself and waits
// for the others.  Sized to at most 8 threads (a host copy saturates the memory bus well before that).
class CopyPool {
public:
    static CopyPool& get()
    {
        static CopyPool pool;
        return pool;
    }

    void copy(char* dst, const char* src, std::size_t n)
    {
        const std::size_t parts = std::min<std::size_t>(workers_.size() + 1, n / PIECE);
        if (parts <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        Batch b;
        b.left = int(parts - 1);
        const std::size_t share = (n / parts + 63) / 64 * 64;
        {
            std::lock_guard<std::mutex> lock(m_);
            for (std::size_t p = 1; p < parts; ++p) {
                const std::size_t off = p * share;
                tasks_.push_back({dst + off, src + off, off < n ? std::min(share, n - off) : 0, &b});
            }
        }
        cv_.notify_all();
        std::memcpy(dst, src, std::min(share, n));
        std::unique_lock<std::mutex> lock(b.m);
        b.cv.wait(lock, [&] { return b.left == 0; });
    }

private:
    struct Batch {
        std::mutex m;
        std::condition_variable cv;
        int left = 0;
    };
    struct Task {
        char* dst;
        const char* src;
        std::size_t n;
        Batch* batch;
    };

    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned want = std::min(7u, hw > 1 ? hw - 1 : 0u);
        for (unsigned i = 0; i < want; ++i) {
            try {
                workers_.emplace_back([this] { run(); });
            } catch (...) {
                break; // fewer threads: the copies still complete on the caller
            }
        }
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lock(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }

    void run()
    {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lock(m_);
                cv_.wait(lock, [&] { return stop_ || !tasks_.empty(); });
                if (stop_ && tasks_.empty()) return;
                t = tasks_.front();
                tasks_.pop_front();
            }
            if (t.n) std::memcpy(t.dst, t.src, t.n);
            std::lock_guard<std::mutex> lock(t.batch->m);
            if (--t.batch->left == 0) t.batch->cv.notify_all();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Task> tasks_;
    bool stop_ = false;
};

// The pinned chunks of one host thread and the events that mark each chunk's last DMA.
struct Ring {
    char* pinned = nullptr;
    hipEvent_t done[RING] = {};
    bool busy[RING] = {};
    ~Ring()
    {
        for (int i = 0; i < RING; ++i) {
            if (busy[i]) (void)hipEventSynchronize(done[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
        }
        if (pinned) (void)hipHostFree(pinned);
    }
    hipError_t ready()
    {
        if (!pinned)
            if (hipError_t e = hipHostMalloc((void**)&pinned, CHUNK * RING, hipHostMallocDefault)) {
                pinned = nullptr;
                return e;
            }
        for (auto& e : done)
            if (!e)
                if (hipError_t x = hipEventCreateWithFlags(&e, hipEventDisableTiming)) return x;
        return hipSuccess;
    }
    // chunk k's memory once its previous DMA has finished
    hipError_t take(int k, char** p)
    {
        if (busy[k]) {
            if (hipError_t e = hipEventSynchronize(done[k])) return e;
            busy[k] = false;
        }
        *p = pinned + std::size_t(k) * CHUNK;
        return hipSuccess;
    }
    hipError_t mark(int k, hipStream_t s)
    {
        if (hipError_t e = hipEventRecord(done[k], s)) return e;
        busy[k] = true;
        return hipSuccess;
    }
};

Ring& ring()
{
    thread_local Ring r;
    return r;
}

} // namespace

void parallel_host_copy(void* dst, const void* src, std::size_t bytes)
{
    if (bytes) CopyPool::get().copy(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
}

hipError_t copy_to_device(void* dev, const void* host, std::size_t bytes, hipStream_t s)
{
    if (!bytes) return hipSuccess;
    Ring& r = ring();
    if (hipError_t e = r.ready()) return e;
    static thread_local int next = 0;
    for (std::size_t off = 0; off < bytes; off += CHUNK) {
        const std::size_t len = std::min(CHUNK, bytes - off);
        const int k = next;
        next = (next + 1) % RING;
        char* p = nullptr;
        if (hipError_t e = r.take(k, &p)) return e;
        parallel_host_copy(p, static_cast<const char*>(host) + off, len);
        if (hipError_t e = hipMemcpyAsync(static_cast<char*>(dev) + off, p, len, hipMemcpyHostToDevice, s)) return e;
        if (hipError_t e = r.mark(k, s)) return e;
    }
    return hipSuccess;
}

hipError_t copy_to_host(void* host, const void* dev, std::size_t bytes, hipStream_t s)
{
    if (!bytes) return hipSuccess;
    Ring& r = ring();
    if (hipError_t e = r.ready()) return e;
    const std::size_t chunks = (bytes + CHUNK - 1) / CHUNK;
    // every slot's earlier DMA (an upload of this thread) has finished before a download reuses it
    for (int k = 0; k < RING; ++k) {
        char* p = nullptr;
        if (hipError_t e = r.take(k, &p)) return e;
    }
    std::size_t queued = 0;
    for (std::size_t c = 0; c < chunks; ++c) {
        while (queued < chunks && queued < c + RING) {
            const int k = int(queued % RING);
            const std::size_t off = queued * CHUNK, len = std::min(CHUNK, bytes - off);
            if (hipError_t e = hipMemcpyAsync(r.pinned + std::size_t(k) * CHUNK, static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost, s))
                return e;
            if (hipError_t e = r.mark(k, s)) return e;
            ++queued;
        }
        const int k = int(c % RING);
        char* p = nullptr;
        if (hipError_t e = r.take(k, &p)) return e;
        const std::size_t off = c * CHUNK, len = std::min(CHUNK, bytes - off);
        parallel_host_copy(static_cast<char*>(host) + off, p, len);
    }
    return hipSuccess;
}

} // namespace lf

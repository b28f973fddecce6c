// The engine's streams: a process-lifetime pool of non-blocking streams handed to threads and
// serving slots, and the fence that orders a free after every stream's earlier work.
//
// Why a pool: a stream destroyed while the device memory pool still records it on a freed block,
// or while a profiler's queue interception still refers to it, leaves the runtime holding a dead
// object. Pooled streams are never destroyed; a thread that ends returns its stream (drained) for
// the next thread.
//
// Why a fence: hipFreeAsync(p, s) orders the free after the work of stream s only. An object that
// several threads used (a sector, a program, an operator shared by the source threads of a
// dynamics walk) can be dropped on a thread whose stream never saw that work -- the release thread
// runs on the legacy null stream, which does not order against non-blocking streams -- and its
// memory would go back to the pool while kernels of another stream still read it. fence() makes
// the calling thread's stream wait for everything enqueued so far on every engine stream and on
// the null stream; the destructors of shared objects call it before their buffers are freed.
//
// Fences and graph captures exclude each other: a fence records an event on every pooled stream,
// and an event recorded on a stream that another thread is capturing becomes a node of that capture
// (the fencing stream's wait on it then joins the capture, and the free that follows would be
// recorded into the other thread's graph). A capture therefore holds the pool's lock from its
// begin to its end (capture_guard), and a fence takes the same lock.
#include <atomic>
#include <mutex>
#include <vector>

#include "hip_util.h"

namespace qsa {

namespace {

struct Registered {
    hipStream_t stream;
    hipEvent_t marker;
};

struct Pool {
    std::mutex mu;
    std::vector<Registered> all;      // every pooled stream, plus the null stream first
    std::vector<hipStream_t> idle;
    std::atomic<std::uint64_t> fences{0};
    bool null_registered = false;

    void ensure_null_locked() {
        if (null_registered) return;
        hipEvent_t e = nullptr;
        QSA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        all.push_back({nullptr, e});
        null_registered = true;
    }
};

Pool& pool() {
    static Pool* p = new Pool;   // never destroyed: streams outlive static destruction order
    return *p;
}

}  // namespace

hipStream_t pooled_stream_acquire() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    P.ensure_null_locked();
    if (!P.idle.empty()) {
        hipStream_t s = P.idle.back();
        P.idle.pop_back();
        return s;
    }
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    QSA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (err != hipSuccess) {
        P.idle.push_back(s);   // keep the stream; the fence list stays consistent
        QSA_HIP(err);
    }
    P.all.push_back({s, e});
    return s;
}

void pooled_stream_return(hipStream_t s) noexcept {
    if (!s) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    P.idle.push_back(s);
}

void pooled_stream_release(hipStream_t s) noexcept {
    if (!s) return;
    (void)hipStreamSynchronize(s);
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    P.idle.push_back(s);
}

void fence_engine_streams() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return;   // nothing can have run
    }
    Pool& P = pool();
    const hipStream_t cur = engine_stream();
    std::lock_guard<std::mutex> lock(P.mu);
    P.ensure_null_locked();
    for (const Registered& r : P.all) {
        if (r.stream == cur) continue;
        QSA_HIP(hipEventRecord(r.marker, r.stream));
        QSA_HIP(hipStreamWaitEvent(cur, r.marker, 0));
    }
    P.fences.fetch_add(1, std::memory_order_relaxed);
}

std::unique_lock<std::mutex> capture_guard() { return std::unique_lock<std::mutex>(pool().mu); }

std::uint64_t fence_count() { return pool().fences.load(std::memory_order_relaxed); }

std::size_t pooled_stream_count() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    return P.all.empty() ? 0 : P.all.size() - 1;
}

}  // namespace qsa

// Errors, device buffers and the stream the engine runs on.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

namespace qsa {

/// Status codes of the C-ABI; the Python layer maps them onto the reference's error classes
/// (include/ed/core/errors.h:24-50): InvalidRequest, ResourceLimit, DeviceUnavailable,
/// DeviceUnsupported, ConvergenceError.
enum Status : int {
    kOk = 0,
    kInvalid = 1,
    kResourceLimit = 2,
    kDeviceUnavailable = 3,
    kDeviceUnsupported = 4,
    kRuntime = 5,
    kConvergence = 6,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? kResourceLimit
                     : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? kDeviceUnavailable
                                                                               : kRuntime;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define QSA_HIP(x) ::qsa::hip_check((x), #x)

namespace detail {
inline hipStream_t& tls_stream() {
    thread_local hipStream_t s = nullptr;
    return s;
}
}  // namespace detail

/// The stream every engine launch, copy and allocation of the calling thread goes on: the null
/// stream unless a StreamScope routes the thread elsewhere (the serving loop gives each of its
/// slots a stream of its own, so requests in flight overlap on the device).
inline hipStream_t engine_stream() { return detail::tls_stream(); }

class StreamScope {
public:
    explicit StreamScope(hipStream_t s) : prev_(detail::tls_stream()) { detail::tls_stream() = s; }
    ~StreamScope() { detail::tls_stream() = prev_; }
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;

private:
    hipStream_t prev_;
};

/// Stream-ordered device memory on the engine stream (the device's default pool, kept by the
/// process instead of returned at every synchronisation). A buffer is freed in the order of the
/// freeing thread's stream: an object whose buffers other streams may still read calls
/// fence_before_release() first (every shared object's destructor does).
void device_alloc(void** p, std::size_t bytes);
void device_free(void* p) noexcept;

/// A non-blocking stream from the engine's pool (created on demand, never destroyed while the
/// process runs); release drains it and hands it back for the next thread or slot.
hipStream_t pooled_stream_acquire();
void pooled_stream_release(hipStream_t s) noexcept;
/// Hands a stream back without any HIP call (for thread-exit destructors, where a profiler's or the
/// runtime's own per-thread state may already be gone): work still queued on it stays ordered
/// before whatever the next owner enqueues.
void pooled_stream_return(hipStream_t s) noexcept;
/// The calling thread's engine stream waits for all work enqueued so far on every pooled stream
/// and on the null stream. Objects that several threads may have used call it before freeing
/// their device memory (hipFreeAsync orders a free on the freeing stream only).
void fence_engine_streams();
/// Held from hipStreamBeginCapture to hipStreamEndCapture on a pooled stream: no fence records an
/// event on any engine stream meanwhile. Only kernel launches may happen under it (no allocation,
/// no stream acquisition: the lock is the pool's own).
std::unique_lock<std::mutex> capture_guard();
/// fence_engine_streams for destructors: errors are swallowed (and cleared).
inline void fence_before_release() noexcept {
    try {
        fence_engine_streams();
    } catch (...) {
        (void)hipGetLastError();
    }
}
/// Fences taken so far (tests), and the pooled streams created.
std::uint64_t fence_count();
std::size_t pooled_stream_count();

/// Synchronous copies through the engine stream.
inline void copy_to_host(void* dst, const void* src, std::size_t bytes) {
    if (!bytes) return;
    QSA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, engine_stream()));
    QSA_HIP(hipStreamSynchronize(engine_stream()));
}
inline void copy_to_device(void* dst, const void* src, std::size_t bytes) {
    if (!bytes) return;
    QSA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, engine_stream()));
    QSA_HIP(hipStreamSynchronize(engine_stream()));
}

/// One device allocation, freed with the object.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { reset(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void reset(std::size_t bytes) {
        release();
        if (bytes == 0) return;
        device_alloc(&p_, bytes);
        n_ = bytes;
    }
    void release() noexcept {
        if (p_) device_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    template <class T = void> T* get() const { return static_cast<T*>(p_); }
    std::size_t bytes() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    void* p_ = nullptr;
    std::size_t n_ = 0;
};

/// Upload a host array into a fresh device buffer (synchronous).
template <class T> DeviceBuffer upload(const T* host, std::size_t n) {
    DeviceBuffer b(n * sizeof(T));
    copy_to_device(b.get(), host, n * sizeof(T));
    return b;
}

/// Fails with DeviceUnavailable when no HIP device is visible.
void require_device();

}  // namespace qsa

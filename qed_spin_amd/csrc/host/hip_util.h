// Errors, device buffers and the stream the engine runs on.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
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
        const hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) {
            p_ = nullptr;
            (void)hipGetLastError();
            throw Error(e == hipErrorOutOfMemory ? kResourceLimit : kRuntime,
                        "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        n_ = bytes;
    }
    void release() noexcept {
        if (p_) (void)hipFree(p_);
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
    if (n) QSA_HIP(hipMemcpy(b.get(), host, n * sizeof(T), hipMemcpyHostToDevice));
    return b;
}

/// The stream every engine launch goes on (the null stream of the current device: launches from
/// one host thread are ordered, as the reference's legacy default stream orders them).
inline hipStream_t engine_stream() { return nullptr; }

/// Fails with DeviceUnavailable when no HIP device is visible.
void require_device();

}  // namespace qsa

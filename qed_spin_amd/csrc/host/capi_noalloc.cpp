// qedh_dot_many_noalloc: the Gram-Schmidt coefficients of qedh_dot_many for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops). Its column table,
// coefficients and reduction partials live in one file-scope device array, `scratch`, of a fixed
// size; the library holds it on every device it runs on until the process ends. The kernels are
// qedh_dot_many's own, so the two entry points give the same coefficients bit for bit.
//
// Every call on one device shares that device's array and leaves the last call's contents there.
// Calls that overlap on one device (two threads, two streams) race on it. Nothing in the array
// needs to start at zero: the reduction writes every partial before the finish reads it.
#include "../../../include/qedh.h"

#include <mutex>
#include <string>

#include "backend.h"

namespace qsa {
int capi_status(const std::exception& e);
void capi_set_error(const std::string& m);
}  // namespace qsa

namespace {

constexpr std::size_t kTableBytes = QEDH_NOALLOC_MAX_BASIS * sizeof(void*);
constexpr std::size_t kCoeffBytes = QEDH_NOALLOC_MAX_BASIS * sizeof(qsa::dev::Zd);
constexpr std::size_t kOutBytes = QEDH_NOALLOC_MAX_BASIS * sizeof(double);
constexpr std::size_t kCoeffOff = kTableBytes;
constexpr std::size_t kOutOff = kCoeffOff + kCoeffBytes;
constexpr std::size_t kPartialOff = kOutOff + kOutBytes;
constexpr std::size_t kScratchBytes = kPartialOff + qsa::dev::kBlasScratchMaxBytes;
static_assert(kCoeffOff % 256 == 0 && kOutOff % 256 == 0 && kPartialOff % 256 == 0, "aligned regions");

// Host code must not use this name as an address: it is a host-side stand-in. The device address
// comes from hipGetSymbolAddress, once per device (scratch_of).
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

/// This device's copy of `scratch` (the current device's), looked up on its first use.
unsigned char* scratch_of() {
    int dev = -1;
    QSA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw qsa::Error(qsa::kDeviceUnsupported, "dot_many_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        QSA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw qsa::Error(qsa::kRuntime, "dot_many_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

__global__ void k_real_parts(const qsa::dev::Zd* c, double* out, int m) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) out[j] = c[j].re;
}

}  // namespace

extern "C" int qedh_dot_many_noalloc(int complex_vectors, const void* const* basis, size_t num_basis, const void* v,
                                     size_t n, double* coeffs_out) {
    try {
        qsa::capi_set_error("");
        if (!num_basis) return 0;
        if (!basis || !coeffs_out) throw qsa::Error(qsa::kInvalid, "dot_many_noalloc: basis and coeffs_out are required");
        if (num_basis > QEDH_NOALLOC_MAX_BASIS)
            throw qsa::Error(qsa::kInvalid, "dot_many_noalloc: num_basis " + std::to_string(num_basis) +
                                                " exceeds QEDH_NOALLOC_MAX_BASIS (" +
                                                std::to_string(QEDH_NOALLOC_MAX_BASIS) + ")");
        const int words = complex_vectors ? 2 : 1;
        if (!n) {
            for (size_t j = 0; j < num_basis * static_cast<size_t>(words); ++j) coeffs_out[j] = 0.0;
            return 0;
        }
        if (!v) throw qsa::Error(qsa::kInvalid, "dot_many_noalloc: v is null");
        for (size_t j = 0; j < num_basis; ++j)
            if (!basis[j]) throw qsa::Error(qsa::kInvalid, "dot_many_noalloc: basis vector is null");
        if (qsa::dev::blas_scratch_bytes(256) > qsa::dev::kBlasScratchMaxBytes)
            throw qsa::Error(qsa::kRuntime, "dot_many_noalloc: reduction scratch larger than its region");
        unsigned char* base = scratch_of();
        auto* table = reinterpret_cast<const void**>(base);
        auto* coeffs = reinterpret_cast<qsa::dev::Zd*>(base + kCoeffOff);
        auto* packed = reinterpret_cast<double*>(base + kOutOff);
        void* partial = base + kPartialOff;
        const hipStream_t st = qsa::engine_stream();
        QSA_HIP(hipMemcpyAsync(table, basis, num_basis * sizeof(void*), hipMemcpyHostToDevice, st));
        qsa::dev::dot_many_table(complex_vectors != 0, table, num_basis, v, n, partial, coeffs, st);
        const void* src = coeffs;
        if (!complex_vectors) {
            const int m = static_cast<int>(num_basis);
            hipLaunchKernelGGL(k_real_parts, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, st, coeffs,
                               packed, m);
            src = packed;
        }
        QSA_HIP(hipGetLastError());
        QSA_HIP(hipMemcpyAsync(coeffs_out, src, num_basis * static_cast<size_t>(words) * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        QSA_HIP(hipStreamSynchronize(st));
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

// The dense lane on the device: whole spectra of square sector blocks by rocSOLVER's Hermitian
// divide-and-conquer eigensolver (zheevd), blocks of one dimension in strided batches -- the
// reference's DenseBatch lane (cuSOLVER Xsyevd in batches, src/gpu/little_group.cu:38-211,
// include/ed/gpu/little_group.h:28-69), which place() picks for dense blocks of at least
// kDeviceDenseMinDim = 1024 states under device='auto' (include/ed/core/device.h).
#include "dense_device.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <map>
#include <string>

#include "../device/launch.h"

namespace qsa {

namespace {

void rb_check(rocblas_status s, const char* what) {
    if (s == rocblas_status_success) return;
    if (s == rocblas_status_memory_error)
        throw Error(kResourceLimit, std::string("dense device lane: ") + what + ": device memory exhausted");
    throw Error(kRuntime, std::string("dense device lane: ") + what + " failed (rocblas status " +
                              std::to_string(static_cast<int>(s)) + ")");
}

struct Handle {
    rocblas_handle h = nullptr;
    explicit Handle(hipStream_t st) {
        rb_check(rocblas_create_handle(&h), "rocblas_create_handle");
        rb_check(rocblas_set_stream(h, st), "rocblas_set_stream");
    }
    ~Handle() {
        if (h) (void)rocblas_destroy_handle(h);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

}  // namespace

std::vector<std::vector<double>> device_spectra(const std::vector<const SectorOperator*>& ops,
                                                std::uint64_t batch_bytes) {
    std::vector<std::vector<double>> out(ops.size());
    std::map<std::int64_t, std::vector<std::size_t>> by_dim;
    for (std::size_t b = 0; b < ops.size(); ++b) {
        const std::int64_t n = ops[b]->dim();
        if (!ops[b]->same()) throw Error(kInvalid, "dense device lane: a block must be square");
        if (ops[b]->lane() != kLaneCsr) throw Error(kDeviceUnsupported, "dense device lane: a block needs its CSR");
        if (n > kDenseDeviceMaxDim)
            throw Error(kInvalid, "dense device lane: block above " + std::to_string(kDenseDeviceMaxDim) + " states");
        if (n > 0) by_dim[n].push_back(b);
    }
    if (by_dim.empty()) return out;
    const hipStream_t st = engine_stream();
    Handle handle(st);
    for (const auto& [n, idx] : by_dim) {
        const std::size_t nn = static_cast<std::size_t>(n) * static_cast<std::size_t>(n);
        const std::size_t per = std::max<std::size_t>(1, batch_bytes / (nn * sizeof(dev::Zd)));
        for (std::size_t i0 = 0; i0 < idx.size(); i0 += per) {
            const std::size_t cnt = std::min(per, idx.size() - i0);
            DeviceBuffer A(cnt * nn * sizeof(dev::Zd)), D(cnt * static_cast<std::size_t>(n) * 8),
                E(cnt * static_cast<std::size_t>(n) * 8), info(cnt * sizeof(rocblas_int));
            for (std::size_t q = 0; q < cnt; ++q)
                dev::csr_to_dense(ops[idx[i0 + q]]->device_csr(), A.get<dev::Zd>() + q * nn, st);
            QSA_HIP(hipGetLastError());
            rb_check(rocsolver_zheevd_strided_batched(
                         handle.h, rocblas_evect_none, rocblas_fill_lower, static_cast<rocblas_int>(n),
                         reinterpret_cast<rocblas_double_complex*>(A.get()), static_cast<rocblas_int>(n),
                         static_cast<rocblas_stride>(nn), D.get<double>(), static_cast<rocblas_stride>(n),
                         E.get<double>(), static_cast<rocblas_stride>(n), info.get<rocblas_int>(),
                         static_cast<rocblas_int>(cnt)),
                     "rocsolver_zheevd_strided_batched");
            std::vector<rocblas_int> ih(cnt);
            std::vector<double> dh(cnt * static_cast<std::size_t>(n));
            copy_to_host(ih.data(), info.get(), cnt * sizeof(rocblas_int));
            copy_to_host(dh.data(), D.get(), dh.size() * 8);
            for (std::size_t q = 0; q < cnt; ++q) {
                if (ih[q] != 0)
                    throw Error(kConvergence, "dense device lane: zheevd did not converge (info " +
                                                  std::to_string(ih[q]) + ")");
                out[idx[i0 + q]].assign(dh.begin() + static_cast<std::ptrdiff_t>(q * n),
                                        dh.begin() + static_cast<std::ptrdiff_t>((q + 1) * n));
            }
        }
    }
    return out;
}

}  // namespace qsa

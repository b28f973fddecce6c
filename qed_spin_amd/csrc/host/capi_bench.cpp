// qedh_matvec_benchmark: the apply timed with HIP events on the engine stream.
#include "../../../include/qedh.h"

#include <cstring>
#include <string>
#include <vector>

#include "../device/launch.h"
#include "sector_operator.h"

struct qedh_matvec_s {
    std::shared_ptr<qsa::SectorOperator> op;
};

namespace qsa {
int capi_status(const std::exception& e);
void capi_set_error(const std::string& m);
}  // namespace qsa

extern "C" int qedh_matvec_benchmark(qedh_matvec m, int k, int block, int complex_vectors, int warmup, int iters,
                                     double* avg_seconds) {
    try {
        if (!m || !avg_seconds || k < 1 || k > 8 || iters < 1 || warmup < 0)
            throw qsa::Error(qsa::kInvalid, "qedh_matvec_benchmark: bad arguments");
        const auto& op = *m->op;
        const bool cv = complex_vectors != 0 || !op.is_real();
        const std::size_t elem = cv ? 16 : 8;
        const std::size_t rows = static_cast<std::size_t>(op.dim()), cols = static_cast<std::size_t>(op.col_dim());
        const hipStream_t st = qsa::engine_stream();
        std::vector<qsa::DeviceBuffer> xs, ys;
        qsa::DeviceBuffer X, Y, part(qsa::dev::blas_scratch_bytes(1)), nrm(8);
        std::vector<const void*> ins;
        std::vector<void*> outs;
        // random unit input vectors (Gaussian entries, normalised on the device)
        auto random_unit = [&](void* x, std::size_t n, std::uint64_t seed) {
            qsa::dev::fill_gaussian(cv, x, n, seed, st);
            qsa::dev::nrm2(cv, x, n, part.get(), nrm.get<double>(), st);
            double h = 1.0;
            QSA_HIP(hipMemcpyAsync(&h, nrm.get(), 8, hipMemcpyDeviceToHost, st));
            QSA_HIP(hipStreamSynchronize(st));
            if (h > 0.0) qsa::dev::scale(cv, qsa::dev::Zd{1.0 / h, 0.0}, x, n, st);
        };
        if (block) {
            X.reset(cols * k * elem);
            Y.reset(rows * k * elem);
            random_unit(X.get(), cols * k, 0x51ED2701ULL);
        } else {
            for (int i = 0; i < k; ++i) {
                xs.emplace_back(cols * elem);
                ys.emplace_back(rows * elem);
                random_unit(xs.back().get(), cols, 0x51ED2701ULL + static_cast<std::uint64_t>(i));
                ins.push_back(xs.back().get());
                outs.push_back(ys.back().get());
            }
        }
        auto once = [&] {
            if (block) op.apply_block(k, X.get(), Y.get(), cv);
            else op.apply_multi(ins.data(), outs.data(), k, cv);
        };
        for (int i = 0; i < warmup; ++i) once();
        hipEvent_t a, b;
        QSA_HIP(hipEventCreate(&a));
        QSA_HIP(hipEventCreate(&b));
        QSA_HIP(hipEventRecord(a, st));
        for (int i = 0; i < iters; ++i) once();
        QSA_HIP(hipEventRecord(b, st));
        QSA_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        QSA_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *avg_seconds = static_cast<double>(ms) * 1e-3 / iters;
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

extern "C" int qedh_matvec_export_csr_rows(qedh_matvec m, int64_t r0, int64_t r1, uint64_t* row_ptr, uint32_t* col,
                                           double* val) {
    try {
        if (!m) throw qsa::Error(qsa::kInvalid, "matvec is null");
        std::vector<std::uint64_t> rp;
        std::vector<std::uint32_t> c;
        std::vector<double> v;
        m->op->csr_download_rows(r0, r1, rp, c, v);
        if (row_ptr) std::memcpy(row_ptr, rp.data(), rp.size() * 8);
        if (col && !c.empty()) std::memcpy(col, c.data(), c.size() * 4);
        if (val && !v.empty()) std::memcpy(val, v.data(), v.size() * 8);
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

extern "C" int qedh_matvec_set_tuning(qedh_matvec m, int lanes, int flags) {
    try {
        if (!m) throw qsa::Error(qsa::kInvalid, "matvec is null");
        if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
            throw qsa::Error(qsa::kInvalid, "lanes must be 0 (keep), 4, 8, 16, 32 or 64");
        if (lanes) m->op->set_lanes(lanes);
        m->op->set_nontemporal((flags & QEDH_TUNE_NONTEMPORAL) != 0);
        m->op->set_xcd_rows((flags & QEDH_TUNE_XCD_ROWS) != 0);
        m->op->set_unroll((flags & QEDH_TUNE_UNROLL16) ? 16 : (flags & QEDH_TUNE_UNROLL8) ? 8 : (flags & QEDH_TUNE_UNROLL) ? 4
                          : (flags & QEDH_TUNE_UNROLL2) ? 2 : 1);
        if ((flags & QEDH_TUNE_SELL) || m->op->sell()) m->op->build_sell((flags & QEDH_TUNE_SELL_DIAG) != 0);
        m->op->set_sell_sums((flags & QEDH_TUNE_SELL_SUMS1) ? 1 : 4);
        m->op->set_use_sell((flags & QEDH_TUNE_CSR) == 0);
        m->op->set_workgroup((flags & QEDH_TUNE_WG1024) ? 1024 : (flags & QEDH_TUNE_WG256) ? 256 : 512);
        m->op->set_global_dict((flags & QEDH_TUNE_GLOBAL_DICT) != 0);
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

extern "C" int qedh_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    try {
        if (!bytes) return 0;
        QSA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, qsa::engine_stream()));
        QSA_HIP(hipStreamSynchronize(qsa::engine_stream()));
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

// The BLAS-1/3 backend of the reference's Krylov drivers behind the C-ABI: BasicBackend<Scalar>
// (include/ed/matvec/backend.h:56-179) and its device implementation CudaBackend
// (include/ed/gpu/cuda_backend.cuh:104-415) on device pointers. Every entry point runs on the calling
// thread's engine stream and returns when its work is done (the backend is synchronous from the
// caller's view, as the reference's is); reductions are the deterministic device reductions the
// engine's own drivers use, with x^H y conjugation on the left.
#include "../../../include/qedh.h"

#include <rocblas/rocblas.h>

#include <complex>
#include <mutex>
#include <string>
#include <vector>

#include "backend.h"

namespace qsa {
int capi_status(const std::exception& e);
void capi_set_error(const std::string& m);
}  // namespace qsa

namespace {

template <class F> int guarded_be(F&& f) {
    try {
        qsa::capi_set_error("");
        f();
        return 0;
    } catch (const qsa::Error& e) {
        qsa::capi_set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc& e) {
        qsa::capi_set_error(std::string("host allocation failed: ") + e.what());
        return qsa::kResourceLimit;
    } catch (const std::exception& e) {
        return qsa::capi_status(e);
    }
}

void need(const void* p, const char* what) {
    if (!p) throw qsa::Error(qsa::kInvalid, std::string(what) + " is null");
}

qsa::dev::Zd scalar_of(const double* a, int cplx) {
    need(a, "scalar");
    return qsa::dev::Zd{a[0], cplx ? a[1] : 0.0};
}

/// Per-thread device objects (reduction scratch, a rocBLAS handle) come from process-wide pools and
/// go back to them when the thread ends; nothing is freed or destroyed at thread exit, where a
/// profiler's or the runtime's own per-thread state may already be gone.
template <class T> struct ThreadPool {
    std::mutex mu;
    std::vector<T*> idle;
    T* take() {
        std::lock_guard<std::mutex> lock(mu);
        if (idle.empty()) return nullptr;
        T* t = idle.back();
        idle.pop_back();
        return t;
    }
    void give(T* t) noexcept {
        if (!t) return;
        std::lock_guard<std::mutex> lock(mu);
        idle.push_back(t);
    }
};
template <class T> ThreadPool<T>& pool_of() {
    static ThreadPool<T>* p = new ThreadPool<T>;   // never destroyed
    return *p;
}
template <class T> struct Lease {
    T* t = nullptr;
    ~Lease() { pool_of<T>().give(t); }
};

/// Reduction scratch and result slots of the calling thread (its engine stream only).
struct ThreadScratch {
    qsa::DeviceBuffer partial, result;
    ThreadScratch() : partial(qsa::dev::blas_scratch_bytes(256)), result(64) {}
};
ThreadScratch& scratch() {
    thread_local Lease<ThreadScratch> lease;
    if (!lease.t) {
        lease.t = pool_of<ThreadScratch>().take();
        if (!lease.t) lease.t = new ThreadScratch;
    }
    return *lease.t;
}

void sync() { QSA_HIP(hipStreamSynchronize(qsa::engine_stream())); }

void read_result(double* out, int words) {
    QSA_HIP(hipMemcpyAsync(out, scratch().result.get(), static_cast<std::size_t>(words) * 8, hipMemcpyDeviceToHost,
                           qsa::engine_stream()));
    sync();
}

void rb(rocblas_status s, const char* what) {
    if (s == rocblas_status_success) return;
    throw qsa::Error(s == rocblas_status_memory_error ? qsa::kResourceLimit : qsa::kRuntime,
                     std::string("qedh_gemm: ") + what + " failed (rocblas status " + std::to_string(static_cast<int>(s)) +
                         ")");
}

rocblas_operation op_of(char c, int cplx) {
    switch (c) {
    case 'N': case 'n': return rocblas_operation_none;
    case 'T': case 't': return rocblas_operation_transpose;
    case 'C': case 'c': return cplx ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
    default: throw qsa::Error(qsa::kInvalid, std::string("qedh_gemm: op must be N, T or C, got '") + c + "'");
    }
}

/// The calling thread's rocBLAS handle (from the pool on first use, back to it when the thread ends).
struct BlasHandle {
    rocblas_handle h = nullptr;
};
rocblas_handle blas_handle() {
    thread_local Lease<BlasHandle> lease;
    if (!lease.t) {
        lease.t = pool_of<BlasHandle>().take();
        if (!lease.t) {
            auto* fresh = new BlasHandle;
            const rocblas_status s = rocblas_create_handle(&fresh->h);
            if (s != rocblas_status_success) {
                delete fresh;
                rb(s, "rocblas_create_handle");
            }
            lease.t = fresh;
        }
    }
    rb(rocblas_set_stream(lease.t->h, qsa::engine_stream()), "rocblas_set_stream");
    return lease.t->h;
}

}  // namespace

extern "C" {

int qedh_fill_zero(int complex_vectors, void* x, size_t n) {
    return guarded_be([&] {
        if (!n) return;
        need(x, "x");
        qsa::dev::fill_zero(complex_vectors != 0, x, n, qsa::engine_stream());
        sync();
    });
}

int qedh_axpy(int complex_vectors, const double* alpha, const void* x, void* y, size_t n) {
    return guarded_be([&] {
        if (!n) return;
        need(x, "x");
        need(y, "y");
        qsa::dev::axpy(complex_vectors != 0, scalar_of(alpha, complex_vectors), x, y, n, qsa::engine_stream());
        sync();
    });
}

int qedh_scale(int complex_vectors, const double* alpha, void* x, size_t n) {
    return guarded_be([&] {
        if (!n) return;
        need(x, "x");
        qsa::dev::scale(complex_vectors != 0, scalar_of(alpha, complex_vectors), x, n, qsa::engine_stream());
        sync();
    });
}

int qedh_axpby(int complex_vectors, const double* alpha, const void* x, const double* beta, void* y, size_t n) {
    return guarded_be([&] {
        if (!n) return;
        need(x, "x");
        need(y, "y");
        qsa::dev::axpby(complex_vectors != 0, scalar_of(alpha, complex_vectors), x, scalar_of(beta, complex_vectors), y,
                        n, qsa::engine_stream());
        sync();
    });
}

int qedh_dot(int complex_vectors, const void* x, const void* y, size_t n, double* out) {
    return guarded_be([&] {
        need(out, "out");
        out[0] = 0.0;
        if (complex_vectors) out[1] = 0.0;
        if (!n) return;
        need(x, "x");
        need(y, "y");
        auto& S = scratch();
        qsa::dev::dot(complex_vectors != 0, x, y, n, S.partial.get(), S.result.get<qsa::dev::Zd>(), qsa::engine_stream());
        double r[2];
        read_result(r, 2);
        out[0] = r[0];
        if (complex_vectors) out[1] = r[1];
    });
}

int qedh_nrm2(int complex_vectors, const void* x, size_t n, double* out) {
    return guarded_be([&] {
        need(out, "out");
        *out = 0.0;
        if (!n) return;
        need(x, "x");
        auto& S = scratch();
        qsa::dev::nrm2(complex_vectors != 0, x, n, S.partial.get(), S.result.get<double>(), qsa::engine_stream());
        read_result(out, 1);
    });
}

int qedh_axpy_dot(int complex_vectors, const double* alpha, const void* x, void* y, const void* z, size_t n,
                  double* out) {
    return guarded_be([&] {
        need(out, "out");
        out[0] = 0.0;
        if (complex_vectors) out[1] = 0.0;
        if (!n) return;
        need(x, "x");
        need(y, "y");
        need(z, "z");
        auto& S = scratch();
        qsa::dev::axpy_dot(complex_vectors != 0, scalar_of(alpha, complex_vectors), x, y, z, n, S.partial.get(),
                           S.result.get<qsa::dev::Zd>(), qsa::engine_stream());
        double r[2];
        read_result(r, 2);
        out[0] = r[0];
        if (complex_vectors) out[1] = r[1];
    });
}

int qedh_axpy_nrm2(int complex_vectors, const double* alpha, const void* x, void* y, size_t n, double* out) {
    return guarded_be([&] {
        need(out, "out");
        *out = 0.0;
        if (!n) return;
        need(x, "x");
        need(y, "y");
        auto& S = scratch();
        qsa::dev::axpy_nrm2(complex_vectors != 0, scalar_of(alpha, complex_vectors), x, y, n, S.partial.get(),
                            S.result.get<double>(), qsa::engine_stream());
        read_result(out, 1);
    });
}

int qedh_dot_many(int complex_vectors, const void* const* basis, size_t num_basis, const void* v, size_t n,
                  double* coeffs_out) {
    return guarded_be([&] {
        if (!num_basis) return;
        need(basis, "basis");
        need(coeffs_out, "coeffs_out");
        const int words = complex_vectors ? 2 : 1;
        if (!n) {
            for (size_t j = 0; j < num_basis * static_cast<size_t>(words); ++j) coeffs_out[j] = 0.0;
            return;
        }
        need(v, "v");
        for (size_t j = 0; j < num_basis; ++j) need(basis[j], "basis vector");
        const hipStream_t st = qsa::engine_stream();
        qsa::DeviceBuffer table(num_basis * sizeof(void*)), coeffs(num_basis * sizeof(qsa::dev::Zd));
        QSA_HIP(hipMemcpyAsync(table.get(), basis, table.bytes(), hipMemcpyHostToDevice, st));
        qsa::dev::dot_many_table(complex_vectors != 0, table.get<const void* const>(), num_basis, v, n,
                                 scratch().partial.get(), coeffs.get<qsa::dev::Zd>(), st);
        std::vector<qsa::dev::Zd> h(num_basis);
        QSA_HIP(hipMemcpyAsync(h.data(), coeffs.get(), coeffs.bytes(), hipMemcpyDeviceToHost, st));
        sync();
        for (size_t j = 0; j < num_basis; ++j) {
            coeffs_out[j * words] = h[j].re;
            if (complex_vectors) coeffs_out[j * 2 + 1] = h[j].im;
        }
    });
}

int qedh_axpy_many(int complex_vectors, const double* alphas, const void* const* basis, size_t num_basis, void* v,
                   size_t n) {
    return guarded_be([&] {
        if (!num_basis || !n) return;
        need(basis, "basis");
        need(alphas, "alphas");
        need(v, "v");
        for (size_t j = 0; j < num_basis; ++j) need(basis[j], "basis vector");
        const hipStream_t st = qsa::engine_stream();
        std::vector<qsa::dev::Zd> a(num_basis);
        for (size_t j = 0; j < num_basis; ++j)
            a[j] = complex_vectors ? qsa::dev::Zd{alphas[2 * j], alphas[2 * j + 1]} : qsa::dev::Zd{alphas[j], 0.0};
        qsa::DeviceBuffer table(num_basis * sizeof(void*)), coeffs(num_basis * sizeof(qsa::dev::Zd));
        QSA_HIP(hipMemcpyAsync(table.get(), basis, table.bytes(), hipMemcpyHostToDevice, st));
        QSA_HIP(hipMemcpyAsync(coeffs.get(), a.data(), coeffs.bytes(), hipMemcpyHostToDevice, st));
        qsa::dev::axpy_many_table(complex_vectors != 0, coeffs.get<qsa::dev::Zd>(), table.get<const void* const>(),
                                  num_basis, v, n, st);
        sync();
    });
}

int qedh_gemm(int complex_vectors, char op_a, char op_b, size_t m, size_t n, size_t k, const double* alpha,
              const void* A, size_t lda, const void* B, size_t ldb, const double* beta, void* Cm, size_t ldc) {
    return guarded_be([&] {
        const int cv = complex_vectors ? 1 : 0;
        const rocblas_operation oa = op_of(op_a, cv), ob = op_of(op_b, cv);
        if (!m || !n) return;
        need(Cm, "C");
        if (k) {
            need(A, "A");
            need(B, "B");
        }
        const size_t rows_a = (oa == rocblas_operation_none) ? m : k;
        const size_t rows_b = (ob == rocblas_operation_none) ? k : n;
        if (lda < (rows_a ? rows_a : 1) || ldb < (rows_b ? rows_b : 1) || ldc < m)
            throw qsa::Error(qsa::kInvalid, "qedh_gemm: a leading dimension is smaller than its matrix's rows");
        const size_t lim = static_cast<size_t>(0x7FFFFFFF);
        if (m > lim || n > lim || k > lim || lda > lim || ldb > lim || ldc > lim)
            throw qsa::Error(qsa::kInvalid, "qedh_gemm: a size exceeds the library's 32-bit range");
        rocblas_handle h = blas_handle();
        const auto I = [](size_t x) { return static_cast<rocblas_int>(x); };
        if (cv) {
            const rocblas_double_complex a{alpha ? alpha[0] : 1.0, alpha ? alpha[1] : 0.0};
            const rocblas_double_complex b{beta ? beta[0] : 0.0, beta ? beta[1] : 0.0};
            rb(rocblas_zgemm(h, oa, ob, I(m), I(n), I(k), &a, static_cast<const rocblas_double_complex*>(A), I(lda),
                             static_cast<const rocblas_double_complex*>(B), I(ldb), &b,
                             static_cast<rocblas_double_complex*>(Cm), I(ldc)),
               "rocblas_zgemm");
        } else {
            const double a = alpha ? alpha[0] : 1.0, b = beta ? beta[0] : 0.0;
            rb(rocblas_dgemm(h, oa, ob, I(m), I(n), I(k), &a, static_cast<const double*>(A), I(lda),
                             static_cast<const double*>(B), I(ldb), &b, static_cast<double*>(Cm), I(ldc)),
               "rocblas_dgemm");
        }
        sync();
    });
}

int qedh_stream_stats(uint64_t* fences, int64_t* pooled_streams) {
    return guarded_be([&] {
        if (fences) *fences = qsa::fence_count();
        if (pooled_streams) *pooled_streams = static_cast<int64_t>(qsa::pooled_stream_count());
    });
}

}  // extern "C"

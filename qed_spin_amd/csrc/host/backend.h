// The device backend of the Krylov drivers: the reference's BasicBackend<Scalar> contract
// (include/ed/matvec/backend.h:56-179) on HIP. Vectors live in HBM; every reduction is a
// deterministic device reduction whose scalar comes back to the host when the driver asks for it
// (one small copy), as cuBLAS host-pointer mode does in the reference's CudaBackend.
#pragma once

#include <complex>
#include <memory>
#include <type_traits>
#include <vector>

#include "../device/launch.h"
#include "hip_util.h"

namespace qsa {

template <class Scalar> class HipBackend {
    static_assert(std::is_same_v<Scalar, double> || std::is_same_v<Scalar, std::complex<double>>,
                  "HipBackend: Scalar is double or std::complex<double>");

public:
    using scalar_type = Scalar;
    static constexpr bool kComplex = std::is_same_v<Scalar, std::complex<double>>;

    HipBackend() : scratch_(dev::blas_scratch_bytes(256)), result_(64 * 16) {}

    struct Deleter {
        void operator()(Scalar* p) const noexcept { device_free(p); }
    };
    using UniqueVec = std::unique_ptr<Scalar, Deleter>;

    Scalar* allocate(std::size_t n) const {
        void* p = nullptr;
        if (n == 0) return nullptr;
        device_alloc(&p, n * sizeof(Scalar));
        return static_cast<Scalar*>(p);
    }
    UniqueVec make_zero_vector(std::size_t n) const {
        UniqueVec v(allocate(n));
        fill_zero(v.get(), n);
        return v;
    }
    void fill_zero(Scalar* p, std::size_t n) const { dev::fill_zero(kComplex, p, n, st()); }
    void copy(const Scalar* src, Scalar* dst, std::size_t n) const {
        if (n) QSA_HIP(hipMemcpyAsync(dst, src, n * sizeof(Scalar), hipMemcpyDeviceToDevice, st()));
    }
    void copy_from_host(const Scalar* h, Scalar* d, std::size_t n) const {
        if (n) QSA_HIP(hipMemcpyAsync(d, h, n * sizeof(Scalar), hipMemcpyHostToDevice, st()));
        QSA_HIP(hipStreamSynchronize(st()));
    }
    void copy_to_host(const Scalar* d, Scalar* h, std::size_t n) const {
        if (n) QSA_HIP(hipMemcpyAsync(h, d, n * sizeof(Scalar), hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
    }

    void axpy(Scalar a, const Scalar* x, Scalar* y, std::size_t n) const { dev::axpy(kComplex, z(a), x, y, n, st()); }
    void scale(Scalar a, Scalar* x, std::size_t n) const { dev::scale(kComplex, z(a), x, n, st()); }
    void axpby(Scalar a, const Scalar* x, Scalar b, Scalar* y, std::size_t n) const {
        dev::axpby(kComplex, z(a), x, z(b), y, n, st());
    }
    /// x^H y
    Scalar dot(const Scalar* x, const Scalar* y, std::size_t n) const {
        dev::dot(kComplex, x, y, n, scratch_.get(), result_.get<dev::Zd>(), st());
        return read_scalar();
    }
    double nrm2(const Scalar* x, std::size_t n) const {
        dev::nrm2(kComplex, x, n, scratch_.get(), result_.get<double>(), st());
        return read_double(0);
    }
    /// y += a x; returns z^H y
    Scalar axpy_dot(Scalar a, const Scalar* x, Scalar* y, const Scalar* zz, std::size_t n) const {
        dev::axpy_dot(kComplex, z(a), x, y, zz, n, scratch_.get(), result_.get<dev::Zd>(), st());
        return read_scalar();
    }
    /// y += a x; returns ||y||
    double axpy_nrm2(Scalar a, const Scalar* x, Scalar* y, std::size_t n) const {
        dev::axpy_nrm2(kComplex, z(a), x, y, n, scratch_.get(), result_.get<double>(), st());
        return read_double(0);
    }

    /// One classical Gram-Schmidt pass against the m columns of the contiguous block V (column j
    /// at V + j ld): w -= V (V^H w). Returns (||w|| before, ||w|| after). The coefficients stay on
    /// the device (the host never needs them).
    std::pair<double, double> cgs_pass(const Scalar* V, std::size_t ld, std::size_t m, Scalar* w, std::size_t n) const {
        if (coeff_.bytes() < m * sizeof(dev::Zd)) coeff_.reset(std::max<std::size_t>(m, 64) * 2 * sizeof(dev::Zd));
        double* norms = result_.get<double>() + 4;
        dev::dot_many(kComplex, V, ld, m, w, n, scratch_.get(), coeff_.get<dev::Zd>(), norms, st());
        dev::axpy_many(kComplex, coeff_.get<dev::Zd>(), V, ld, m, w, n, scratch_.get(), norms + 1, st(), -1.0);
        double h[2];
        QSA_HIP(hipMemcpyAsync(h, norms, 16, hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
        return {std::sqrt(h[0]), std::sqrt(h[1])};
    }

    /// Two unconditional classical Gram-Schmidt passes of w against the m >= 1 contiguous columns
    /// of V. Returns the coefficient of the last column summed over both passes (V_{m-1}^H w as w
    /// was) and ||w|| after, read back together in one synchronisation.
    std::pair<Scalar, double> cgs2_tail(const Scalar* V, std::size_t ld, std::size_t m, Scalar* w, std::size_t n) const {
        if (coeff_.bytes() < 2 * m * sizeof(dev::Zd)) coeff_.reset(std::max<std::size_t>(m, 64) * 2 * sizeof(dev::Zd));
        dev::Zd* c1 = coeff_.get<dev::Zd>();
        dev::Zd* c2 = c1 + m;
        double* norm = result_.get<double>() + 8;
        dev::dot_many(kComplex, V, ld, m, w, n, scratch_.get(), c1, nullptr, st());
        dev::axpy_many(kComplex, c1, V, ld, m, w, n, scratch_.get(), nullptr, st(), -1.0);
        dev::dot_many(kComplex, V, ld, m, w, n, scratch_.get(), c2, nullptr, st());
        dev::axpy_many(kComplex, c2, V, ld, m, w, n, scratch_.get(), norm, st(), -1.0);
        dev::Zd tail[2];
        double nn = 0.0;
        QSA_HIP(hipMemcpyAsync(&tail[0], c1 + (m - 1), sizeof(dev::Zd), hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipMemcpyAsync(&tail[1], c2 + (m - 1), sizeof(dev::Zd), hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipMemcpyAsync(&nn, norm, sizeof(double), hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
        return {from_z(dev::Zd{tail[0].re + tail[1].re, tail[0].im + tail[1].im}), std::sqrt(nn)};
    }

    /// coeffs_out[j] = V_j^H v for the m columns of V (host result).
    void dot_many_block(const Scalar* V, std::size_t ld, std::size_t m, const Scalar* v, std::size_t n,
                        Scalar* coeffs_out) const {
        if (coeff_.bytes() < m * sizeof(dev::Zd)) coeff_.reset(std::max<std::size_t>(m, 64) * 2 * sizeof(dev::Zd));
        dev::dot_many(kComplex, V, ld, m, v, n, scratch_.get(), coeff_.get<dev::Zd>(), nullptr, st());
        std::vector<dev::Zd> h(m);
        QSA_HIP(hipMemcpyAsync(h.data(), coeff_.get(), m * sizeof(dev::Zd), hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
        for (std::size_t j = 0; j < m; ++j) coeffs_out[j] = from_z(h[j]);
    }
    /// v += sum_j coeffs[j] V_j (host coefficients).
    void axpy_many_block(const Scalar* coeffs, const Scalar* V, std::size_t ld, std::size_t m, Scalar* v,
                         std::size_t n) const {
        if (coeff_.bytes() < m * sizeof(dev::Zd)) coeff_.reset(std::max<std::size_t>(m, 64) * 2 * sizeof(dev::Zd));
        std::vector<dev::Zd> h(m);
        for (std::size_t j = 0; j < m; ++j) h[j] = z(coeffs[j]);
        QSA_HIP(hipMemcpyAsync(coeff_.get(), h.data(), m * sizeof(dev::Zd), hipMemcpyHostToDevice, st()));
        dev::axpy_many(kComplex, coeff_.get<dev::Zd>(), V, ld, m, v, n, scratch_.get(), nullptr, st());
        QSA_HIP(hipStreamSynchronize(st()));
    }

    void synchronize() const { QSA_HIP(hipStreamSynchronize(st())); }

private:
    static hipStream_t st() { return engine_stream(); }
    static dev::Zd z(Scalar a) {
        if constexpr (kComplex) return dev::Zd{a.real(), a.imag()};
        else return dev::Zd{a, 0.0};
    }
    static Scalar from_z(dev::Zd v) {
        if constexpr (kComplex) return Scalar(v.re, v.im);
        else return v.re;
    }
    Scalar read_scalar() const {
        dev::Zd h;
        QSA_HIP(hipMemcpyAsync(&h, result_.get(), sizeof h, hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
        return from_z(h);
    }
    double read_double(int slot) const {
        double h;
        QSA_HIP(hipMemcpyAsync(&h, result_.get<double>() + slot, sizeof h, hipMemcpyDeviceToHost, st()));
        QSA_HIP(hipStreamSynchronize(st()));
        return h;
    }
    DeviceBuffer scratch_, result_;
    mutable DeviceBuffer coeff_;
};

}  // namespace qsa

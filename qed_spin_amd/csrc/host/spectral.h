// Sampled thermodynamics (FTLM) and the T = 0 continued fraction on the device backend:
// the reference's include/ed/thermal/ftlm.h:165-480, include/ed/thermal/sample_seed.h:37-89 and
// include/ed/dynamics/cf.h:39-161. Random starts are drawn on the host with the same recipe and the
// same standard library (splitmix64 -> std::mt19937 -> std::normal_distribution), so a seed gives
// the reference's vectors.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <random>
#include <vector>

#include "krylov.h"

namespace qsa {

using cplx = std::complex<double>;

inline std::uint64_t splitmix_seed(std::uint64_t base, std::uint64_t idx) {
    std::uint64_t z = base + 0x9E3779B97F4A7C15ULL * (idx + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/// The engine of sample s: std::mt19937 of the low 32 bits of splitmix64(base + phi (s + 1)).
inline std::mt19937 sample_engine(std::uint64_t base, std::uint64_t s) {
    return std::mt19937(static_cast<std::mt19937::result_type>(splitmix_seed(base, s)));
}

/// n i.i.d. complex Gaussians Complex(nd(gen), nd(gen)) from one normal_distribution, normalised
/// (sample_seed.h:78-88). The reference's compiler (GCC on x86-64) evaluates the constructor's
/// second argument first, so each entry's first draw is its imaginary part: for a real H the two
/// orders give the same Lanczos coefficients (one vector is i conj of the other), for a complex H
/// only this one reproduces the reference's samples.
inline std::vector<cplx> gaussian_vector(std::size_t n, std::mt19937& gen) {
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<cplx> v(n);
    for (auto& x : v) {
        const double im = nd(gen);
        const double re = nd(gen);
        x = cplx(re, im);
    }
    // The 2-norm with the scaling of the reference BLAS dznrm2 (scale / sum-of-squares update).
    double scale = 0.0, ssq = 1.0;
    for (const auto& x : v)
        for (double c : {x.real(), x.imag()}) {
            if (c != 0.0) {
                const double a = std::abs(c);
                if (scale < a) {
                    ssq = 1.0 + ssq * (scale / a) * (scale / a);
                    scale = a;
                } else {
                    ssq += (a / scale) * (a / scale);
                }
            }
        }
    const double norm = scale * std::sqrt(ssq);
    if (norm > 0.0) {
        const cplx s(1.0 / norm, 0.0);
        for (auto& x : v) x *= s;
    }
    return v;
}

struct Curves {
    std::vector<double> lnZ, E, V;
    double e_min = 0.0;
};

struct SampleMoments {
    double e_min = 0.0;
    std::vector<double> Z, E1, E2;
};

inline SampleMoments sample_moments(const std::vector<double>& ritz, const std::vector<double>& w,
                                    const std::vector<double>& betas) {
    SampleMoments m;
    m.e_min = std::numeric_limits<double>::infinity();
    for (std::size_t i = 0; i < ritz.size(); ++i)
        if (w[i] > 0.0) m.e_min = std::min(m.e_min, ritz[i]);
    if (!std::isfinite(m.e_min)) m.e_min = *std::min_element(ritz.begin(), ritz.end());
    const std::size_t nT = betas.size();
    m.Z.resize(nT);
    m.E1.resize(nT);
    m.E2.resize(nT);
    for (std::size_t t = 0; t < nT; ++t) {
        double z = 0.0, e1 = 0.0, e2 = 0.0;
        for (std::size_t i = 0; i < ritz.size(); ++i) {
            const double x = ritz[i] - m.e_min;
            const double b = w[i] * std::exp(-betas[t] * x);
            z += b;
            e1 += x * b;
            e2 += x * x * b;
        }
        m.Z[t] = z;
        m.E1[t] = e1;
        m.E2[t] = e2;
    }
    return m;
}

inline Curves combine_samples(const std::vector<SampleMoments>& S, std::uint64_t n, const std::vector<double>& betas) {
    Curves c;
    const std::size_t nT = betas.size(), R = S.size();
    double e_ref = S.front().e_min;
    for (const auto& s : S) e_ref = std::min(e_ref, s.e_min);
    c.e_min = e_ref;
    const double ln_n = std::log(static_cast<double>(n));
    c.lnZ.resize(nT);
    c.E.resize(nT);
    c.V.resize(nT);
    for (std::size_t t = 0; t < nT; ++t) {
        const double beta = betas[t];
        double z = 0.0, e1 = 0.0, e2 = 0.0;
        for (const auto& s : S) {
            const double d = s.e_min - e_ref, r = std::exp(-beta * d);
            z += s.Z[t] * r;
            e1 += (s.E1[t] + d * s.Z[t]) * r;
            e2 += (s.E2[t] + 2.0 * d * s.E1[t] + d * d * s.Z[t]) * r;
        }
        z /= static_cast<double>(R);
        e1 /= static_cast<double>(R);
        e2 /= static_cast<double>(R);
        if (z > 1e-300) {
            const double m1 = e1 / z, m2 = e2 / z;
            c.lnZ[t] = ln_n + std::log(z) - beta * e_ref;
            c.E[t] = e_ref + m1;
            c.V[t] = std::max(m2 - m1 * m1, 0.0);
        } else {
            c.lnZ[t] = -beta * e_ref;
            c.E[t] = e_ref;
            c.V[t] = 0.0;
        }
    }
    return c;
}

/// The exact thermodynamics of a block from its whole spectrum (include/ed/thermal/curves.h:38-65):
/// Boltzmann weights about the lowest level, E from the first moment, V by a second pass about E.
inline Curves exact_curves(const std::vector<double>& ev, const std::vector<double>& betas) {
    Curves c;
    const double e0 = *std::min_element(ev.begin(), ev.end());
    c.e_min = e0;
    for (double bt : betas) {
        double z = 0.0, e = 0.0;
        for (double x : ev) {
            const double w = std::exp(-bt * (x - e0));
            z += w;
            e += w * (x - e0);
        }
        c.lnZ.push_back(std::log(z) - bt * e0);
        c.E.push_back(e0 + e / z);
        double v = 0.0;
        for (double x : ev) {
            const double dx = x - e0 - e / z;
            v += std::exp(-bt * (x - e0)) * dx * dx;
        }
        c.V.push_back(v / z);
    }
    return c;
}

/// The whole system from its blocks (src/engine/thermal.cpp:597-640): block weights w_j
/// (multiplicities), p_j = w_j Z_j / max, lnZ by log-sum-exp, E the weighted mean, C = beta^2 times
/// the law-of-total-variance V, S = lnZ + beta E, F = -lnZ / beta.
struct SystemCurves {
    std::vector<double> lnZ, E, C, S, F;
};
inline SystemCurves combine_blocks(const std::vector<Curves>& blocks, const std::vector<double>& weights,
                                   const std::vector<double>& betas) {
    SystemCurves out;
    const std::size_t nT = betas.size(), nb = blocks.size();
    out.lnZ.resize(nT);
    out.E.resize(nT);
    out.C.resize(nT);
    out.S.resize(nT);
    out.F.resize(nT);
    std::vector<double> p(nb);
    for (std::size_t i = 0; i < nT; ++i) {
        double mx = -std::numeric_limits<double>::infinity();
        for (std::size_t j = 0; j < nb; ++j) mx = std::max(mx, std::log(weights[j]) + blocks[j].lnZ[i]);
        double z = 0.0, e = 0.0;
        for (std::size_t j = 0; j < nb; ++j) {
            p[j] = std::exp(std::log(weights[j]) + blocks[j].lnZ[i] - mx);
            z += p[j];
            e += p[j] * blocks[j].E[i];
        }
        e /= z;
        double var = 0.0;
        for (std::size_t j = 0; j < nb; ++j) {
            const double d = blocks[j].E[i] - e;
            var += p[j] * (blocks[j].V[i] + d * d);
        }
        var /= z;
        const double bt = betas[i];
        out.lnZ[i] = mx + std::log(z);
        out.E[i] = e;
        out.C[i] = bt * bt * var;
        out.S[i] = out.lnZ[i] + bt * e;
        out.F[i] = -out.lnZ[i] / bt;
    }
    return out;
}

struct FtlmOptions {
    std::size_t samples = 40, krylov = 100;
    double breakdown_tol = 0.0;
    std::vector<double> betas;
    std::uint64_t seed = 1;
};

/// FTLM of one block, one sample after another: LocalDGKS3 Lanczos without a basis per sample,
/// the tridiagonal's Ritz values and first-component weights, the moments, then the
/// Jensen-correct combination. Complex vectors throughout (the reference's sample vectors are
/// complex whatever the block).
template <class Matvec>
Curves ftlm_block(const HipBackend<cplx>& be, Matvec&& H, std::size_t n, const FtlmOptions& o,
                  LanczosTimings* timings = nullptr) {
    std::vector<SampleMoments> mom;
    auto v0 = be.make_zero_vector(n);
    for (std::size_t s = 0; s < o.samples; ++s) {
        std::mt19937 rng = sample_engine(o.seed, s);
        const auto h = gaussian_vector(n, rng);
        be.copy_from_host(h.data(), v0.get(), n);
        LanczosOptions<cplx> ko;
        ko.max_iter = o.krylov;
        if (o.breakdown_tol > 0.0) ko.breakdown_tol = o.breakdown_tol;
        ko.reorth = Reorth::LocalDGKS3;
        ko.keep_basis = false;
        auto r = lanczos_run(be, H, n, v0.get(), ko);
        if (timings) {
            timings->apply_s += r.timings.apply_s;
            timings->applies += r.timings.applies;
            timings->total_s += r.timings.total_s;
        }
        if (r.alpha.empty()) continue;
        std::vector<double> values, weights;
        tridiag_weights(r.alpha.data(), r.beta.data() + 1, r.alpha.size(), values, weights);
        mom.push_back(sample_moments(values, weights, o.betas));
    }
    if (mom.empty()) throw Error(kConvergence, "ftlm: every sample failed");
    return combine_samples(mom, n, o.betas);
}

/// S(w) = -Im G(w + i eta) / pi, G = norm_sq / (z - a0 - b1^2 / (z - a1 - ...)), bottom-up.
inline std::vector<double> continued_fraction(const std::vector<double>& alpha, const std::vector<double>& beta,
                                              const std::vector<double>& omega, double eta, double norm_sq) {
    std::vector<double> S(omega.size(), 0.0);
    const std::size_t M = alpha.size();
    if (M == 0) return S;
    const double kPi = 3.14159265358979323846;
    for (std::size_t iw = 0; iw < omega.size(); ++iw) {
        const cplx z(omega[iw], eta);
        cplx G(0.0, 0.0);
        for (std::size_t nn = M - 1; nn >= 1; --nn) {
            const double b2 = nn < beta.size() ? beta[nn] * beta[nn] : 0.0;
            const cplx d = z - cplx(alpha[nn], 0.0) - G;
            G = std::abs(d) > 1e-300 ? cplx(b2, 0.0) / d : cplx(0.0, 0.0);
        }
        const cplx d = z - cplx(alpha[0], 0.0) - G;
        const cplx G0 = std::abs(d) > 1e-300 ? cplx(norm_sq, 0.0) / d : cplx(0.0, 0.0);
        S[iw] = -G0.imag() / kPi;
    }
    return S;
}

struct CfResult {
    std::vector<double> S, alpha, beta;
    double energy_shift = 0.0, norm_sq = 0.0, convergence_change = 0.0;
};

/// The continued-fraction kernel from a device vector phi (cf.h:103-161): normalise, Lanczos
/// without reorthogonalisation or basis, shift by E_shift (or the lowest Ritz value), evaluate,
/// and the half-depth change as the convergence measure.
template <class Matvec>
CfResult cf_from_vector(const HipBackend<cplx>& be, Matvec&& H, std::size_t n, const cplx* phi_dev,
                        const std::vector<double>& omega, double eta, std::size_t krylov, bool has_shift,
                        double shift, double tol) {
    CfResult R;
    auto phi = be.make_zero_vector(n);
    be.copy(phi_dev, phi.get(), n);
    const double nrm = be.nrm2(phi.get(), n);
    R.norm_sq = nrm * nrm;
    if (nrm < 1e-14) {
        R.S.assign(omega.size(), 0.0);
        return R;
    }
    be.scale(cplx(1.0 / nrm, 0.0), phi.get(), n);
    LanczosOptions<cplx> ko;
    ko.max_iter = krylov;
    ko.reorth = Reorth::None;
    ko.keep_basis = false;
    ko.breakdown_tol = tol;
    auto r = lanczos_run(be, H, n, phi.get(), ko);
    R.alpha = r.alpha;
    R.beta = r.beta;
    if (r.alpha.empty()) {
        R.S.assign(omega.size(), 0.0);
        return R;
    }
    double E = shift;
    if (!has_shift) E = tridiag_eig(r.alpha, r.beta, r.alpha.size(), false).values.front();
    std::vector<double> a = r.alpha;
    for (auto& x : a) x -= E;
    R.energy_shift = E;
    R.S = continued_fraction(a, r.beta, omega, eta, R.norm_sq);
    if (a.size() >= 4) {
        const std::size_t h = a.size() / 2;
        std::vector<double> a2(a.begin(), a.begin() + static_cast<std::ptrdiff_t>(h));
        std::vector<double> b2(r.beta.begin(), r.beta.begin() + static_cast<std::ptrdiff_t>(std::min(r.beta.size(), h + 1)));
        const auto Sh = continued_fraction(a2, b2, omega, eta, R.norm_sq);
        double smax = 0.0, dmax = 0.0;
        for (std::size_t i = 0; i < R.S.size(); ++i) {
            smax = std::max(smax, std::abs(R.S[i]));
            dmax = std::max(dmax, std::abs(R.S[i] - Sh[i]));
        }
        R.convergence_change = smax > 0.0 ? dmax / smax : 0.0;
    }
    return R;
}

}  // namespace qsa

// Thermodynamics from Lanczos data and the T = 0 continued fraction, on the device backend.
//
// - Seeds: the sample recipe of the reference (include/ed/thermal/sample_seed.h:37-89): splitmix64
//   of the base seed and the sample index, its low 32 bits seeding std::mt19937, complex Gaussian
//   entries from std::normal_distribution (the standard library the reference also uses, so a seed
//   gives the reference's vectors), normalised.
// - FTLM: a block's samples are pooled as one list of Ritz pairs (value, squared first component)
//   with the sample boundaries kept; every curve is a weighted sum over the pool about its lowest
//   weighted Ritz value. Z(beta) = dim / R * sum w e^{-beta theta}; E the weighted mean; V the
//   weighted second moment about E (a second pass, no cancellation of raw moments). This is the
//   same estimator as the reference's per-sample moments recombined about a common shift
//   (include/ed/thermal/ftlm.h:165-235), grouped differently.
// - Exact curves of a whole spectrum (include/ed/thermal/curves.h:38-65) and the system's curves
//   from its blocks by log-sum-exp and the law of total variance (src/engine/thermal.cpp:597-640).
// - The continued fraction G(z) = |phi|^2 / (z - a_0 - b_1^2 / (z - a_1 - ...)) of a Lanczos run
//   (include/ed/dynamics/cf.h:39-161), evaluated for all frequencies level by level from the bottom.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <random>
#include <vector>

#include "lanczos.h"
#include "tridiag.h"

namespace qsa {

using cplx = std::complex<double>;

// ---- seeds -------------------------------------------------------------------------------------

/// splitmix64 (Steele, Lea and Flood 2014) of base + phi (idx + 1), phi the 64-bit golden ratio.
inline std::uint64_t splitmix_seed(std::uint64_t base, std::uint64_t idx) {
    std::uint64_t x = base + 0x9E3779B97F4A7C15ULL * (idx + 1);
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/// The generator of sample `s` of a block seeded `base`.
inline std::mt19937 sample_engine(std::uint64_t base, std::uint64_t s) {
    return std::mt19937(static_cast<std::mt19937::result_type>(splitmix_seed(base, s)));
}

/// The base seed of one sector of a walk: splitmix64 of the walk's seed and the sector's key.
inline std::uint64_t block_seed(std::uint64_t base, std::uint64_t key) { return splitmix_seed(base, key); }

/// The Euclidean norm of 2n doubles by the scaled sum of squares of the reference BLAS nrm2
/// (a running scale, the largest magnitude so far, and the sum of squares relative to it), which
/// neither overflows nor underflows on the way.
inline double scaled_norm(const double* x, std::size_t count) {
    double big = 0.0, rel = 1.0;
    for (std::size_t i = 0; i < count; ++i) {
        const double a = std::abs(x[i]);
        if (a == 0.0) continue;
        if (a > big) {
            const double r = big / a;
            rel = 1.0 + rel * r * r;
            big = a;
        } else {
            const double r = a / big;
            rel += r * r;
        }
    }
    return big * std::sqrt(rel);
}

/// n complex Gaussian entries of one normal_distribution on `gen`, normalised. Each entry is the
/// reference's Complex(nd(gen), nd(gen)), whose compiler (GCC on x86-64) evaluates the second
/// argument first: the first draw of an entry is its imaginary part. A real H gives the same
/// Lanczos coefficients in either order; a complex H reproduces the reference's samples only in
/// this one.
inline std::vector<cplx> gaussian_vector(std::size_t n, std::mt19937& gen) {
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<cplx> v(n);
    for (cplx& x : v) {
        const double im = nd(gen);
        x = cplx(nd(gen), im);
    }
    const double norm = scaled_norm(reinterpret_cast<const double*>(v.data()), 2 * n);
    if (norm > 0.0) {
        const double inv = 1.0 / norm;
        for (cplx& x : v) x *= cplx(inv, 0.0);
    }
    return v;
}

// ---- curves ------------------------------------------------------------------------------------

/// lnZ, E and V = <(H - E)^2> per beta, and the lowest (weighted) level they are measured from.
struct Curves {
    std::vector<double> lnZ, E, V;
    double e_min = 0.0;
};

/// The Ritz pairs of a block's FTLM samples, end to end.
class RitzPool {
public:
    void add_sample(const std::vector<double>& values, const std::vector<double>& weights) {
        theta_.insert(theta_.end(), values.begin(), values.end());
        weight_.insert(weight_.end(), weights.begin(), weights.end());
        ++samples_;
    }
    std::size_t samples() const { return samples_; }
    std::size_t size() const { return theta_.size(); }

    /// The sampled curves of a block of `dim` states: the pooled pairs weighted by w e^{-beta
    /// (theta - e0)}, e0 the lowest value of positive weight (the lowest of all when none has),
    /// Z scaled by dim / samples.
    Curves curves(std::uint64_t dim, const std::vector<double>& betas) const {
        Curves c;
        if (samples_ == 0 || theta_.empty()) throw Error(kConvergence, "ftlm: no sample produced Ritz values");
        double e0 = std::numeric_limits<double>::infinity(), lowest = e0;
        for (std::size_t k = 0; k < theta_.size(); ++k) {
            lowest = std::min(lowest, theta_[k]);
            if (weight_[k] > 0.0) e0 = std::min(e0, theta_[k]);
        }
        if (!std::isfinite(e0)) e0 = lowest;
        c.e_min = e0;
        const double log_scale = std::log(static_cast<double>(dim)) - std::log(static_cast<double>(samples_));
        std::vector<double> boltz(theta_.size());
        for (double beta : betas) {
            double z = 0.0, e1 = 0.0;
            for (std::size_t k = 0; k < theta_.size(); ++k) {
                boltz[k] = weight_[k] * std::exp(-beta * (theta_[k] - e0));
                z += boltz[k];
                e1 += boltz[k] * (theta_[k] - e0);
            }
            if (!(z > 1e-300)) {   // nothing weighs at this temperature: the block sits at e0
                c.lnZ.push_back(-beta * e0);
                c.E.push_back(e0);
                c.V.push_back(0.0);
                continue;
            }
            const double mean = e1 / z;
            double e2 = 0.0;
            for (std::size_t k = 0; k < theta_.size(); ++k) {
                const double d = theta_[k] - e0 - mean;
                e2 += boltz[k] * d * d;
            }
            c.lnZ.push_back(log_scale + std::log(z) - beta * e0);
            c.E.push_back(e0 + mean);
            c.V.push_back(e2 / z);
        }
        return c;
    }

private:
    std::vector<double> theta_, weight_;
    std::size_t samples_ = 0;
};

/// The exact curves of a block from its whole spectrum: the pool of one sample holding every level
/// with unit weight, and no dim / samples factor.
inline Curves exact_curves(const std::vector<double>& levels, const std::vector<double>& betas) {
    RitzPool pool;
    pool.add_sample(levels, std::vector<double>(levels.size(), 1.0));
    return pool.curves(1, betas);
}

/// The system's curves from blocks of multiplicity w_j: lnZ = log sum_j w_j Z_j (by log-sum-exp),
/// E the Z-weighted mean, C = beta^2 (the within-block variances plus the spread of the block
/// means), S = lnZ + beta E, F = -lnZ / beta.
struct SystemCurves {
    std::vector<double> lnZ, E, C, S, F;
};
inline SystemCurves combine_blocks(const std::vector<Curves>& blocks, const std::vector<double>& weights,
                                   const std::vector<double>& betas) {
    SystemCurves out;
    const std::size_t nb = blocks.size();
    std::vector<double> logw(nb), p(nb);
    for (std::size_t j = 0; j < nb; ++j) logw[j] = std::log(weights[j]);
    for (std::size_t i = 0; i < betas.size(); ++i) {
        double top = -std::numeric_limits<double>::infinity();
        for (std::size_t j = 0; j < nb; ++j) top = std::max(top, logw[j] + blocks[j].lnZ[i]);
        double z = 0.0, ez = 0.0;
        for (std::size_t j = 0; j < nb; ++j) {
            p[j] = std::exp(logw[j] + blocks[j].lnZ[i] - top);
            z += p[j];
            ez += p[j] * blocks[j].E[i];
        }
        const double e = ez / z;
        double var = 0.0;
        for (std::size_t j = 0; j < nb; ++j) {
            const double spread = blocks[j].E[i] - e;
            var += p[j] * (blocks[j].V[i] + spread * spread);
        }
        const double beta = betas[i], lnz = top + std::log(z);
        out.lnZ.push_back(lnz);
        out.E.push_back(e);
        out.C.push_back(beta * beta * var / z);
        out.S.push_back(lnz + beta * e);
        out.F.push_back(-lnz / beta);
    }
    return out;
}

// ---- FTLM of one block, sample by sample ---------------------------------------------------------

struct FtlmOptions {
    std::size_t samples = 40, krylov = 100;
    double breakdown_tol = 0.0;
    std::vector<double> betas;
    std::uint64_t seed = 1;
};

/// FTLM of one block without batching: each sample's start drawn on the host, a LocalDGKS3 Lanczos
/// run without a basis, its Ritz values and weights pooled. Complex vectors throughout (the samples
/// are complex whatever the block).
template <class Matvec>
Curves ftlm_block(const HipBackend<cplx>& be, Matvec&& H, std::size_t n, const FtlmOptions& o,
                  LanczosTimings* timings = nullptr) {
    RitzPool pool;
    auto start = be.make_zero_vector(n);
    LanczosOptions<cplx> lo;
    lo.max_iter = o.krylov;
    if (o.breakdown_tol > 0.0) lo.breakdown_tol = o.breakdown_tol;
    lo.reorth = Reorth::LocalDGKS3;
    lo.keep_basis = false;
    for (std::size_t s = 0; s < o.samples; ++s) {
        std::mt19937 gen = sample_engine(o.seed, s);
        const std::vector<cplx> h = gaussian_vector(n, gen);
        be.copy_from_host(h.data(), start.get(), n);
        const auto run = lanczos_run(be, H, n, start.get(), lo);
        if (timings) {
            timings->apply_s += run.timings.apply_s;
            timings->applies += run.timings.applies;
            timings->total_s += run.timings.total_s;
        }
        if (run.alpha.empty()) continue;
        std::vector<double> values, weights;
        tridiag_weights(run.alpha.data(), run.beta.data() + 1, run.alpha.size(), values, weights);
        pool.add_sample(values, weights);
    }
    return pool.curves(n, o.betas);
}

// ---- continued fraction ------------------------------------------------------------------------

/// S(omega) = -Im G(omega + i eta) / pi for the tridiagonal (alpha[0..M), beta[1..M)) and weight
/// norm_sq, every frequency at once: starting from the deepest level, g <- beta_l^2 / (z - alpha_l -
/// g) for l = M-1 .. 1, then G = norm_sq / (z - alpha_0 - g). A level whose denominator vanishes
/// (only possible at eta = 0, exactly on a pole) contributes nothing.
inline std::vector<double> continued_fraction(const std::vector<double>& alpha, const std::vector<double>& beta,
                                              const std::vector<double>& omega, double eta, double norm_sq) {
    const std::size_t nw = omega.size(), M = alpha.size();
    std::vector<double> S(nw, 0.0);
    if (M == 0) return S;
    std::vector<cplx> g(nw, cplx(0.0, 0.0));
    auto level = [&](std::size_t l, double numer) {
        for (std::size_t k = 0; k < nw; ++k) {
            const cplx d = cplx(omega[k] - alpha[l], eta) - g[k];
            g[k] = std::abs(d) > 1e-300 ? cplx(numer, 0.0) / d : cplx(0.0, 0.0);
        }
    };
    for (std::size_t l = M - 1; l >= 1; --l) level(l, l < beta.size() ? beta[l] * beta[l] : 0.0);
    level(0, norm_sq);
    constexpr double kInvPi = 0.31830988618379067154;
    for (std::size_t k = 0; k < nw; ++k) S[k] = -g[k].imag() * kInvPi;
    return S;
}

struct CfResult {
    std::vector<double> S, alpha, beta;
    double energy_shift = 0.0, norm_sq = 0.0, convergence_change = 0.0;
};

/// The spectral function of one start vector phi (device): Lanczos without reorthogonalisation from
/// phi / |phi|, the diagonal shifted by E (given, or the lowest Ritz value of the run), the continued
/// fraction at every frequency, and as the convergence measure the largest change against the
/// fraction cut at half depth, relative to the largest |S|.
template <class Matvec>
CfResult cf_from_vector(const HipBackend<cplx>& be, Matvec&& H, std::size_t n, const cplx* phi_dev,
                        const std::vector<double>& omega, double eta, std::size_t krylov, bool has_shift,
                        double shift, double tol) {
    CfResult R;
    R.S.assign(omega.size(), 0.0);
    auto phi = be.make_zero_vector(n);
    be.copy(phi_dev, phi.get(), n);
    const double norm = be.nrm2(phi.get(), n);
    R.norm_sq = norm * norm;
    if (norm < 1e-14) return R;
    be.scale(cplx(1.0 / norm, 0.0), phi.get(), n);
    LanczosOptions<cplx> lo;
    lo.max_iter = krylov;
    lo.reorth = Reorth::None;
    lo.keep_basis = false;
    lo.breakdown_tol = tol;
    const auto run = lanczos_run(be, H, n, phi.get(), lo);
    R.alpha = run.alpha;
    R.beta = run.beta;
    const std::size_t M = run.alpha.size();
    if (M == 0) return R;
    R.energy_shift = has_shift ? shift : tridiag_eig(run.alpha, run.beta, M, false).values.front();
    std::vector<double> shifted(M);
    for (std::size_t i = 0; i < M; ++i) shifted[i] = run.alpha[i] - R.energy_shift;
    R.S = continued_fraction(shifted, run.beta, omega, eta, R.norm_sq);
    if (M >= 4) {
        const std::size_t half = M / 2;
        const std::vector<double> a_half(shifted.begin(), shifted.begin() + static_cast<std::ptrdiff_t>(half));
        const auto S_half = continued_fraction(a_half, run.beta, omega, eta, R.norm_sq);
        double top = 0.0, change = 0.0;
        for (std::size_t k = 0; k < omega.size(); ++k) {
            top = std::max(top, std::abs(R.S[k]));
            change = std::max(change, std::abs(R.S[k] - S_half[k]));
        }
        R.convergence_change = top > 0.0 ? change / top : 0.0;
    }
    return R;
}

}  // namespace qsa

// The block eigensolver lanes on the device backend: the k = 1 Lanczos scan with its ghost-aware
// convergence gate (src/engine/block_solve.cpp:263-394, 528-598) and thick-restart Krylov-Schur
// for k > 1 (include/ed/krylov/krylov_schur.h:95-312, block_solve.cpp:433-510), plus the
// seeds and tolerances they use (include/ed/core/numerics.h:20-37).
#pragma once

#include <algorithm>
#include <cmath>
#include <future>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#include "lanczos.h"
#include "tridiag.h"

namespace qsa {

inline constexpr double kEps = 2.220446049250313e-16;
inline constexpr double kBreakdownRel = 64.0 * kEps;
inline constexpr double kLockRel = 1e-10;
inline constexpr std::uint64_t kKsStartSeed = 0x51ED0B70ULL;
inline double scale_or_one(double b) { return b > 0.0 ? b : 1.0; }

/// A unit-variance Gaussian start from std::mt19937_64(seed) (block_solve.cpp:268-277 staged_seed,
/// lanczos.h gaussian_entry). A complex entry is the reference's Complex(nd(gen), nd(gen)); C++
/// leaves the order of the two argument evaluations unspecified and the reference's compiler (GCC on
/// x86-64) evaluates the second first: the first draw is the imaginary part.
template <class Scalar> std::vector<Scalar> gaussian_start(std::size_t n, std::uint64_t seed) {
    std::vector<Scalar> h(n);
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (auto& c : h) {
        if constexpr (std::is_same_v<Scalar, double>) {
            c = nd(gen);
        } else {
            // Complex(nd(gen), nd(gen)): the reference's compiler evaluates the second argument
            // first, so the first draw is the imaginary part
            const double im = nd(gen);
            const double re = nd(gen);
            c = Scalar(re, im);
        }
    }
    return h;
}

/// Successive Gaussian vectors of one std::mt19937_64 stream, a fresh normal_distribution per
/// vector (krylov_schur.h:74-75, 240-276). With prefetch the first vector is drawn on a host thread
/// from construction on, while the caller's device work runs; later ones are drawn when taken.
template <class Scalar> class SeedStream {
public:
    SeedStream(std::uint64_t seed, std::size_t n, bool prefetch) : gen_(seed), n_(n) {
        if (prefetch) next_ = std::async(std::launch::async, [this] { return draw(); });
    }
    ~SeedStream() {
        if (next_.valid()) next_.wait();
    }
    SeedStream(const SeedStream&) = delete;
    SeedStream& operator=(const SeedStream&) = delete;
    std::vector<Scalar> take() { return next_.valid() ? next_.get() : draw(); }

private:
    std::vector<Scalar> draw() {
        std::vector<Scalar> h(n_);
        std::normal_distribution<double> nd(0.0, 1.0);
        for (auto& z : h) {
            if constexpr (std::is_same_v<Scalar, double>) {
                z = nd(gen_);
            } else {   // imaginary part first, as gaussian_start
                const double im = nd(gen_);
                const double re = nd(gen_);
                z = Scalar(re, im);
            }
        }
        return h;
    }
    std::mt19937_64 gen_;
    std::size_t n_;
    std::future<std::vector<Scalar>> next_;
};

/// The lowest LEVELS of Ritz values w[0..m) (ascending): values within 1e-9 scale of a level's first
/// copy are that level, converged when any copy's Paige bound is <= 1e-7 scale; the walk stops at
/// the first unconverged level (block_solve.cpp:335-357).
template <class Bound>
bool lowest_levels(std::size_t m, const double* w, Bound&& bound, double scale, std::size_t k, std::vector<double>* out) {
    std::size_t found = 0;
    for (std::size_t j = 0; j < m && found < k;) {
        const double first = w[j];
        std::size_t i = j;
        bool converged = false;
        double value = first;
        for (; i < m && std::abs(w[i] - first) <= 1e-9 * scale; ++i)
            if (!converged && bound(i) <= 1e-7 * scale) {
                converged = true;
                value = w[i];
            }
        if (!converged) return false;
        if (out) out->push_back(value);
        ++found;
        j = i;
    }
    return found >= k;
}

/// The leading invariant block: the step before the first off-diagonal at 1e-12 of the largest
/// |alpha|, |beta| (block_solve.cpp:366-376).
inline std::size_t leading_block(const std::vector<double>& alpha, const std::vector<double>& beta, std::size_t m) {
    double scale = 0.0;
    for (std::size_t i = 0; i < m; ++i) {
        scale = std::max(scale, std::abs(alpha[i]));
        if (i + 1 < beta.size()) scale = std::max(scale, std::abs(beta[i + 1]));
    }
    for (std::size_t i = 1; i < m; ++i)
        if (std::abs(beta[i]) <= 1e-12 * scale) return i;
    return m;
}

inline bool lowest_level(const std::vector<double>& alpha, const std::vector<double>& beta, std::size_t m,
                         std::vector<double>* out) {
    const double beta_m = beta.size() > m ? std::abs(beta[m]) : 0.0;
    const TridiagEig f = tridiag_eig(alpha, beta, m, true);
    const double scale = std::max({std::abs(f.values.front()), std::abs(f.values.back()), 1e-300});
    return lowest_levels(
        m, f.values.data(), [&](std::size_t j) { return beta_m * std::abs(f.z(m - 1, j)); }, scale, 1, out);
}

struct BlockSolution {
    std::vector<double> values;
    bool converged = false;
    bool whole = false;
    std::size_t applies = 0;
    LanczosTimings timings;
};

/// k = 1: the pure three-term recurrence (full reorthogonalisation only when the run may span the
/// block) with the Paige gate every step.
template <class Scalar, class Matvec>
BlockSolution lowest_scan(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n, std::size_t max_iter) {
    LanczosOptions<Scalar> o;
    o.max_iter = std::min(n, max_iter);
    o.reorth = n <= o.max_iter ? Reorth::FullCGS2 : Reorth::None;
    o.keep_basis = o.reorth == Reorth::FullCGS2;
    o.convergence_check = [](const std::vector<double>& a, const std::vector<double>& b) {
        const std::size_t m_all = a.size();
        const std::size_t m = leading_block(a, b, m_all);
        if (m == m_all && m_all < 3) return false;
        return lowest_level(a, b, m, nullptr);
    };
    o.convergence_check_interval = 1;
    const auto h = gaussian_start<Scalar>(n, kKsStartSeed);
    auto v0 = be.make_zero_vector(n);
    be.copy_from_host(h.data(), v0.get(), n);
    auto r = lanczos_run(be, H, n, v0.get(), o);
    BlockSolution sol;
    sol.applies = r.timings.applies;
    sol.timings = r.timings;
    if (r.alpha.empty()) return sol;
    const std::size_t m = leading_block(r.alpha, r.beta, r.alpha.size());
    sol.converged = lowest_level(r.alpha, r.beta, m, &sol.values);
    return sol;
}

/// The pruning estimate of a block (block_solve.cpp:630-655): 40 Lanczos steps without
/// reorthogonalisation from the fixed start gaussian_start(n, 0xE57A7E); theta is the lowest Ritz
/// value (an upper bound on the block's lowest level), residual |beta_m| |z_(m-1),0| its residual
/// bound. The caller prunes on theta - residual.
struct BlockEstimate {
    double theta = -std::numeric_limits<double>::infinity();
    double residual = 0.0;
    std::size_t applies = 0;
};
inline constexpr std::uint64_t kEstimateSeed = 0xE57A7EULL;

template <class Scalar, class Matvec>
BlockEstimate estimate_lowest(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n) {
    BlockEstimate est;
    if (n == 0) return est;
    LanczosOptions<Scalar> o;
    o.max_iter = std::min<std::size_t>(40, n);
    o.reorth = Reorth::None;
    o.keep_basis = false;
    const auto h = gaussian_start<Scalar>(n, kEstimateSeed);
    auto v0 = be.make_zero_vector(n);
    be.copy_from_host(h.data(), v0.get(), n);
    auto r = lanczos_run(be, H, n, v0.get(), o);
    est.applies = r.timings.applies;
    if (r.alpha.empty()) return est;
    const std::size_t m = r.alpha.size();
    const TridiagEig t = tridiag_eig(r.alpha, r.beta, m, true);
    est.theta = t.values.front();
    est.residual = r.beta.size() > m ? std::abs(r.beta[m]) * std::abs(t.z(m - 1, 0)) : 0.0;
    return est;
}

struct KsOptions {
    std::size_t num_eigs = 1;
    std::size_t max_iter = 100;        ///< cycle length hint
    double tolerance = 1e-10;
    std::size_t max_restarts = 30;
    double breakdown_tol = 1e-13;
    std::uint64_t max_subspace_vectors = 0;
    bool probe_degeneracy = true;
};

inline std::size_t krylov_subspace_dim(std::size_t nev, std::size_t requested, std::uint64_t global_dim,
                                       std::uint64_t max_vectors) {
    std::size_t m = 2 * nev + 20;
    if (requested > m) m = requested;
    if (max_vectors > 0 && m > max_vectors) m = static_cast<std::size_t>(max_vectors);
    if (global_dim > 0 && m > global_dim) m = static_cast<std::size_t>(global_dim);
    if (m < nev + 1) m = nev + 1;
    return m;
}

struct KsResult {
    std::vector<double> eigenvalues;
    std::size_t applies = 0, restarts = 0;
    bool converged = false, exhausted = false;
};

/// Thick-restart Krylov-Schur with locking, fresh deflated starts and the degeneracy probe.
template <class Scalar, class Matvec>
KsResult krylov_schur(const HipBackend<Scalar>& be, Matvec&& matvec, std::size_t n, const Scalar* seed,
                      const KsOptions& opts, Scalar* vectors_out = nullptr, std::size_t vectors_cap = 0) {
    KsResult R;
    if (n == 0) {
        R.converged = R.exhausted = true;
        return R;
    }
    const std::size_t k = std::max<std::size_t>(1, opts.num_eigs);
    const std::size_t m_cap = krylov_subspace_dim(k, opts.max_iter, n, opts.max_subspace_vectors);
    const std::size_t p_want = k + std::max<std::size_t>(k / 2, 8);
    const std::size_t m = std::max<std::size_t>(1, std::min<std::size_t>({m_cap, 2 * p_want + 20, n}));
    const std::size_t p_keep = std::min(p_want, m > 1 ? m - 1 : std::size_t{1});
    auto Vown = be.make_zero_vector(n * (m + 1));
    Scalar* V = Vown.get();
    auto col = [&](std::size_t j) { return V + j * n; };
    auto w = be.make_zero_vector(n);
    const std::size_t found_cap = 2 * k + 4;
    auto Fown = be.make_zero_vector(n * found_cap);
    std::size_t n_found = 0;
    std::vector<double> found_vals;
    DeviceBuffer Sdev;
    // two unconditional classical passes against columns [0, j) and the found vectors; returns the
    // accumulated coefficients on the columns
    auto orthogonalize = [&](Scalar* x, std::size_t j) {
        std::vector<Scalar> c(j, Scalar(0)), c2(std::max<std::size_t>(j, n_found));
        for (int pass = 0; pass < 2 && (j + n_found) > 0; ++pass) {
            if (j) {
                be.dot_many_block(V, n, j, x, n, c2.data());
                for (std::size_t i = 0; i < j; ++i) {
                    c[i] += c2[i];
                    c2[i] = -c2[i];
                }
                be.axpy_many_block(c2.data(), V, n, j, x, n);
            }
            if (n_found) {
                be.dot_many_block(Fown.get(), n, n_found, x, n, c2.data());
                for (std::size_t i = 0; i < n_found; ++i) c2[i] = -c2[i];
                be.axpy_many_block(c2.data(), Fown.get(), n, n_found, x, n);
            }
        }
        return c;
    };
    auto rotate = [&](const std::vector<double>& S, std::size_t mm, std::size_t cols, Scalar* out) {
        Sdev = upload(S.data(), mm * cols);
        dev::block_rotate(HipBackend<Scalar>::kComplex, V, n, mm, Sdev.get<double>(), cols, out, n, engine_stream());
        QSA_HIP(hipGetLastError());
        be.synchronize();
    };
    enum Outcome { kConverged, kBudget, kNullSeed, kNothingBelow };
    auto search = [&](std::size_t want, std::size_t budget, double give_up) -> Outcome {
        orthogonalize(col(0), 0);
        const double nrm = be.nrm2(col(0), n);
        if (nrm < 1e-13) return kNullSeed;
        be.scale(Scalar(1.0 / nrm), col(0), n);
        std::size_t p = 0;
        std::vector<double> theta, arrow;
        for (std::size_t cycle = 0; cycle < budget; ++cycle) {
            std::vector<double> T(m * m, 0.0);
            for (std::size_t i = 0; i < p; ++i) {
                T[i + i * m] = theta[i];
                T[i + p * m] = T[p + i * m] = arrow[i];
            }
            std::size_t mm = m;
            double beta_last = 0.0;
            for (std::size_t j = p; j < m; ++j) {
                matvec(col(j), w.get());
                ++R.applies;
                const auto c = orthogonalize(w.get(), j + 1);
                T[j + j * m] = std::real(c[j]);
                const double beta = be.nrm2(w.get(), n);
                if (beta <= opts.breakdown_tol) {
                    mm = j + 1;
                    beta_last = 0.0;
                    break;
                }
                if (j + 1 < m) T[j + (j + 1) * m] = T[(j + 1) + j * m] = beta;
                be.copy(w.get(), col(j + 1), n);
                be.scale(Scalar(1.0 / beta), col(j + 1), n);
                beta_last = beta;
            }
            ++R.restarts;
            std::vector<double> Tm(mm * mm);
            for (std::size_t c = 0; c < mm; ++c)
                for (std::size_t r = 0; r < mm; ++r) Tm[c * mm + r] = T[c * m + r];
            const TridiagEig e = symmetric_eig(std::move(Tm), mm);
            if (cycle == 0 && e.values[0] >= give_up) return kNothingBelow;
            const std::size_t top = std::min(want, mm);
            std::size_t conv = 0;
            while (conv < top && std::abs(beta_last) * std::abs(e.vectors[conv * mm + (mm - 1)]) < opts.tolerance) ++conv;
            const bool done = conv == top;
            if (done || cycle + 1 == budget) {
                if (conv > 0) {
                    if (n_found + conv > found_cap) throw Error(kRuntime, "krylov_schur: found-vector capacity");
                    rotate(e.vectors, mm, conv, Fown.get() + n_found * n);
                    for (std::size_t i = 0; i < conv; ++i) found_vals.push_back(e.values[i]);
                    n_found += conv;
                }
                return done ? kConverged : kBudget;
            }
            p = std::min(p_keep, mm - 1);
            {
                auto U = be.make_zero_vector(n * p);
                rotate(e.vectors, mm, p, U.get());
                be.copy(U.get(), col(0), n * p);
            }
            be.copy(col(mm), col(p), n);
            theta.assign(e.values.begin(), e.values.begin() + static_cast<std::ptrdiff_t>(p));
            arrow.resize(p);
            for (std::size_t i = 0; i < p; ++i) arrow[i] = beta_last * e.vectors[i * mm + (mm - 1)];
        }
        return kBudget;
    };
    // fresh deflated starts, and the degeneracy probe's starts (its first one drawn on a host thread
    // while the search runs)
    SeedStream<Scalar> fresh_gen(0xF8E5A7C3ULL, n, false);
    SeedStream<Scalar> probe_gen(0xDE6E4E7AULL, n, opts.probe_degeneracy);
    auto fresh_seed = [&](SeedStream<Scalar>& gen) {
        const std::vector<Scalar> host = gen.take();
        be.copy_from_host(host.data(), col(0), n);
    };
    be.copy(seed, col(0), n);
    bool exhausted = false, stalled = false;
    const std::size_t budget = opts.max_restarts;
    int null_starts = 0;
    while (found_vals.size() < k && R.restarts < budget) {
        const Outcome o = search(k - found_vals.size(), budget - R.restarts, std::numeric_limits<double>::infinity());
        if (o == kNullSeed) {
            if (++null_starts >= 2) {
                exhausted = true;
                break;
            }
            fresh_seed(fresh_gen);
            continue;
        }
        null_starts = 0;
        if (o == kBudget) {
            stalled = true;
            break;
        }
        if (found_vals.size() < k) fresh_seed(fresh_gen);
    }
    bool converged = !stalled && found_vals.size() >= k;
    if (!found_vals.empty() && !exhausted && opts.probe_degeneracy) {
        for (std::size_t round = 0; round < k; ++round) {
            const double top = *std::max_element(found_vals.begin(), found_vals.end());
            const double gap = std::max(10.0 * opts.tolerance, 1e-8 * std::abs(top));
            fresh_seed(probe_gen);
            const std::size_t before = found_vals.size();
            const Outcome o = search(1, std::max<std::size_t>(opts.max_restarts, 1), top - gap);
            if (o == kNothingBelow || o == kNullSeed) break;
            if (found_vals.size() == before) {
                converged = false;
                break;
            }
            if (found_vals.back() >= top - gap) {
                found_vals.pop_back();
                --n_found;
                break;
            }
            while (found_vals.size() > k) {
                const auto hi = std::max_element(found_vals.begin(), found_vals.end()) - found_vals.begin();
                // drop the highest: move the last found vector into its slot
                const std::size_t last = n_found - 1;
                if (static_cast<std::size_t>(hi) != last) {
                    be.copy(Fown.get() + last * n, Fown.get() + static_cast<std::size_t>(hi) * n, n);
                    found_vals[static_cast<std::size_t>(hi)] = found_vals[last];
                }
                found_vals.pop_back();
                --n_found;
            }
        }
    }
    // ascending, the locked vectors in the same order
    std::vector<std::size_t> order(found_vals.size());
    for (std::size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) { return found_vals[a] < found_vals[b]; });
    R.eigenvalues.resize(order.size());
    for (std::size_t i = 0; i < order.size(); ++i) R.eigenvalues[i] = found_vals[order[i]];
    if (vectors_out)
        for (std::size_t i = 0; i < order.size() && i < vectors_cap; ++i)
            be.copy(Fown.get() + order[i] * n, vectors_out + i * n, n);
    R.converged = converged;
    R.exhausted = exhausted;
    return R;
}

/// The lowest k levels of one block through Krylov-Schur with the lane's budgets
/// (block_solve.cpp:433-510): iteration budget max(200 k, 2000), cycle min(2k + 60 hint), lock
/// tolerance kLockRel * s_H, breakdown kBreakdownRel * s_H.
template <class Scalar, class Matvec>
BlockSolution krylov_schur_lane(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n, std::size_t k,
                                double norm_bound, std::uint64_t max_iter, Scalar* vectors_out = nullptr,
                                std::future<std::vector<Scalar>>* start = nullptr) {
    const std::uint64_t budget = max_iter > 0 ? max_iter : std::max<std::uint64_t>(200u * k, 2000u);
    const std::size_t per_cycle =
        std::min<std::size_t>(krylov_subspace_dim(k, 2 * k + 60, n, 0), static_cast<std::size_t>(std::max<std::uint64_t>(budget, k + 1)));
    const std::size_t p_keep = k + std::max<std::size_t>(k / 2, 8);
    const std::size_t fresh = std::max<std::size_t>(std::min(per_cycle, 2 * p_keep + 20), p_keep + 1) - p_keep;
    const std::size_t restarts = static_cast<std::size_t>(std::max<std::uint64_t>(1u, budget / fresh));
    const double scale = scale_or_one(norm_bound);
    KsOptions o;
    o.num_eigs = k;
    o.max_iter = per_cycle;
    o.max_restarts = restarts;
    o.tolerance = kLockRel * scale;
    o.breakdown_tol = kBreakdownRel * scale;
    o.max_subspace_vectors = per_cycle;
    // the start: gaussian_start(n, kKsStartSeed), or the same vector the caller began drawing earlier
    const auto h = start && start->valid() ? start->get() : gaussian_start<Scalar>(n, kKsStartSeed);
    if (h.size() != n) throw Error(kInvalid, "krylov_schur: start vector of the wrong length");
    auto v0 = be.make_zero_vector(n);
    be.copy_from_host(h.data(), v0.get(), n);
    const KsResult r = krylov_schur(be, H, n, v0.get(), o, vectors_out, k);
    BlockSolution sol;
    for (std::size_t i = 0; i < r.eigenvalues.size() && sol.values.size() < k; ++i) sol.values.push_back(r.eigenvalues[i]);
    sol.whole = r.exhausted;
    sol.converged = (r.converged || r.exhausted) && (sol.values.size() >= k || r.exhausted);
    sol.applies = r.applies;
    return sol;
}

}  // namespace qsa

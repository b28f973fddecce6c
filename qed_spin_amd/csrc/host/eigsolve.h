// The block eigensolver lanes on the device backend: the lowest level of a block (k = 1), the
// pruning estimate, and the k lowest levels with their vectors (k > 1). They replace the reference's
// solve_block_lowest / solve_block_eigenpairs lanes (src/engine/block_solve.cpp:263-655) and its
// Krylov-Schur (include/ed/krylov/krylov_schur.h); eigenvalues agree with those lanes and with a
// dense diagonalisation to the lock tolerance, not through the same iteration.
//
// Design:
// - Start vectors are drawn on the device (a counter-based Gaussian: splitmix64 of the seed and
//   the entry index through Box-Muller), so no host thread draws or uploads a start.
// - k = 1: Lanczos without reorthogonalisation (full CGS2 when the run may span the block), its
//   tridiagonal diagonalised every step; done when the lowest Ritz value, or a copy of it that
//   rounding has produced, has a residual bound |beta_j y_j| below 1e-8 s_H.
// - k > 1: thick-restart Lanczos in Krylov-Schur form with locking at every restart. One device
//   block holds the locked vectors and the active basis side by side -- locked vectors fill from
//   the middle of the block downwards, the active basis from the middle upwards -- so a
//   Gram-Schmidt pass against both is one tall-skinny dot_many / axpy_many over one contiguous
//   range of columns. After each cycle the Ritz pairs are computed, the lowest ones whose residual
//   is below the lock tolerance move into the locked range at once, and the next few Ritz vectors
//   are kept as the new active basis (their couplings to the residual vector in the arrow row).
//   When k values are locked, a deflated restart from a fresh start checks that no level below the
//   highest locked one was missed (a degenerate copy a single Krylov space cannot see); a level
//   found below it replaces the highest, and the check repeats.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "lanczos.h"
#include "tridiag.h"

namespace qsa {

inline constexpr double kEps = 2.220446049250313e-16;
inline constexpr double kBreakdownRel = 64.0 * kEps;   ///< a Krylov run breaks down at 64 eps s_H
inline constexpr double kLockRel = 1e-10;              ///< residual that locks a level, times s_H
inline constexpr double kScanRel = 1e-8;               ///< residual that ends the k = 1 scan, times s_H
inline double scale_or_one(double s) { return s > 0.0 ? s : 1.0; }

/// Seeds of the device start vectors (any fixed values: results depend on them only below the
/// tolerances).
inline constexpr std::uint64_t kStartSeed = 0x5EED0001ULL;
inline constexpr std::uint64_t kEstimateStartSeed = 0x5EED0002ULL;
inline constexpr std::uint64_t kFreshSeedBase = 0x5EED1000ULL;

template <class Scalar> void device_gaussian(Scalar* v, std::size_t n, std::uint64_t seed) {
    dev::fill_gaussian(HipBackend<Scalar>::kComplex, v, n, seed, engine_stream());
    QSA_HIP(hipGetLastError());
}

struct BlockSolution {
    std::vector<double> values;
    bool converged = false;
    bool whole = false;   ///< the locked vectors span the block: every level was found
    std::size_t applies = 0;
    LanczosTimings timings;
};

// ---- k = 1 ---------------------------------------------------------------------------------------

/// Whether the lowest Ritz value of the tridiagonal (alpha, beta) has converged: the values within
/// 1e-9 s_H of the lowest are one level (Lanczos without reorthogonalisation repeats a converged
/// value), converged when one of them has |beta_m y_(m-1)| <= tol; `value` gets that one.
/// Ritz values (ascending) of the m x m tridiagonal (alpha[0..m), beta[1..m)) and the squared LAST
/// components of their eigenvectors, in O(m^2): the first components of the same tridiagonal with
/// its order reversed.
inline void ritz_last_weights(const std::vector<double>& alpha, const std::vector<double>& beta, std::size_t m,
                              std::vector<double>& values, std::vector<double>& last2) {
    std::vector<double> a(m), b(m > 0 ? m - 1 : 0);
    for (std::size_t i = 0; i < m; ++i) a[i] = alpha[m - 1 - i];
    for (std::size_t i = 0; i + 1 < m; ++i) b[i] = beta[m - 1 - i];   // couples reversed i and i + 1
    tridiag_weights(a.data(), b.data(), m, values, last2);
}

inline bool lowest_ritz_converged(const std::vector<double>& alpha, const std::vector<double>& beta, double s_H,
                                  double tol, double* value) {
    const std::size_t m = alpha.size();
    if (m == 0) return false;
    std::vector<double> theta, last2;
    ritz_last_weights(alpha, beta, m, theta, last2);
    const double tail = beta.size() > m ? std::abs(beta[m]) : 0.0;
    double best = std::numeric_limits<double>::infinity();
    for (std::size_t i = 0; i < m && theta[i] - theta[0] <= 1e-9 * s_H; ++i) {
        const double bound = tail * std::sqrt(last2[i]);
        if (bound < best) {
            best = bound;
            if (value) *value = theta[i];
        }
    }
    return best <= tol;
}

template <class Scalar, class Matvec>
BlockSolution lowest_scan(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n, std::size_t max_iter,
                          double norm_bound) {
    BlockSolution sol;
    if (n == 0) return sol;
    const double s_H = scale_or_one(norm_bound);
    LanczosOptions<Scalar> o;
    o.max_iter = std::min(n, max_iter);
    o.reorth = n <= o.max_iter ? Reorth::FullCGS2 : Reorth::None;
    o.keep_basis = o.reorth == Reorth::FullCGS2;
    o.breakdown_tol = kBreakdownRel * s_H;
    o.convergence_check = [&](const std::vector<double>& a, const std::vector<double>& b) {
        return a.size() >= 2 && lowest_ritz_converged(a, b, s_H, kScanRel * s_H, nullptr);
    };
    o.convergence_check_interval = 1;
    auto v0 = be.make_zero_vector(n);
    device_gaussian(v0.get(), n, kStartSeed);
    const auto run = lanczos_run(be, H, n, v0.get(), o);
    sol.applies = run.timings.applies;
    sol.timings = run.timings;
    if (run.alpha.empty()) return sol;
    const std::size_t m = run.alpha.size();
    const bool invariant = run.beta.size() > m && run.beta[m] < o.breakdown_tol;   // an exact invariant subspace
    double value = 0.0;
    if (invariant || m == n) {
        value = tridiag_eig(run.alpha, run.beta, m, false).values.front();
        sol.converged = true;
    } else {
        sol.converged = lowest_ritz_converged(run.alpha, run.beta, s_H, kScanRel * s_H, &value);
    }
    sol.values.push_back(value);
    return sol;
}

/// The pruning estimate of a block: 40 Lanczos steps without reorthogonalisation from a fixed
/// start; theta is the lowest Ritz value (an upper bound on the block's lowest level) and residual
/// |beta_m y_(m-1),0| its residual bound. The caller prunes on theta - residual.
struct BlockEstimate {
    double theta = -std::numeric_limits<double>::infinity();
    double residual = 0.0;
    std::size_t applies = 0;
};

template <class Scalar, class Matvec>
BlockEstimate estimate_lowest(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n) {
    BlockEstimate est;
    if (n == 0) return est;
    LanczosOptions<Scalar> o;
    o.max_iter = std::min<std::size_t>(40, n);
    o.reorth = Reorth::None;
    o.keep_basis = false;
    auto v0 = be.make_zero_vector(n);
    device_gaussian(v0.get(), n, kEstimateStartSeed);
    const auto run = lanczos_run(be, H, n, v0.get(), o);
    est.applies = run.timings.applies;
    const std::size_t m = run.alpha.size();
    if (m == 0) return est;
    std::vector<double> theta, last2;
    ritz_last_weights(run.alpha, run.beta, m, theta, last2);
    est.theta = theta.front();
    est.residual = run.beta.size() > m ? std::abs(run.beta[m]) * std::sqrt(last2.front()) : 0.0;
    return est;
}

// ---- k > 1 -------------------------------------------------------------------------------------

/// Thick-restart Lanczos with locking (Krylov-Schur form) for the k lowest levels of a Hermitian
/// operator of dimension n, and their vectors.
template <class Scalar, class Matvec> class LockingLanczos {
public:
    struct Settings {
        std::size_t k = 1;
        std::size_t basis = 0;        ///< active basis size m (0: chosen from k)
        std::size_t budget = 2000;    ///< H applies in all
        double lock_tol = 1e-10;      ///< residual that locks a Ritz pair
        double breakdown_tol = 1e-13;
    };

    LockingLanczos(const HipBackend<Scalar>& be, Matvec& H, std::size_t n, const Settings& s)
        : be_(be), H_(H), n_(n), s_(s) {
        k_ = std::max<std::size_t>(1, std::min(s.k, n));
        m_ = s.basis ? s.basis : std::max<std::size_t>(3 * k_ + 30, 40);
        m_ = std::max<std::size_t>(std::min(m_, n), 1);
        mid_ = k_ + 1;   // locked columns [mid_ - locked_, mid_), active columns [mid_, mid_ + m_ + 1)
        Q_ = be.make_zero_vector(n * (mid_ + m_ + 1));
        w_ = be.make_zero_vector(n);
        keep_ = be.make_zero_vector(n * m_);
    }

    /// Runs to k locked levels (verified against missed degenerate copies) or the apply budget.
    /// Returns the levels ascending; vectors_out (device, k columns of n) gets their vectors.
    BlockSolution solve(const Scalar* start, Scalar* vectors_out) {
        BlockSolution sol;
        be_.copy(start, col(0), n_);
        std::uint64_t fresh = 0;
        bool exhausted = false;
        // main search, resumed from fresh deflated starts until k levels are locked
        while (locked_ < k_ && applies_ < s_.budget) {
            const Search r = search(k_ - locked_);
            if (r == Search::kNullStart) {
                exhausted = true;
                break;
            }
            if (locked_ < k_) device_gaussian(col(0), n_, kFreshSeedBase + fresh++);
        }
        bool converged = locked_ >= k_ || exhausted;
        // verification: a deflated search from a fresh start must find nothing below the top
        for (std::size_t round = 0; converged && !exhausted && round < k_; ++round) {
            const double top = *std::max_element(value_.begin(), value_.end());
            const double gap = 100.0 * s_.lock_tol;
            device_gaussian(col(0), n_, kFreshSeedBase + fresh++);
            const std::size_t before = locked_;
            const Search r = search(1, top - gap);
            if (r == Search::kNullStart) {
                exhausted = true;
                break;
            }
            if (r == Search::kNothingBelow) break;
            if (locked_ == before) {   // the budget ran out before the check converged
                converged = false;
                break;
            }
            const double found = value_.back();
            if (found >= top - gap) {   // nothing new below: the check passes
                drop(locked_ - 1);
                break;
            }
            drop(static_cast<std::size_t>(std::max_element(value_.begin(), value_.end()) - value_.begin()));
        }
        std::vector<std::size_t> order(locked_);
        std::iota(order.begin(), order.end(), std::size_t{0});
        std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) { return value_[a] < value_[b]; });
        for (std::size_t i = 0; i < order.size() && i < k_; ++i) {
            sol.values.push_back(value_[order[i]]);
            if (vectors_out) be_.copy(locked_col(order[i]), vectors_out + i * n_, n_);
        }
        sol.whole = exhausted;
        sol.converged = converged && (sol.values.size() >= k_ || exhausted);
        sol.applies = applies_;
        return sol;
    }

private:
    enum class Search { kDone, kBudget, kNullStart, kNothingBelow };

    Scalar* col(std::size_t j) { return Q_.get() + (mid_ + j) * n_; }
    /// Locked vector i (0 = the first locked) sits just below the active range, i-th from it.
    Scalar* locked_col(std::size_t i) { return Q_.get() + (mid_ - 1 - i) * n_; }
    Scalar* lowest_locked() { return Q_.get() + (mid_ - locked_) * n_; }

    /// Forget locked vector i: the most recent one takes its slot.
    void drop(std::size_t i) {
        const std::size_t last = locked_ - 1;
        if (i != last) {
            be_.copy(locked_col(last), locked_col(i), n_);
            value_[i] = value_[last];
        }
        value_.pop_back();
        --locked_;
    }

    /// Columns [mid_ - locked_, mid_ + j) as one block: the locked vectors and active columns < j.
    std::pair<Scalar, double> orthogonalize(Scalar* x, std::size_t j) {
        const std::size_t cols = locked_ + j;
        if (cols == 0) return {Scalar(0), be_.nrm2(x, n_)};
        return be_.cgs2_tail(lowest_locked(), n_, cols, x, n_);
    }

    /// Lock `want` more levels starting from the vector in active column 0. With a finite give_up,
    /// the search ends after its first cycle when that cycle shows nothing below give_up.
    Search search(std::size_t want, double give_up = std::numeric_limits<double>::infinity()) {
        const std::size_t target = locked_ + want;
        bool first_cycle = true;
        {
            const double nrm = orthogonalize(col(0), 0).second;
            if (nrm < 1e-12) return Search::kNullStart;
            be_.scale(Scalar(1.0 / nrm), col(0), n_);
        }
        std::size_t kept = 0;            // Ritz vectors carried over at the head of the active basis
        std::vector<double> theta, arrow;   // their values and couplings to column `kept`
        while (applies_ < s_.budget) {
            const std::size_t m = std::min(m_, n_ - locked_);
            if (m == 0) return Search::kNullStart;
            // projected matrix, column-major m x m: the kept block diag(theta) with its arrow, then
            // the Lanczos tridiagonal
            std::vector<double> T(m * m, 0.0);
            for (std::size_t i = 0; i < kept; ++i) {
                T[i * m + i] = theta[i];
                T[kept * m + i] = T[i * m + kept] = arrow[i];
            }
            std::size_t size = m;
            double tail = 0.0;
            for (std::size_t j = kept; j < m; ++j) {
                H_(col(j), w_.get());
                ++applies_;
                const auto [diag, beta] = orthogonalize(w_.get(), j + 1);
                T[j * m + j] = std::real(diag);
                if (beta <= s_.breakdown_tol || j + 1 == n_ - locked_) {   // invariant subspace
                    size = j + 1;
                    tail = 0.0;
                    break;
                }
                if (j + 1 < m) T[j * m + j + 1] = T[(j + 1) * m + j] = beta;
                be_.copy(w_.get(), col(j + 1), n_);
                be_.scale(Scalar(1.0 / beta), col(j + 1), n_);
                tail = beta;
            }
            std::vector<double> Ts(size * size);
            for (std::size_t c = 0; c < size; ++c)
                for (std::size_t r = 0; r < size; ++r) Ts[c * size + r] = T[c * m + r];
            const TridiagEig e = symmetric_eig(std::move(Ts), size);
            // lock the lowest Ritz pairs in order while their residual bound is small enough
            std::size_t lock = 0;
            while (locked_ + lock < target && lock < size &&
                   tail * std::abs(e.vectors[lock * size + size - 1]) <= s_.lock_tol)
                ++lock;
            const bool done = locked_ + lock >= target;
            if (!done && first_cycle && give_up < std::numeric_limits<double>::infinity()) {
                // a check for a missed level: after one cycle the lowest Ritz value, less its
                // residual bound, is still at or above the level to beat -- nothing lies below
                const double r0 = tail * std::abs(e.vectors[size - 1]);
                if (lock == 0 && e.values[0] - r0 >= give_up) return Search::kNothingBelow;
            }
            first_cycle = false;
            const std::size_t remaining = target - (locked_ + lock);
            const std::size_t carry =
                done || size < lock + 2 ? 0 : std::min(std::max<std::size_t>(remaining + 8, 10), size - lock - 1);
            rotate(e, size, lock, carry);
            for (std::size_t i = 0; i < lock; ++i) value_.push_back(e.values[i]);
            locked_ += lock;
            if (done) return Search::kDone;
            if (tail == 0.0) {
                // the space was invariant: the remaining Ritz vectors are exact but higher; go on
                // from a fresh start orthogonal to everything locked
                return locked_ < target ? restart_fresh(target) : Search::kDone;
            }
            kept = std::min(carry, size - lock);
            theta.assign(e.values.begin() + static_cast<std::ptrdiff_t>(lock),
                         e.values.begin() + static_cast<std::ptrdiff_t>(lock + kept));
            arrow.resize(kept);
            for (std::size_t i = 0; i < kept; ++i) arrow[i] = tail * e.vectors[(lock + i) * size + size - 1];
            be_.copy(col(size), col(kept), n_);   // the residual direction follows the kept vectors
        }
        return Search::kBudget;
    }

    Search restart_fresh(std::size_t target) {
        device_gaussian(col(0), n_, kFreshSeedBase + 0x100 + locked_);
        return search(target - locked_);
    }

    /// Ritz vectors of the active columns [0, size): the first `lock` into the locked range, the
    /// next `carry` to the head of the active basis.
    void rotate(const TridiagEig& e, std::size_t size, std::size_t lock, std::size_t carry) {
        const std::size_t cols = lock + carry;
        if (cols == 0) return;
        std::vector<double> Y(e.vectors.begin(), e.vectors.begin() + static_cast<std::ptrdiff_t>(cols * size));
        DeviceBuffer Yd = upload(Y.data(), Y.size());
        // locked vectors: slots locked_ .. locked_ + lock - 1 lie below the active range in reverse
        // order, so each goes through the carry buffer
        dev::block_rotate(HipBackend<Scalar>::kComplex, col(0), n_, size, Yd.get<double>(), cols, keep_.get(), n_,
                          engine_stream());
        QSA_HIP(hipGetLastError());
        for (std::size_t i = 0; i < lock; ++i) be_.copy(keep_.get() + i * n_, locked_col(locked_ + i), n_);
        if (carry) be_.copy(keep_.get() + lock * n_, col(0), n_ * carry);
        be_.synchronize();
    }

    const HipBackend<Scalar>& be_;
    Matvec& H_;
    std::size_t n_, k_ = 1, m_ = 1, mid_ = 2, locked_ = 0, applies_ = 0;
    Settings s_;
    typename HipBackend<Scalar>::UniqueVec Q_, w_, keep_;
    std::vector<double> value_;
};

/// The lowest k levels of one block with the lane's budgets: at most max(200 k, 2000) applies
/// (max_iter when given), lock tolerance 1e-10 s_H, breakdown 64 eps s_H.
template <class Scalar, class Matvec>
BlockSolution krylov_schur_lane(const HipBackend<Scalar>& be, Matvec&& H, std::size_t n, std::size_t k,
                                double norm_bound, std::uint64_t max_iter, Scalar* vectors_out = nullptr) {
    BlockSolution sol;
    if (n == 0) {
        sol.converged = sol.whole = true;
        return sol;
    }
    const double s_H = scale_or_one(norm_bound);
    typename LockingLanczos<Scalar, Matvec>::Settings st;
    st.k = std::min(k, n);
    st.budget = static_cast<std::size_t>(max_iter > 0 ? max_iter : std::max<std::uint64_t>(200u * k, 2000u));
    st.lock_tol = kLockRel * s_H;
    st.breakdown_tol = kBreakdownRel * s_H;
    LockingLanczos<Scalar, Matvec> solver(be, H, n, st);
    auto v0 = be.make_zero_vector(n);
    device_gaussian(v0.get(), n, kStartSeed);
    return solver.solve(v0.get(), vectors_out);
}

}  // namespace qsa

// The Lanczos recurrence on device vectors: the reference's single Lanczos algorithm
// (include/ed/krylov/lanczos.h:266-576) with its three reorthogonalisation policies, breakdown rule,
// convergence hook and per-vector hook, on the HIP backend. The kept basis is one contiguous
// column block in HBM, so a full Gram-Schmidt pass is one tall-skinny dot kernel and one update
// kernel over it (each streaming the block once), with the norms before and after folded in.
#pragma once

#include <chrono>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "backend.h"

namespace qsa {

enum class Reorth : int { None = 0, FullCGS2 = 1, LocalDGKS3 = 2 };

/// DGKS "twice is enough": a second pass only when the first removed more than 1 - 1/sqrt 2 of ||w||.
inline constexpr double kDgksKappa = 0.7071067811865476;

template <class Scalar> struct LanczosOptions {
    std::size_t max_iter = 100;
    double breakdown_tol = 1e-300;
    Reorth reorth = Reorth::FullCGS2;
    bool keep_basis = true;
    std::function<bool(const std::vector<double>&, const std::vector<double>&)> convergence_check;
    std::size_t convergence_check_interval = 0;
    std::function<void(std::size_t, const Scalar*)> on_vector;
    double local_ortho_threshold = 1.49011611938476562e-08;
    /// Optional caller-owned basis block (capacity columns of length n); its first n_aux columns
    /// are fixed vectors every Gram-Schmidt pass also projects out (Krylov-Schur's locked set).
    Scalar* basis_block = nullptr;
    std::size_t basis_capacity = 0;
    std::size_t n_aux = 0;
};

struct LanczosTimings {
    double apply_s = 0.0, reorth_s = 0.0, vector_s = 0.0, total_s = 0.0;
    std::size_t applies = 0;
};

template <class Scalar> struct LanczosResult {
    std::vector<double> alpha;   ///< iters
    std::vector<double> beta;    ///< iters + 1, beta[0] = 0
    std::size_t iters = 0;
    typename HipBackend<Scalar>::UniqueVec basis_owned;
    Scalar* basis = nullptr;     ///< column j of the Krylov basis at basis + j n (kept basis only)
    std::size_t basis_cols = 0;
    LanczosTimings timings;
};

/// A matvec that can also run a step's first pass fused (first_pass(v, v_prev, beta, w, alpha):
/// w = H v - beta v_prev, alpha = <v, w>, false when it cannot this time).
template <class M, class S, class = void> struct HasFirstPass : std::false_type {};
template <class M, class S>
struct HasFirstPass<M, S,
                    std::void_t<decltype(std::declval<const M&>().first_pass(
                        static_cast<const S*>(nullptr), static_cast<const S*>(nullptr), 0.0, static_cast<S*>(nullptr),
                        std::declval<S&>()))>> : std::true_type {};

template <class Scalar, class Matvec>
LanczosResult<Scalar> lanczos_run(const HipBackend<Scalar>& be, Matvec&& matvec, std::size_t n, const Scalar* v0,
                                  const LanczosOptions<Scalar>& o) {
    using clock = std::chrono::steady_clock;
    const auto t_start = clock::now();
    LanczosResult<Scalar> R;
    const bool cgs = o.reorth == Reorth::FullCGS2;
    if (cgs && !o.keep_basis) throw Error(kInvalid, "lanczos: FullCGS2 requires keep_basis");
    if (o.max_iter == 0) return R;
    auto v_prev = be.make_zero_vector(n);
    auto v_curr = be.make_zero_vector(n);
    auto w = be.make_zero_vector(n);
    const std::size_t cap = std::min<std::size_t>(o.max_iter, n);
    Scalar* block = nullptr;
    std::size_t aux = 0;
    if (o.keep_basis) {
        if (o.basis_block) {
            if (o.basis_capacity < o.n_aux + cap) throw Error(kInvalid, "lanczos: basis block too small");
            block = o.basis_block;
            aux = o.n_aux;
        } else {
            R.basis_owned.reset(be.allocate(n * cap));
            block = R.basis_owned.get();
        }
    }
    R.beta.push_back(0.0);
    be.copy(v0, v_curr.get(), n);
    const double n0 = be.nrm2(v_curr.get(), n);
    if (!(n0 > 0.0)) throw Error(kInvalid, "lanczos: initial vector has non-positive norm");
    be.scale(Scalar(1.0 / n0), v_curr.get(), n);
    std::size_t kept = 0;
    if (block) {
        be.copy(v_curr.get(), block + (aux + kept) * n, n);
        ++kept;
    }
    const auto secs = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    for (std::size_t j = 0; j < cap; ++j) {
        if (o.on_vector) o.on_vector(j, v_curr.get());
        const auto t0 = clock::now();
        // the apply, and alpha_j = <v_j, H v_j - beta_j v_{j-1}>: in the apply's epilogue when the
        // matvec offers it, else apply, then axpy_dot
        Scalar aj{};
        bool fused = false;
        if constexpr (HasFirstPass<std::decay_t<Matvec>, Scalar>::value)
            fused = matvec.first_pass(v_curr.get(), j > 0 ? v_prev.get() : nullptr, R.beta[j], w.get(), aj);
        if (!fused) {
            matvec(v_curr.get(), w.get());
            be.synchronize();
        }
        const auto t1 = clock::now();
        R.timings.apply_s += secs(t0, t1);
        ++R.timings.applies;
        if (!fused)
            aj = (j > 0) ? be.axpy_dot(Scalar(-R.beta[j]), v_prev.get(), w.get(), v_curr.get(), n)
                         : be.dot(v_curr.get(), w.get(), n);
        R.alpha.push_back(std::real(aj));
        const bool local = o.reorth == Reorth::LocalDGKS3 && j > 0;
        Scalar overlap{};
        if (local) overlap = be.axpy_dot(Scalar(-R.alpha[j]), v_curr.get(), w.get(), v_curr.get(), n);
        else be.axpy(Scalar(-R.alpha[j]), v_curr.get(), w.get(), n);
        bool defer = false;
        Scalar defer_coef{};
        const auto t2 = clock::now();
        R.timings.vector_s += secs(t1, t2);
        if (cgs && aux + kept > 0) {
            const auto [n_before, n_after] = be.cgs_pass(block, n, aux + kept, w.get(), n);
            if (n_after < kDgksKappa * n_before) (void)be.cgs_pass(block, n, aux + kept, w.get(), n);
        } else if (local) {
            if (std::abs(overlap) > o.local_ortho_threshold) {
                defer = true;
                defer_coef = -overlap;
            }
        }
        const auto t3 = clock::now();
        R.timings.reorth_s += secs(t2, t3);
        const double bnext = defer ? be.axpy_nrm2(defer_coef, v_curr.get(), w.get(), n) : be.nrm2(w.get(), n);
        R.beta.push_back(bnext);
        if (bnext < o.breakdown_tol) {
            R.timings.vector_s += secs(t3, clock::now());
            break;
        }
        be.scale(Scalar(1.0 / bnext), w.get(), n);
        std::swap(v_prev, v_curr);
        std::swap(v_curr, w);
        R.timings.vector_s += secs(t3, clock::now());
        if (o.convergence_check && o.convergence_check_interval > 0 && ((j + 1) % o.convergence_check_interval == 0))
            if (o.convergence_check(R.alpha, R.beta)) break;
        if (j + 1 >= cap) break;
        if (block) {
            be.copy(v_curr.get(), block + (aux + kept) * n, n);
            ++kept;
        }
    }
    be.synchronize();
    R.iters = R.alpha.size();
    if (block) {
        R.basis = block + aux * n;
        R.basis_cols = kept;
    }
    R.timings.total_s = secs(t_start, clock::now());
    return R;
}

}  // namespace qsa

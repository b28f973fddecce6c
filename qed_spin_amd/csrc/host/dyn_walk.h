// Finite-temperature dynamical correlations S_{AA}(omega, T) by FTLM over a whole walk of source
// sectors in the batched engine: the reference's ftlm_dynamics_kernel (include/ed/dynamics/
// ftlm_dynamics.h:110-397) as the dynamics verb runs it per source for T > 0
// (src/engine/dynamics.cpp:693-1257), restated for a device that holds many sources at once.
//
// For each sample r of a source (drawn by the reference's seed recipe) and each channel -- a probe A
// from the source to a target sector -- the estimator is
//   S(omega, T) = sum_i e^{-beta e_i} <r|psi_i> sum_j <psi_i|A^dag|phi_j> <phi_j|A|r> L_eta(omega - (l_j - e_i)),
//   Z(T)        = sum_i e^{-beta e_i} |<r|psi_i>|^2,
// with (e_i, psi_i) the Ritz pairs of the source Lanczos run from r and (l_j, phi_j) those of the
// target run from A r / |A r|; the samples of a source are combined about their lowest weighted
// Ritz value and scaled by dim_src / samples.
//
// Design: every source of a wave and all its samples run through one packed multi-vector Lanczos
// with kept bases (BatchedLanczos::run_kept), and so do all channels' target runs, started from the
// packed probe applied to the sources' start vectors. The overlaps <A v_a|w_b> of every pair are
// products of the kept bases laid out as column panels, split over chunks of 8192 rows so that a
// pair of 1e6 rows fills the device as well as many small pairs do: every chunk of every pair in
// one launch (64 x 64 tiles on the FP64 matrix cores, LDS-staged slices, the tiles of a chunk on one
// XCD), then the chunks of a pair summed in order. (For these shapes -- 40..200 vectors against 40..200, 1e2..1e6 rows -- the
// library's FP64 GEMM, batched per pair, measured slower.) Only the small tridiagonal algebra and
// the Lorentzian sums run on the host. Waves are cut so that the kept vectors fit a device-memory
// budget (by default two fifths of the device's memory, at most nine tenths of what is free).
#pragma once

#include <memory>
#include <vector>

#include "batched.h"

namespace qsa {

struct DynChannel {
    int source = 0;                               ///< index of the source sector
    std::shared_ptr<const SectorOperator> H;      ///< H on the target sector (square, CSR lane)
    std::shared_ptr<const SectorOperator> A;      ///< the probe, source sector -> target sector (CSR lane)
};

struct DynWalkOptions {
    int samples = 30, krylov = 200;
    std::vector<std::uint64_t> source_seeds;     ///< per source: its base seed
    std::vector<double> breakdown_tol;           ///< per source; a channel's target runs use its source's
    std::vector<double> temperatures, omega;
    double eta = 0.05;
    std::uint64_t budget_bytes = 0;              ///< device memory for the kept vectors (0: 2/5 of the device's, at most 9/10 of the free)
};

struct DynWalkResult {
    std::vector<std::vector<double>> Z;          ///< [source][T], times dim_src / samples done
    std::vector<std::vector<cplx>> S;            ///< [channel][T nW + w], the same scaling
    std::vector<double> e_min;                   ///< [source] the samples' common reference
    std::vector<int> samples_done;               ///< [source]
    double lanczos_s = 0.0, overlap_s = 0.0, host_s = 0.0, total_s = 0.0, setup_s = 0.0;
    int waves = 0;
};

/// The smallest sector whose Lanczos runs reorthogonalise locally rather than fully: sectors up to
/// max(4 krylov, 2^16) states take FullCGS2 (the reference's rule, ftlm_dynamics.h:52).
inline constexpr std::size_t kFullReorthMaxDim = std::size_t{1} << 16;

DynWalkResult ftlm_dynamics_walk(const std::vector<std::shared_ptr<const SectorOperator>>& sources,
                                 const std::vector<DynChannel>& channels, const DynWalkOptions& o);

}  // namespace qsa

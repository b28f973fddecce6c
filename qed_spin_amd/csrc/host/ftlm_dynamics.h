// Finite-temperature dynamical correlations S_{A B}(omega, T) by FTLM across sectors (the reference's
// ftlm_dynamics_kernel, include/ed/dynamics/ftlm_dynamics.h:110-397, as the dynamics verb drives it
// for T > 0, src/engine/dynamics.cpp:693-1257) on the HIP backend, for the autocorrelation A = B.
//
// Per source sector and sample r (seeded as the reference seeds it): a Lanczos run on H_src from r
// with its basis kept in HBM (Ritz values e_i, eigenvector matrix VH, first components c_i); for
// each target sector the probe reaches, a Lanczos run on H_dst from B r / |B r| (Ritz values
// lambda_j, VS), the overlaps W[a, b] = <A v_a | w_b> (A applied to every source basis vector, then
// one tall-skinny product against the target basis, both on the device), and on the host
//   S_s(omega, T) = sum_i e^{-beta (e_i - smin)} c_i sum_j (sum_b (VH W)_{ib} VS_{jb}) |B r| VS_{j0}
//                   L_eta(omega - (lambda_j - e_i)),   Z_s(T) = sum_i e^{-beta (e_i - smin)} c_i^2.
// The samples are combined about the lowest weighted Ritz value and scaled by dim_src / samples;
// the caller sums sources and divides S by Z.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "lanczos.h"
#include "sector_operator.h"
#include "thermo.h"
#include "tridiag.h"

namespace qsa {

/// The smallest sector whose Lanczos runs reorthogonalise locally (kLocalReorthMinDim of
/// include/ed/dynamics/ftlm_dynamics.h:52); below it, and up to 4 Krylov depths, full CGS2.
inline constexpr std::size_t kLocalReorthMinDimDyn = std::size_t{1} << 16;

struct FtlmDynTarget {
    const SectorOperator* H = nullptr;   ///< H on the target sector
    const SectorOperator* A = nullptr;   ///< the probe from the source to the target
};

struct FtlmDynOptions {
    std::size_t krylov = 200, samples = 30;
    double breakdown_tol = 0.0;
    double eta = 0.05;
    std::uint64_t seed = 1;   ///< the source's base seed (block_seed of the walk's seed and its key)
    std::vector<double> temperatures, omega;
};

struct FtlmDynResult {
    std::vector<double> Z;                 ///< [T], times dim_src
    std::vector<std::vector<cplx>> S;      ///< [target][T nW + w], times dim_src
    double E_min = 0.0;
    std::size_t samples_done = 0;
};

inline FtlmDynResult ftlm_dynamics_source(const SectorOperator& Hsrc, const std::vector<FtlmDynTarget>& targets,
                                          const FtlmDynOptions& o) {
    const std::size_t n = static_cast<std::size_t>(Hsrc.dim());
    if (n == 0) throw Error(kInvalid, "ftlm dynamics: empty source sector");
    if (o.temperatures.empty() || o.omega.empty() || o.samples == 0)
        throw Error(kInvalid, "ftlm dynamics: no temperatures, frequencies or samples");
    constexpr double kInvPi = 0.3183098861837907;
    const std::size_t nW = o.omega.size(), nT = o.temperatures.size(), nt = targets.size();
    HipBackend<cplx> be;
    auto run = [&](const SectorOperator& H, const cplx* v0, std::size_t dim) {
        LanczosOptions<cplx> lo;
        lo.max_iter = std::min(dim, o.krylov);
        lo.reorth = dim > std::max<std::size_t>(4 * o.krylov, kLocalReorthMinDimDyn) ? Reorth::LocalDGKS3
                                                                                    : Reorth::FullCGS2;
        lo.keep_basis = true;
        if (o.breakdown_tol > 0.0) lo.breakdown_tol = o.breakdown_tol;
        return lanczos_run(be, [&](const cplx* in, cplx* out) { H.apply(in, out, true); }, dim, v0, lo);
    };
    struct Sample {
        bool ok = false;
        double smin = 0.0;
        std::vector<double> Z;
        std::vector<std::vector<cplx>> S;
    };
    std::vector<Sample> samples(o.samples);
    auto r = be.make_zero_vector(n);
    for (std::size_t s = 0; s < o.samples; ++s) {
        Sample& out = samples[s];
        std::mt19937 gen = sample_engine(o.seed, s);
        const std::vector<cplx> rh = gaussian_vector(n, gen);
        be.copy_from_host(rh.data(), r.get(), n);
        auto kh = run(Hsrc, r.get(), n);
        if (kh.alpha.empty() || kh.basis_cols < kh.alpha.size()) continue;
        const std::size_t mH = kh.alpha.size();
        const TridiagEig th = tridiag_eig(kh.alpha, kh.beta, mH, true);
        const std::vector<double>& ritz = th.values;
        std::vector<double> c(mH);
        for (std::size_t i = 0; i < mH; ++i) c[i] = th.z(0, i);
        out.smin = std::numeric_limits<double>::infinity();
        for (std::size_t i = 0; i < mH; ++i)
            if (c[i] != 0.0) out.smin = std::min(out.smin, ritz[i]);
        if (!std::isfinite(out.smin)) out.smin = *std::min_element(ritz.begin(), ritz.end());
        std::vector<double> wt(nT * mH);
        out.Z.assign(nT, 0.0);
        for (std::size_t it = 0; it < nT; ++it) {
            const double beta = 1.0 / o.temperatures[it];
            for (std::size_t i = 0; i < mH; ++i) {
                wt[it * mH + i] = std::exp(-beta * (ritz[i] - out.smin));
                out.Z[it] += c[i] * c[i] * wt[it * mH + i];
            }
        }
        out.ok = true;
        out.S.assign(nt, std::vector<cplx>(nT * nW, cplx(0.0, 0.0)));
        for (std::size_t tt = 0; tt < nt; ++tt) {
            const SectorOperator& Ht = *targets[tt].H;
            const SectorOperator& A = *targets[tt].A;
            const std::size_t nd = static_cast<std::size_t>(Ht.dim());
            auto phi = be.make_zero_vector(nd);
            A.apply(r.get(), phi.get(), true);
            const double nphi = be.nrm2(phi.get(), nd);
            if (nphi < 1e-14) continue;   // B annihilates r: no weight in this target
            be.scale(cplx(1.0 / nphi, 0.0), phi.get(), nd);
            auto ks = run(Ht, phi.get(), nd);
            if (ks.alpha.empty() || ks.basis_cols < ks.alpha.size()) continue;
            const std::size_t mS = ks.alpha.size();
            const TridiagEig ts = tridiag_eig(ks.alpha, ks.beta, mS, true);
            // W[a + b mH] = <A v_a | w_b>: A on each source basis vector, then its overlaps with the
            // target basis (one tall-skinny product per row a)
            std::vector<cplx> W(mH * mS), row(mS);
            auto ov = be.make_zero_vector(nd);
            for (std::size_t a = 0; a < mH; ++a) {
                A.apply(kh.basis + a * n, ov.get(), true);
                be.dot_many_block(ks.basis, nd, mS, ov.get(), nd, row.data());
                for (std::size_t b = 0; b < mS; ++b) W[a + b * mH] = std::conj(row[b]);
            }
            // Tm[i, b] = sum_a VH[a, i] W[a, b] (column i of the eigenvector matrix is Ritz vector i)
            std::vector<cplx> Tm(mH * mS, cplx(0.0, 0.0));
            for (std::size_t i = 0; i < mH; ++i)
                for (std::size_t b = 0; b < mS; ++b) {
                    cplx acc(0.0, 0.0);
                    for (std::size_t a = 0; a < mH; ++a) acc += th.z(a, i) * W[a + b * mH];
                    Tm[i * mS + b] = acc;
                }
            auto& St = out.S[tt];
            std::vector<cplx> si(nW);
            for (std::size_t i = 0; i < mH; ++i) {
                if (c[i] == 0.0) continue;
                std::fill(si.begin(), si.end(), cplx(0.0, 0.0));
                for (std::size_t j = 0; j < mS; ++j) {
                    cplx obar(0.0, 0.0);
                    for (std::size_t b = 0; b < mS; ++b) obar += Tm[i * mS + b] * ts.z(b, j);
                    const cplx w_ij = c[i] * obar * (nphi * ts.z(0, j));
                    if (std::abs(w_ij) < 1e-300) continue;
                    const double E_ij = ts.values[j] - ritz[i];
                    for (std::size_t iw = 0; iw < nW; ++iw) {
                        const double d = o.omega[iw] - E_ij;
                        si[iw] += w_ij * ((o.eta * kInvPi) / (d * d + o.eta * o.eta));
                    }
                }
                for (std::size_t it = 0; it < nT; ++it) {
                    const double wi = wt[it * mH + i];
                    if (wi < 1e-300) continue;
                    for (std::size_t iw = 0; iw < nW; ++iw) St[it * nW + iw] += wi * si[iw];
                }
            }
        }
    }
    // the samples about the common reference, in sample order
    FtlmDynResult R;
    R.Z.assign(nT, 0.0);
    R.S.assign(nt, std::vector<cplx>(nT * nW, cplx(0.0, 0.0)));
    double E_min = std::numeric_limits<double>::infinity();
    for (const Sample& smp : samples)
        if (smp.ok) E_min = std::min(E_min, smp.smin);
    for (const Sample& smp : samples) {
        if (!smp.ok) continue;
        for (std::size_t it = 0; it < nT; ++it) {
            const double f = std::exp(-(1.0 / o.temperatures[it]) * (smp.smin - E_min));
            R.Z[it] += f * smp.Z[it];
            for (std::size_t tt = 0; tt < nt; ++tt)
                for (std::size_t iw = 0; iw < nW; ++iw) R.S[tt][it * nW + iw] += f * smp.S[tt][it * nW + iw];
        }
        ++R.samples_done;
    }
    if (R.samples_done > 0) {
        const double scale = static_cast<double>(n) / static_cast<double>(R.samples_done);
        for (auto& z : R.Z) z *= scale;
        for (auto& St : R.S)
            for (auto& v : St) v *= scale;
    }
    R.E_min = std::isfinite(E_min) ? E_min : 0.0;
    return R;
}

}  // namespace qsa

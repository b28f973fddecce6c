// qedh_batch_lanczos, qedh_ftlm_batched: the batched Krylov engine behind the C-ABI.
#include "../../../include/qedh.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "batched.h"
#include "dense_device.h"
#include "dyn_walk.h"
#include "serve.h"
#include "tridiag.h"

struct qedh_matvec_s {
    std::shared_ptr<qsa::SectorOperator> op;
};

namespace qsa {
int capi_status(const std::exception& e);
void capi_set_error(const std::string& m);
}  // namespace qsa

namespace {

template <class F> int guarded_batch(F&& f) {
    try {
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

std::vector<std::shared_ptr<const qsa::SectorOperator>> gather_blocks(const qedh_matvec* blocks, int n) {
    if (!blocks || n < 1) throw qsa::Error(qsa::kInvalid, "batched: no blocks");
    std::vector<std::shared_ptr<const qsa::SectorOperator>> out;
    out.reserve(static_cast<std::size_t>(n));
    for (int b = 0; b < n; ++b) {
        if (!blocks[b]) throw qsa::Error(qsa::kInvalid, "batched: block " + std::to_string(b) + " is null");
        out.push_back(blocks[b]->op);
    }
    return out;
}

}  // namespace

extern "C" int qedh_batch_lanczos(const qedh_matvec* blocks, int n_blocks, int R, int krylov, const double* start,
                                  const double* breakdown_tol, int reorth, int use_graph, double* alpha, double* beta,
                                  int* steps, double* timings) {
    return guarded_batch([&] {
        if (!start || !breakdown_tol || !alpha || !beta || !steps)
            throw qsa::Error(qsa::kInvalid, "qedh_batch_lanczos: null argument");
        const auto ops = gather_blocks(blocks, n_blocks);
        qsa::BatchedTimings t;
        const auto t0 = std::chrono::steady_clock::now();
        qsa::BatchedLanczos L(ops, R);
        t.pack_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        L.set_start(reinterpret_cast<const qsa::cplx*>(start));
        std::vector<double> a, b;
        std::vector<int> st;
        L.run(krylov, std::vector<double>(breakdown_tol, breakdown_tol + n_blocks), reorth != 0, use_graph != 0, a, b, st,
              t);
        std::copy(a.begin(), a.end(), alpha);
        std::copy(b.begin(), b.end(), beta);
        std::copy(st.begin(), st.end(), steps);
        if (timings) {
            timings[0] = t.pack_s;
            timings[1] = t.steps_s;
            timings[2] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}

extern "C" int qedh_ftlm_batched(const qedh_matvec* blocks, int n_blocks, int samples, int krylov,
                                 const uint64_t* block_seeds, const double* breakdown_tol, int n_beta,
                                 const double* betas, int use_graph, double* lnZ, double* E, double* V, double* e_min,
                                 double* stats) {
    return guarded_batch([&] {
        if (!block_seeds || !betas || !lnZ || !E || !V || n_beta < 1)
            throw qsa::Error(qsa::kInvalid, "qedh_ftlm_batched: null argument or no temperature");
        const auto ops = gather_blocks(blocks, n_blocks);
        qsa::BatchedFtlmOptions o;
        o.samples = samples;
        o.krylov = krylov;
        o.betas.assign(betas, betas + n_beta);
        o.block_seeds.assign(block_seeds, block_seeds + n_blocks);
        o.use_graph = use_graph != 0;
        for (int b = 0; b < n_blocks; ++b) {
            double tol = 0.0;
            if (breakdown_tol) {
                tol = breakdown_tol[b];
            } else {
                const double s = ops[b]->norm_bound();
                tol = 64.0 * std::numeric_limits<double>::epsilon() * (s > 0.0 ? s : 1.0);
            }
            o.breakdown_tol.push_back(tol);
        }
        qsa::BatchedTimings t;
        const auto c = qsa::batched_ftlm(ops, o, t);
        for (int b = 0; b < n_blocks; ++b) {
            for (int i = 0; i < n_beta; ++i) {
                lnZ[b * n_beta + i] = c[b].lnZ[i];
                E[b * n_beta + i] = c[b].E[i];
                V[b * n_beta + i] = c[b].V[i];
            }
            if (e_min) e_min[b] = c[b].e_min;
        }
        if (stats) {
            stats[0] = t.pack_s;
            stats[1] = t.seed_s;
            stats[2] = t.steps_s;
            stats[3] = t.post_s;
            stats[4] = t.total_s;
            stats[5] = static_cast<double>(t.applies);
            stats[6] = static_cast<double>(t.rows);
            stats[7] = static_cast<double>(t.nnz);
            stats[8] = static_cast<double>(t.schedule);
            stats[9] = t.tune_s;
        }
    });
}

/* Exact thermodynamics of small blocks: each block's CSR is expanded to a dense Hermitian matrix on
 * the host and its whole spectrum solved (the reference's host dense lane for blocks at most
 * dense_max_dim, src/engine/thermal.cpp:197-200), blocks in parallel on the host threads. */
namespace {

/// Every block's whole spectrum, ascending. Blocks of at least device_min_dim states (and every
/// block above the host limit of 8192) on the device lane (rocSOLVER, batched per dimension); the
/// others expanded to dense Hermitian matrices on this thread (device calls stay on one thread)
/// and solved on the host pool.
std::vector<std::vector<double>> dense_spectra(const std::vector<std::shared_ptr<const qsa::SectorOperator>>& ops,
                                               const char* who, std::int64_t device_min_dim) {
    std::vector<const qsa::SectorOperator*> dev_ops;
    std::vector<std::size_t> dev_idx, host_idx;
    for (std::size_t b = 0; b < ops.size(); ++b) {
        const auto& op = ops[b];
        if (!op->same()) throw qsa::Error(qsa::kInvalid, std::string(who) + ": a block must be square");
        if (op->lane() != qsa::kLaneCsr)
            throw qsa::Error(qsa::kDeviceUnsupported, std::string(who) + ": a block needs its CSR (csr lane)");
        if (op->dim() > qsa::kDenseDeviceMaxDim)
            throw qsa::Error(qsa::kInvalid, std::string(who) + ": block above " +
                                                std::to_string(qsa::kDenseDeviceMaxDim) + " states");
        if (op->dim() >= device_min_dim || op->dim() > 8192) {
            dev_ops.push_back(op.get());
            dev_idx.push_back(b);
        } else {
            host_idx.push_back(b);
        }
    }
    std::vector<std::vector<double>> ev(ops.size());
    if (!dev_ops.empty()) {
        auto d = qsa::device_spectra(dev_ops);
        for (std::size_t q = 0; q < dev_idx.size(); ++q) ev[dev_idx[q]] = std::move(d[q]);
    }
    std::vector<std::vector<std::complex<double>>> dense(host_idx.size());
    for (std::size_t q = 0; q < host_idx.size(); ++q) {
        const auto& op = ops[host_idx[q]];
        std::vector<std::uint64_t> rp;
        std::vector<std::uint32_t> col;
        std::vector<double> val;
        op->csr_download(rp, col, val);
        const std::size_t n = static_cast<std::size_t>(op->dim());
        dense[q].assign(n * n, std::complex<double>(0.0, 0.0));
        for (std::size_t r = 0; r < n; ++r)
            for (std::uint64_t e = rp[r]; e < rp[r + 1]; ++e)
                dense[q][col[e] * n + r] += std::complex<double>(val[2 * e], val[2 * e + 1]);
    }
    qsa::parallel_blocks(host_idx.size(), [&](std::size_t q) {
        const std::size_t n = static_cast<std::size_t>(ops[host_idx[q]]->dim());
        ev[host_idx[q]] = qsa::hermitian_eigvals_householder(std::move(dense[q]), n);
    });
    return ev;
}

}  // namespace

extern "C" int qedh_block_spectrum(const qedh_matvec* blocks, int n_blocks, int64_t device_min_dim, double* values) {
    return guarded_batch([&] {
        if (!values) throw qsa::Error(qsa::kInvalid, "qedh_block_spectrum: null argument");
        const auto ev = dense_spectra(gather_blocks(blocks, n_blocks), "qedh_block_spectrum",
                                      device_min_dim < 0 ? qsa::kDeviceDenseMinDim : device_min_dim);
        std::size_t at = 0;
        for (const auto& e : ev) {
            std::copy(e.begin(), e.end(), values + at);
            at += e.size();
        }
    });
}

extern "C" int qedh_exact_thermal(const qedh_matvec* blocks, int n_blocks, int n_beta, const double* betas,
                                  double* lnZ, double* E, double* V, double* e_min) {
    return guarded_batch([&] {
        if (!betas || !lnZ || !E || !V || n_beta < 1)
            throw qsa::Error(qsa::kInvalid, "qedh_exact_thermal: null argument or no temperature");
        const auto ops = gather_blocks(blocks, n_blocks);
        const auto ev = dense_spectra(ops, "qedh_exact_thermal", qsa::kDeviceDenseMinDim);
        const std::vector<double> bt(betas, betas + n_beta);
        std::vector<qsa::Curves> out(ops.size());
        qsa::parallel_blocks(ops.size(), [&](std::size_t b) { out[b] = qsa::exact_curves(ev[b], bt); });
        for (int b = 0; b < n_blocks; ++b) {
            for (int i = 0; i < n_beta; ++i) {
                lnZ[b * n_beta + i] = out[b].lnZ[i];
                E[b * n_beta + i] = out[b].E[i];
                V[b * n_beta + i] = out[b].V[i];
            }
            if (e_min) e_min[b] = out[b].e_min;
        }
    });
}

/* The system's curves from its blocks' (log-sum-exp over weighted blocks, law of total variance):
 * src/engine/thermal.cpp:597-640. Inputs [n_blocks][n_beta]; outputs [n_beta]. */
extern "C" int qedh_thermal_combine(int n_blocks, int n_beta, const double* betas, const double* weights,
                                    const double* lnZ, const double* E, const double* V, double* out_lnZ,
                                    double* out_E, double* out_C, double* out_S, double* out_F) {
    return guarded_batch([&] {
        if (n_blocks < 1 || n_beta < 1 || !betas || !weights || !lnZ || !E || !V || !out_lnZ || !out_E || !out_C ||
            !out_S || !out_F)
            throw qsa::Error(qsa::kInvalid, "qedh_thermal_combine: bad arguments");
        std::vector<qsa::Curves> blocks(static_cast<std::size_t>(n_blocks));
        for (int b = 0; b < n_blocks; ++b) {
            blocks[b].lnZ.assign(lnZ + b * n_beta, lnZ + (b + 1) * n_beta);
            blocks[b].E.assign(E + b * n_beta, E + (b + 1) * n_beta);
            blocks[b].V.assign(V + b * n_beta, V + (b + 1) * n_beta);
        }
        const auto s = qsa::combine_blocks(blocks, std::vector<double>(weights, weights + n_blocks),
                                           std::vector<double>(betas, betas + n_beta));
        for (int i = 0; i < n_beta; ++i) {
            out_lnZ[i] = s.lnZ[i];
            out_E[i] = s.E[i];
            out_C[i] = s.C[i];
            out_S[i] = s.S[i];
            out_F[i] = s.F[i];
        }
    });
}

/* Continued fractions of many (block, vector) pairs at once (cf_spectral_from_vector,
 * include/ed/dynamics/cf.h:103-161, for every probe image the dynamics verb forms,
 * src/engine/dynamics.cpp:528-560): starts[b R + s] is a DEVICE vector of block b's dimension (or
 * NULL: no pair); each is normalised, run through Lanczos without reorthogonalisation in the
 * batched engine, and its spectral function S = -Im G / pi, G the continued fraction of alpha -
 * shift, beta, weight ||phi||^2, evaluated on the host. */
extern "C" int qedh_cf_batched(const qedh_matvec* blocks, int n_blocks, int R, const void* const* starts, int krylov,
                               int has_shift, double energy_shift, double breakdown_tol, int n_omega,
                               const double* omega, double eta, int use_graph, double* S, double* norm_sq,
                               double* shifts, int* steps_out, double* stats) {
    return guarded_batch([&] {
        if (!starts || !omega || !S || n_omega < 1 || krylov < 1)
            throw qsa::Error(qsa::kInvalid, "qedh_cf_batched: bad arguments");
        const auto ops = gather_blocks(blocks, n_blocks);
        const auto t0 = std::chrono::steady_clock::now();
        qsa::BatchedTimings t;
        qsa::BatchedLanczos L(ops, R);
        const auto t_tune = std::chrono::steady_clock::now();
        t.schedule = L.tune_schedule();
        t.tune_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_tune).count();
        L.clear_start();
        const std::size_t pairs = static_cast<std::size_t>(n_blocks) * static_cast<std::size_t>(R);
        for (int b = 0; b < n_blocks; ++b)
            for (int s = 0; s < R; ++s)
                if (starts[b * R + s]) L.place(b, s, static_cast<const qsa::cplx*>(starts[b * R + s]));
        std::vector<double> n2;
        L.normalize(&n2);
        std::vector<int> live(pairs, 0);
        for (std::size_t i = 0; i < pairs; ++i) live[i] = (starts[i] && std::sqrt(n2[i]) >= 1e-14) ? 1 : 0;
        t.pack_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<double> alpha, beta;
        std::vector<int> steps;
        const double tol = breakdown_tol > 0.0 ? breakdown_tol : 1e-12;
        L.run(krylov, std::vector<double>(static_cast<std::size_t>(n_blocks), tol), /*local=*/false, use_graph != 0,
              alpha, beta, steps, t, &live);
        const auto t1 = std::chrono::steady_clock::now();
        const std::vector<double> w(omega, omega + n_omega);
        qsa::parallel_blocks(pairs, [&](std::size_t i) {
            double* out = S + i * static_cast<std::size_t>(n_omega);
            if (norm_sq) norm_sq[i] = live[i] ? n2[i] : 0.0;
            if (steps_out) steps_out[i] = live[i] ? steps[i] : 0;
            if (!live[i]) {
                std::fill(out, out + n_omega, 0.0);
                if (shifts) shifts[i] = 0.0;
                return;
            }
            const int m = steps[i];
            std::vector<double> a(static_cast<std::size_t>(m)), bfull(static_cast<std::size_t>(m) + 1, 0.0);
            for (int j = 0; j < m; ++j) {
                a[j] = alpha[static_cast<std::size_t>(j) * pairs + i];
                bfull[j + 1] = beta[static_cast<std::size_t>(j) * pairs + i];
            }
            double E = energy_shift;
            if (!has_shift) E = qsa::tridiag_eig(a, bfull, a.size(), false).values.front();
            for (auto& x : a) x -= E;
            if (shifts) shifts[i] = E;
            const auto Sv = qsa::continued_fraction(a, bfull, w, eta, n2[i]);
            std::copy(Sv.begin(), Sv.end(), out);
        });
        if (stats) {
            stats[0] = t.pack_s;
            stats[1] = t.steps_s;
            stats[2] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
            stats[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            stats[4] = static_cast<double>(t.applies);
            stats[5] = static_cast<double>(t.rows);
            stats[6] = static_cast<double>(t.nnz);
            stats[7] = static_cast<double>(t.schedule);
            stats[8] = t.tune_s;
        }
    });
}

extern "C" int qedh_serve(const qedh_request_t* reqs, int n, const qedh_serve_options_t* opt, int stride,
                          double* values, double* e0, double* timing, int* status, int64_t* dims, double* stats) {
    return guarded_batch([&] {
        if (!reqs || n < 1 || !opt || stride < 1 || !values)
            throw qsa::Error(qsa::kInvalid, "qedh_serve: bad arguments");
        std::vector<qsa::Request> rs(static_cast<std::size_t>(n));
        for (int i = 0; i < n; ++i) {
            const auto& q = reqs[i];
            auto& r = rs[i];
            r.kind = q.kind;
            r.n_sites = q.n_sites;
            r.j1 = q.j1;
            r.j2 = q.j2;
            r.chi = q.chi;
            r.arrival_s = q.arrival_s;
            r.k = q.k;
            r.samples = q.samples;
            r.depth = q.depth;
            r.krylov = q.krylov;
            r.seed = q.seed;
        }
        qsa::ServeOptions o;
        o.slots = opt->slots;
        o.batched = opt->batched != 0;
        o.cache_sectors = opt->cache_sectors != 0;
        o.realtime = opt->realtime != 0;
        if (opt->hbm_budget) o.hbm_budget = opt->hbm_budget;
        if (opt->cache_budget) o.cache_budget = opt->cache_budget;
        if (opt->n_beta > 0 && opt->betas) o.betas.assign(opt->betas, opt->betas + opt->n_beta);
        if (opt->n_omega > 0 && opt->omega) o.omega.assign(opt->omega, opt->omega + opt->n_omega);
        o.eta = opt->eta;
        o.progress_s = opt->progress_s;
        std::vector<qsa::RequestResult> out;
        const qsa::ServeStats st = qsa::serve(rs, o, out);
        for (int i = 0; i < n; ++i) {
            const auto& r = out[i];
            double* v = values + static_cast<std::size_t>(i) * stride;
            std::fill(v, v + stride, 0.0);
            for (std::size_t j = 0; j < r.values.size() && j < static_cast<std::size_t>(stride); ++j) v[j] = r.values[j];
            if (e0) e0[i] = r.e0;
            if (timing) {
                timing[4 * i] = rs[i].arrival_s;
                timing[4 * i + 1] = r.start_s;
                timing[4 * i + 2] = r.finish_s;
                timing[4 * i + 3] = r.latency_s;
            }
            if (status) status[i] = r.status;
            if (dims) dims[i] = r.dim;
        }
        if (stats) {
            const double s[16] = {st.wall_s, st.solves_per_s, st.p50_s, st.p99_s, st.mean_s, st.max_s,
                                  static_cast<double>(st.completed), static_cast<double>(st.failed),
                                  static_cast<double>(st.peak_reserved), static_cast<double>(st.peak_in_flight),
                                  static_cast<double>(st.sector_builds), static_cast<double>(st.sector_hits),
                                  st.build_s, st.service_s[0], st.service_s[1], st.service_s[2]};
            std::copy(s, s + 16, stats);
        }
    });
}

extern "C" int qedh_request_footprint(const qedh_request_t* req, int batched, uint64_t* bytes) {
    return guarded_batch([&] {
        if (!req || !bytes) throw qsa::Error(qsa::kInvalid, "qedh_request_footprint: null argument");
        qsa::Request r;
        r.kind = req->kind;
        r.n_sites = req->n_sites;
        r.samples = req->samples;
        r.depth = req->depth;
        qsa::ServeOptions o;
        o.batched = batched != 0;
        *bytes = qsa::request_footprint(r, o);
    });
}

namespace {

/// The walk's results into the C-ABI's flat arrays.
void write_walk(const qsa::DynWalkResult& r, std::size_t nT, std::size_t nW, double* Z, double* S, double* e_min,
                int* samples_done) {
    for (std::size_t b = 0; b < r.Z.size(); ++b)
        for (std::size_t i = 0; i < nT; ++i) Z[b * nT + i] = r.Z[b][i];
    if (S)
        for (std::size_t c = 0; c < r.S.size(); ++c)
            for (std::size_t q = 0; q < nT * nW; ++q) {
                S[(c * nT * nW + q) * 2] = r.S[c][q].real();
                S[(c * nT * nW + q) * 2 + 1] = r.S[c][q].imag();
            }
    for (std::size_t b = 0; b < r.e_min.size(); ++b) {
        if (e_min) e_min[b] = r.e_min[b];
        if (samples_done) samples_done[b] = r.samples_done[b];
    }
}

}  // namespace

extern "C" int qedh_ftlm_dynamics_walk(int n_sources, const qedh_matvec* source_h, const uint64_t* source_seeds,
                                       const double* breakdown_tol, int n_channels, const int* channel_source,
                                       const qedh_matvec* channel_h, const qedh_matvec* channel_probe, int samples,
                                       int krylov, int n_t, const double* temperatures, int n_omega,
                                       const double* omega, double eta, uint64_t budget_bytes, double* Z, double* S,
                                       double* e_min, int* samples_done, double* stats) {
    return guarded_batch([&] {
        if (n_sources < 1 || !source_h || !source_seeds || !breakdown_tol || n_channels < 0 ||
            (n_channels > 0 && (!channel_source || !channel_h || !channel_probe || !S)) || samples < 1 || krylov < 1 ||
            n_t < 1 || !temperatures || n_omega < 1 || !omega || !Z)
            throw qsa::Error(qsa::kInvalid, "qedh_ftlm_dynamics_walk: bad arguments");
        const auto src = gather_blocks(source_h, n_sources);
        std::vector<qsa::DynChannel> ch(static_cast<std::size_t>(n_channels));
        for (int c = 0; c < n_channels; ++c) {
            if (!channel_h[c] || !channel_probe[c]) throw qsa::Error(qsa::kInvalid, "qedh_ftlm_dynamics_walk: null channel");
            ch[c].source = channel_source[c];
            ch[c].H = channel_h[c]->op;
            ch[c].A = channel_probe[c]->op;
        }
        qsa::DynWalkOptions o;
        o.samples = samples;
        o.krylov = krylov;
        o.source_seeds.assign(source_seeds, source_seeds + n_sources);
        o.breakdown_tol.assign(breakdown_tol, breakdown_tol + n_sources);
        o.temperatures.assign(temperatures, temperatures + n_t);
        o.omega.assign(omega, omega + n_omega);
        o.eta = eta;
        o.budget_bytes = budget_bytes;
        const qsa::DynWalkResult r = qsa::ftlm_dynamics_walk(src, ch, o);
        write_walk(r, static_cast<std::size_t>(n_t), static_cast<std::size_t>(n_omega), Z, S, e_min, samples_done);
        if (stats) {
            stats[0] = r.lanczos_s;
            stats[1] = r.overlap_s;
            stats[2] = r.host_s;
            stats[3] = r.total_s;
            stats[4] = static_cast<double>(r.waves);
            stats[5] = r.setup_s;
        }
    });
}

extern "C" int qedh_ftlm_dynamics(qedh_matvec source_h, int n_targets, const qedh_matvec* target_h,
                                  const qedh_matvec* probes, int samples, int krylov, uint64_t seed,
                                  double breakdown_tol, int n_t, const double* temperatures, int n_omega,
                                  const double* omega, double eta, double* Z, double* S, double* e_min,
                                  int* samples_done) {
    std::vector<int> src(static_cast<std::size_t>(std::max(n_targets, 0)), 0);
    const double tol = breakdown_tol > 0.0 ? breakdown_tol : 1e-300;
    return qedh_ftlm_dynamics_walk(1, &source_h, &seed, &tol, n_targets, src.data(), target_h, probes, samples, krylov,
                                   n_t, temperatures, n_omega, omega, eta, 0, Z, S, e_min, samples_done, nullptr);
}

extern "C" int qedh_gaussian_vectors(uint64_t base, uint64_t first_sample, int count, int64_t n, double* out) {
    return guarded_batch([&] {
        if (!out || count < 1 || n < 1) throw qsa::Error(qsa::kInvalid, "gaussian vectors: bad arguments");
        const hipStream_t st = qsa::engine_stream();
        const std::size_t c = static_cast<std::size_t>(count);
        std::vector<std::uint64_t> rb(c), dims(c, static_cast<std::uint64_t>(n));
        std::vector<std::uint32_t> seeds(c);
        for (std::size_t i = 0; i < c; ++i) {
            rb[i] = i * static_cast<std::uint64_t>(n);
            seeds[i] = static_cast<std::uint32_t>(qsa::splitmix_seed(base, first_sample + i));
        }
        qsa::DeviceBuffer rb_d(c * 8), dim_d(c * 8), seed_d(c * 4);
        qsa::copy_to_device(rb_d.get(), rb.data(), c * 8);
        qsa::copy_to_device(dim_d.get(), dims.data(), c * 8);
        qsa::copy_to_device(seed_d.get(), seeds.data(), c * 4);
        qsa::dev::sample_gauss(reinterpret_cast<qsa::dev::Zd*>(out), 1, rb_d.get<std::uint64_t>(),
                               dim_d.get<std::uint64_t>(), seed_d.get<std::uint32_t>(), static_cast<std::uint32_t>(c), st);
        QSA_HIP(hipGetLastError());
        QSA_HIP(hipStreamSynchronize(st));
    });
}

extern "C" int qedh_vdot(const void* x, const void* y, int64_t n, int complex_vectors, double* out) {
    return guarded_batch([&] {
        if (!x || !y || !out || n < 0) throw qsa::Error(qsa::kInvalid, "qedh_vdot: bad arguments");
        if (n == 0) {
            out[0] = out[1] = 0.0;
            return;
        }
        if (complex_vectors) {
            qsa::HipBackend<qsa::cplx> be;
            const qsa::cplx d = be.dot(static_cast<const qsa::cplx*>(x), static_cast<const qsa::cplx*>(y),
                                       static_cast<std::size_t>(n));
            out[0] = d.real();
            out[1] = d.imag();
        } else {
            qsa::HipBackend<double> be;
            out[0] = be.dot(static_cast<const double*>(x), static_cast<const double*>(y), static_cast<std::size_t>(n));
            out[1] = 0.0;
        }
    });
}

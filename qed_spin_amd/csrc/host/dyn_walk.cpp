#include "dyn_walk.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <string>
#include <thread>

#include "thermo.h"
#include "tridiag.h"

namespace qsa {

namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

/// The probes of a wave packed block-diagonally: block c maps the channel engine's rows of channel
/// c from the source engine's rows of its source; full complex values, tiles as the engines cut
/// them, so the batched SpMV applies every probe to every sample at once.
class PackedProbe {
public:
    PackedProbe(const std::vector<const SectorOperator*>& probes, const std::vector<std::uint64_t>& row_off,
                const std::vector<std::uint64_t>& col_off, int R)
        : R_(R) {
        const hipStream_t st = engine_stream();
        std::vector<std::uint64_t> nnz_off(1, 0);
        for (const auto* A : probes) {
            if (A->lane() != kLaneCsr) throw Error(kDeviceUnsupported, "ftlm dynamics: a probe needs its device CSR");
            nnz_off.push_back(nnz_off.back() + static_cast<std::uint64_t>(A->csr_info().nnz));
        }
        rows_ = row_off.back();
        row_ptr_.reset((rows_ + 1) * 8);
        col_.reset(std::max<std::uint64_t>(nnz_off.back(), 1) * 4);
        val_.reset(std::max<std::uint64_t>(nnz_off.back(), 1) * sizeof(dev::Zd));
        for (std::size_t c = 0; c < probes.size(); ++c)
            dev::pack_csr(probes[c]->device_csr(), row_off[c], col_off[c], nnz_off[c], row_ptr_.get<std::uint64_t>(),
                          col_.get<std::uint32_t>(), val_.get<dev::Zd>(), nullptr, 0, st);
        QSA_HIP(hipGetLastError());
        std::vector<dev::Tile> tiles;
        std::vector<std::uint32_t> begin(1, 0);
        for (std::size_t c = 0; c < probes.size(); ++c) {
            for (std::uint64_t r = row_off[c]; r < row_off[c + 1]; r += dev::kBatchTileRows) {
                dev::Tile t{};
                t.row_begin = r;
                t.row_end = std::min<std::uint64_t>(r + dev::kBatchTileRows, row_off[c + 1]);
                t.block = static_cast<std::uint32_t>(c);
                tiles.push_back(t);
            }
            begin.push_back(static_cast<std::uint32_t>(tiles.size()));
        }
        n_blocks_ = static_cast<std::uint32_t>(probes.size());
        n_tiles_ = static_cast<std::uint32_t>(tiles.size());
        tiles_ = upload(tiles.data(), tiles.size());
        begin_ = upload(begin.data(), begin.size());
    }
    /// Y (channel layout) = A X (source layout), every block and vector.
    void apply(const dev::Zd* X, dev::Zd* Y) const {
        dev::BatchView B;
        B.rows = rows_;
        B.n_blocks = n_blocks_;
        B.n_tiles = n_tiles_;
        B.R = R_;
        B.tiles = tiles_.get<dev::Tile>();
        B.tile_begin = begin_.get<std::uint32_t>();
        B.row_ptr = row_ptr_.get<std::uint64_t>();
        B.col = col_.get<std::uint32_t>();
        B.val = val_.get<dev::Zd>();
        dev::batch_spmv(B, X, Y, engine_stream());
        QSA_HIP(hipGetLastError());
    }

private:
    int R_;
    std::uint64_t rows_ = 0;
    std::uint32_t n_blocks_ = 0, n_tiles_ = 0;
    DeviceBuffer row_ptr_, col_, val_, tiles_, begin_;
};

/// The Ritz data of one Lanczos run: values ascending, eigenvectors column-major (m x m).
struct Ritz {
    std::size_t m = 0;
    std::vector<double> value, vec;
    double z(std::size_t row, std::size_t col) const { return vec[col * m + row]; }
};

Ritz ritz_of(const std::vector<double>& alpha_hist, const std::vector<double>& beta_hist, std::size_t pairs,
             std::size_t pair, int steps) {
    Ritz R;
    if (steps <= 0) return R;
    const std::size_t m = static_cast<std::size_t>(steps);
    std::vector<double> a(m), b(m + 1, 0.0);
    for (std::size_t j = 0; j < m; ++j) {
        a[j] = alpha_hist[j * pairs + pair];
        b[j + 1] = beta_hist[j * pairs + pair];   // beta_{j+1} couples steps j and j + 1
    }
    TridiagEig e = tridiag_eig(a, b, m, true);
    R.m = m;
    R.value = std::move(e.values);
    R.vec = std::move(e.vectors);
    return R;
}

/// One sample's source-side data: its Ritz pairs, first components and lowest weighted value.
struct SourceSample {
    bool ok = false;
    Ritz ritz;
    double smin = 0.0;
    std::vector<double> Z;   ///< [T], about smin
};

template <class Fn> void parallel_items(std::size_t n, Fn&& fn) {
    parallel_blocks(n, [&](std::size_t i) { fn(i); });
}

}  // namespace

DynWalkResult ftlm_dynamics_walk(const std::vector<std::shared_ptr<const SectorOperator>>& sources,
                                 const std::vector<DynChannel>& channels, const DynWalkOptions& o) {
    const auto t_all = clk::now();
    const std::size_t nS = sources.size(), nC = channels.size();
    if (nS == 0) throw Error(kInvalid, "ftlm dynamics: no source sector");
    if (o.samples < 1 || o.krylov < 1 || o.temperatures.empty() || o.omega.empty())
        throw Error(kInvalid, "ftlm dynamics: samples, krylov, temperatures and frequencies are required");
    if (o.source_seeds.size() != nS || o.breakdown_tol.size() != nS)
        throw Error(kInvalid, "ftlm dynamics: one seed and one breakdown tolerance per source");
    for (const auto& s : sources)
        if (!s || !s->same() || s->lane() != kLaneCsr || s->dim() <= 0)
            throw Error(kInvalid, "ftlm dynamics: every source needs a square CSR operator of a non-empty sector");
    for (const auto& c : channels) {
        if (c.source < 0 || static_cast<std::size_t>(c.source) >= nS || !c.H || !c.A)
            throw Error(kInvalid, "ftlm dynamics: a channel names no source, H or probe");
        if (!c.H->same() || c.H->lane() != kLaneCsr || c.A->col_dim() != sources[c.source]->dim() ||
            c.A->dim() != c.H->dim())
            throw Error(kInvalid, "ftlm dynamics: a probe does not map its source to its target");
    }
    const int M = o.krylov;
    const std::size_t nT = o.temperatures.size(), nW = o.omega.size();
    const std::size_t full_limit = std::max<std::size_t>(4 * static_cast<std::size_t>(M), kFullReorthMaxDim);
    std::uint64_t budget = o.budget_bytes;
    if (!budget) {   // two fifths of the device's memory, at most nine tenths of what is free now
        std::size_t free_b = 0, total_b = 0;
        QSA_HIP(hipMemGetInfo(&free_b, &total_b));
        budget = std::max<std::uint64_t>(std::min<std::uint64_t>(total_b / 5 * 2, free_b / 10 * 9), std::uint64_t{1} << 30);
    }
    std::vector<std::vector<int>> chans_of(nS);
    for (std::size_t c = 0; c < nC; ++c) chans_of[channels[c].source].push_back(static_cast<int>(c));

    DynWalkResult out;
    out.Z.assign(nS, std::vector<double>(nT, 0.0));
    out.S.assign(nC, std::vector<cplx>(nT * nW, cplx(0.0, 0.0)));
    out.e_min.assign(nS, 0.0);
    out.samples_done.assign(nS, 0);
    // every sample's results, about its own smin: combined once every sample is known
    std::vector<std::vector<SourceSample>> src_samples(nS, std::vector<SourceSample>(static_cast<std::size_t>(o.samples)));
    std::vector<std::vector<std::vector<cplx>>> chan_samples(
        nC, std::vector<std::vector<cplx>>(static_cast<std::size_t>(o.samples)));

    const int Rmax = std::min(o.samples, dev::kBatchMaxR);
    const hipStream_t st = engine_stream();
    // The kept vectors live in three device regions that every wave reuses (allocated once: kept
    // vectors of tens of GB allocated per wave in slightly different sizes cost the pool more time
    // than the wave's work): (1) the source bases, later the probe images transposed; (2) the target
    // bases; (3) the probe images, later the target bases transposed. Per sample, with r_s the wave's
    // source rows and r_c its channel rows: 16 (M + 1) (max(r_s, r_c) + 2 r_c) bytes.
    auto rows_c_of = [&](std::size_t b) {
        std::uint64_t r = 0;
        for (int c : chans_of[b]) r += static_cast<std::uint64_t>(channels[c].H->dim());
        return r;
    };
    const std::uint64_t layer_bytes = 16ull * static_cast<std::uint64_t>(M + 1);
    auto wave_bytes = [&](std::uint64_t rs, std::uint64_t rc) { return layer_bytes * (std::max(rs, rc) + 2 * rc); };
    struct Wave {
        std::size_t b0, b1;
        int R;
    };
    // waves: consecutive sources whose kept vectors for Rmax samples fit the budget (a source too
    // large for that forms a wave of its own and takes fewer samples per round)
    std::vector<Wave> waves;
    std::uint64_t need1 = 0, need2 = 0;   // region sizes, in vectors (region 3 equals region 2)
    for (std::size_t b0 = 0; b0 < nS;) {
        std::size_t b1 = b0;
        std::uint64_t rs = 0, rc = 0;
        while (b1 < nS) {
            const std::uint64_t rs2 = rs + static_cast<std::uint64_t>(sources[b1]->dim()), rc2 = rc + rows_c_of(b1);
            if (b1 > b0 && wave_bytes(rs2, rc2) * static_cast<std::uint64_t>(Rmax) > budget) break;
            rs = rs2;
            rc = rc2;
            ++b1;
        }
        const int R = static_cast<int>(
            std::max<std::uint64_t>(1, std::min<std::uint64_t>(Rmax, budget / std::max<std::uint64_t>(wave_bytes(rs, rc), 1))));
        waves.push_back(Wave{b0, b1, R});
        need1 = std::max(need1, std::max(rs, rc) * static_cast<std::uint64_t>(R));
        need2 = std::max(need2, rc * static_cast<std::uint64_t>(R));
        b0 = b1;
    }
    // One allocation holds the three regions. A walk that needs more than half of the budget takes
    // the whole budget, so that consecutive large walks ask the stream-ordered pool for the same size
    // and reuse the block it keeps: memory the pool has given back costs seconds to map again.
    auto up256 = [](std::uint64_t b) { return (b + 255) / 256 * 256; };
    const std::uint64_t bytes1 = up256(need1 * layer_bytes), bytes2 = up256(need2 * layer_bytes);
    std::uint64_t arena_bytes = bytes1 + 2 * bytes2;
    if (arena_bytes > budget / 2 && arena_bytes < budget) arena_bytes = budget;
    DeviceBuffer arena(static_cast<std::size_t>(arena_bytes));
    dev::Zd* const region1 = arena.get<dev::Zd>();
    dev::Zd* const region2 = reinterpret_cast<dev::Zd*>(arena.get<char>() + bytes1);
    dev::Zd* const region3 = reinterpret_cast<dev::Zd*>(arena.get<char>() + bytes1 + bytes2);
    DeviceBuffer partial, Omega;   // grown when a wave needs more, never shrunk
    QSA_HIP(hipStreamSynchronize(st));
    out.setup_s = since(t_all);
    auto ensure = [](DeviceBuffer& b, std::size_t bytes) {
        if (b.bytes() < bytes) b.reset(bytes);
    };
    for (const Wave& wv : waves) {
        const std::size_t b0 = wv.b0, b1 = wv.b1;
        const int Rfit = wv.R;
        for (int s0 = 0; s0 < o.samples; s0 += Rfit) {
            const int R = std::min(Rfit, o.samples - s0);
            ++out.waves;
            // ---- the sources' Lanczos runs with kept bases -------------------------------------
            auto t_lz = clk::now();
            std::vector<std::shared_ptr<const SectorOperator>> wsrc(sources.begin() + static_cast<std::ptrdiff_t>(b0),
                                                                     sources.begin() + static_cast<std::ptrdiff_t>(b1));
            const std::size_t nb = wsrc.size();
            BatchedLanczos S(wsrc, R);
            std::vector<std::uint32_t> seeds(nb * static_cast<std::size_t>(R));
            std::vector<double> tol_s(nb);
            std::vector<int> full_s(nb);
            for (std::size_t b = 0; b < nb; ++b) {
                for (int s = 0; s < R; ++s)
                    seeds[b * R + s] = static_cast<std::uint32_t>(
                        splitmix_seed(o.source_seeds[b0 + b], static_cast<std::uint64_t>(s0 + s)));
                tol_s[b] = o.breakdown_tol[b0 + b];
                full_s[b] = static_cast<std::size_t>(S.dim(static_cast<int>(b))) <= full_limit ? 1 : 0;
            }
            S.sample_start(seeds);
            const std::uint64_t layer_s = S.rows() * static_cast<std::uint64_t>(R);
            dev::Zd* const Vs = region1;
            std::vector<double> aS, bS;
            std::vector<int> stS;
            BatchedTimings ts;
            S.run_kept(M, tol_s, full_s, reinterpret_cast<cplx*>(Vs), aS, bS, stS, ts);
            // ---- the channels' target runs, started from the probe images ----------------------
            std::vector<int> wch;
            for (std::size_t b = b0; b < b1; ++b)
                for (int c : chans_of[b]) wch.push_back(c);
            const std::size_t nc = wch.size();
            std::vector<double> n2;
            std::vector<double> aC, bC;
            std::vector<int> stC;
            std::vector<cplx> omega_h;
            std::vector<std::uint64_t> crow(1, 0);
            if (nc) {
                std::vector<std::shared_ptr<const SectorOperator>> wtgt;
                std::vector<const SectorOperator*> probes;
                std::vector<std::uint64_t> colo;
                std::vector<double> tol_c;
                std::vector<int> full_c;
                for (int c : wch) {
                    wtgt.push_back(channels[c].H);
                    probes.push_back(channels[c].A.get());
                    crow.push_back(crow.back() + static_cast<std::uint64_t>(channels[c].H->dim()));
                    colo.push_back(S.row_offset(channels[c].source - static_cast<int>(b0)));
                    tol_c.push_back(o.breakdown_tol[channels[c].source]);
                    full_c.push_back(static_cast<std::size_t>(channels[c].H->dim()) <= full_limit ? 1 : 0);
                }
                const PackedProbe P(probes, crow, colo, R);
                BatchedLanczos T(wtgt, R);
                P.apply(Vs, reinterpret_cast<dev::Zd*>(T.start_buffer()));
                T.normalize(&n2);
                std::vector<int> live(nc * static_cast<std::size_t>(R));
                for (std::size_t i = 0; i < live.size(); ++i) live[i] = std::sqrt(n2[i]) >= 1e-14 ? 1 : 0;
                const std::uint64_t layer_c = T.rows() * static_cast<std::uint64_t>(R);
                dev::Zd* const Wb = region2;
                BatchedTimings tc;
                T.run_kept(M, tol_c, full_c, reinterpret_cast<cplx*>(Wb), aC, bC, stC, tc, &live);
                out.lanczos_s += since(t_lz);
                // ---- overlaps <A v_a | w_b>: the probe on every source layer, then one GEMM per pair
                const auto t_ov = clk::now();
                int mH = 0;
                for (int v : stS) mH = std::max(mH, v);
                // region 3 <- A (region 1); region 1 <- region 3 transposed; region 3 <- region 2 transposed
                dev::Zd* const AV = region3;
                for (int j = 0; j < mH; ++j)
                    P.apply(Vs + static_cast<std::uint64_t>(j) * layer_s, AV + static_cast<std::uint64_t>(j) * layer_c);
                dev::Zd* const AVt = region1;
                dev::batch_layers_transpose(AV, T.rows(), R, std::max(mH, 1), AVt, st);
                dev::Zd* const Wt = region3;
                dev::batch_layers_transpose(Wb, T.rows(), R, M + 1, Wt, st);
                QSA_HIP(hipGetLastError());
                const std::size_t per = static_cast<std::size_t>(M) * static_cast<std::size_t>(M);
                ensure(Omega, nc * static_cast<std::size_t>(R) * per * sizeof(dev::Zd));
                // every live pair's rows in chunks of at most 8192: the overlaps as split products
                const std::uint64_t rows_t = T.rows();
                const int kav = std::max(mH, 1);
                std::vector<dev::OverlapChunk> chunks;
                std::vector<std::uint32_t> chunk_begin(1, 0);
                std::vector<int> ma(nc * static_cast<std::size_t>(R), 0), mb(ma.size(), 0);
                for (std::size_t ci = 0; ci < nc; ++ci) {
                    const int src = channels[wch[ci]].source - static_cast<int>(b0);
                    for (int s = 0; s < R; ++s) {
                        const std::size_t p = ci * R + s;
                        const int mh = stS[static_cast<std::size_t>(src) * R + s];
                        const int ms = stC[p];
                        if (mh > 0 && ms > 0 && live[p]) {
                            ma[p] = mh;
                            mb[p] = ms;
                            const std::uint64_t n_rows = crow[ci + 1] - crow[ci];
                            for (std::uint64_t r0 = 0; r0 < n_rows; r0 += 8192) {
                                dev::OverlapChunk q{};
                                q.x_base = static_cast<std::uint64_t>(s) * kav * rows_t + crow[ci] + r0;
                                q.y_base = static_cast<std::uint64_t>(s) * (M + 1) * rows_t + crow[ci] + r0;
                                q.rows = static_cast<std::uint32_t>(std::min<std::uint64_t>(8192, n_rows - r0));
                                q.pair = static_cast<std::uint32_t>(p);
                                chunks.push_back(q);
                            }
                        }
                        chunk_begin.push_back(static_cast<std::uint32_t>(chunks.size()));
                    }
                }
                if (!chunks.empty()) {
                    // every chunk's product in one launch, then the chunks of every pair summed in order
                    ensure(partial, chunks.size() * per * sizeof(dev::Zd));
                    DeviceBuffer chunks_d = upload(chunks.data(), chunks.size());
                    DeviceBuffer begin_d = upload(chunk_begin.data(), chunk_begin.size());
                    DeviceBuffer ma_d = upload(ma.data(), ma.size()), mb_d = upload(mb.data(), mb.size());
                    dev::overlap_chunks(AVt, rows_t, Wt, rows_t,
                                        chunks_d.get<dev::OverlapChunk>(), static_cast<std::uint32_t>(chunks.size()),
                                        ma_d.get<int>(), mb_d.get<int>(), M, partial.get<dev::Zd>(), st);
                    dev::overlap_reduce(partial.get<dev::Zd>(), begin_d.get<std::uint32_t>(), ma_d.get<int>(),
                                        mb_d.get<int>(), static_cast<std::uint32_t>(ma.size()), M,
                                        Omega.get<dev::Zd>(), st);
                    QSA_HIP(hipGetLastError());
                    QSA_HIP(hipStreamSynchronize(st));
                }
                omega_h.resize(nc * static_cast<std::size_t>(R) * per);
                copy_to_host(omega_h.data(), Omega.get(), omega_h.size() * sizeof(cplx));
                out.overlap_s += since(t_ov);
            } else {
                out.lanczos_s += since(t_lz);
            }
            // ---- host: Ritz pairs, Boltzmann factors, Lorentzian sums -------------------------
            const auto t_host = clk::now();
            const std::size_t pairs_s = nb * static_cast<std::size_t>(R);
            parallel_items(pairs_s, [&](std::size_t i) {
                const std::size_t b = i / R, s = i % R;
                SourceSample& out_s = src_samples[b0 + b][static_cast<std::size_t>(s0) + s];
                out_s.ritz = ritz_of(aS, bS, pairs_s, i, stS[i]);
                const Ritz& r = out_s.ritz;
                if (r.m == 0) return;
                double smin = std::numeric_limits<double>::infinity(), lo = smin;
                for (std::size_t k = 0; k < r.m; ++k) {
                    lo = std::min(lo, r.value[k]);
                    if (r.z(0, k) != 0.0) smin = std::min(smin, r.value[k]);
                }
                out_s.smin = std::isfinite(smin) ? smin : lo;
                out_s.Z.assign(nT, 0.0);
                for (std::size_t it = 0; it < nT; ++it) {
                    const double beta = 1.0 / o.temperatures[it];
                    for (std::size_t k = 0; k < r.m; ++k)
                        out_s.Z[it] += r.z(0, k) * r.z(0, k) * std::exp(-beta * (r.value[k] - out_s.smin));
                }
                out_s.ok = true;
            });
            const std::size_t per = static_cast<std::size_t>(M) * static_cast<std::size_t>(M);
            const double eta = o.eta, inv_pi = 0.31830988618379067154;
            parallel_items(nc * static_cast<std::size_t>(R), [&](std::size_t i) {
                const std::size_t ci = i / R, s = i % R;
                const int c = wch[ci];
                const int src = channels[c].source;
                const SourceSample& ss = src_samples[src][static_cast<std::size_t>(s0) + s];
                std::vector<cplx>& acc = chan_samples[c][static_cast<std::size_t>(s0) + s];
                acc.assign(nT * nW, cplx(0.0, 0.0));
                const int ms = stC[ci * R + s];
                if (!ss.ok || ms <= 0 || std::sqrt(n2[ci * R + s]) < 1e-14) return;
                const Ritz tg = ritz_of(aC, bC, nc * static_cast<std::size_t>(R), ci * R + s, ms);
                const Ritz& sr = ss.ritz;
                const std::size_t mh = sr.m, mt = tg.m;
                const cplx* Om = omega_h.data() + i * per;   // mh x mt, column-major
                // X = Omega VS (mh x mt), then Mm = VH^T X
                std::vector<cplx> X(mh * mt, cplx(0.0, 0.0)), Mm(mh * mt, cplx(0.0, 0.0));
                for (std::size_t j = 0; j < mt; ++j)
                    for (std::size_t q = 0; q < mt; ++q) {
                        const double v = tg.z(q, j);
                        for (std::size_t a = 0; a < mh; ++a) X[j * mh + a] += Om[q * mh + a] * v;
                    }
                for (std::size_t j = 0; j < mt; ++j)
                    for (std::size_t k = 0; k < mh; ++k) {
                        cplx sum(0.0, 0.0);
                        for (std::size_t a = 0; a < mh; ++a) sum += sr.z(a, k) * X[j * mh + a];
                        Mm[j * mh + k] = sum;
                    }
                const double nphi = std::sqrt(n2[ci * R + s]);
                std::vector<cplx> line(nW);
                for (std::size_t k = 0; k < mh; ++k) {
                    const double ck = sr.z(0, k);
                    if (ck == 0.0) continue;
                    std::fill(line.begin(), line.end(), cplx(0.0, 0.0));
                    for (std::size_t j = 0; j < mt; ++j) {
                        const cplx w = ck * Mm[j * mh + k] * (nphi * tg.z(0, j));
                        if (std::abs(w) < 1e-300) continue;
                        const double de = tg.value[j] - sr.value[k];
                        for (std::size_t q = 0; q < nW; ++q) {
                            const double x = o.omega[q] - de;
                            line[q] += w * (eta * inv_pi / (x * x + eta * eta));
                        }
                    }
                    for (std::size_t it = 0; it < nT; ++it) {
                        const double f = std::exp(-(sr.value[k] - ss.smin) / o.temperatures[it]);
                        if (f < 1e-300) continue;
                        for (std::size_t q = 0; q < nW; ++q) acc[it * nW + q] += f * line[q];
                    }
                }
            });
            out.host_s += since(t_host);
        }
    }
    // every source's samples combined about their common reference, in sample order
    const auto t_host = clk::now();
    for (std::size_t b = 0; b < nS; ++b) {
        double ref = std::numeric_limits<double>::infinity();
        int done = 0;
        for (const auto& ss : src_samples[b])
            if (ss.ok) {
                ref = std::min(ref, ss.smin);
                ++done;
            }
        out.samples_done[b] = done;
        out.e_min[b] = done ? ref : 0.0;
        if (!done) continue;
        const double scale = static_cast<double>(sources[b]->dim()) / static_cast<double>(done);
        for (std::size_t it = 0; it < nT; ++it) {
            double z = 0.0;
            for (const auto& ss : src_samples[b])
                if (ss.ok) z += std::exp(-(ss.smin - ref) / o.temperatures[it]) * ss.Z[it];
            out.Z[b][it] = z * scale;
        }
        for (int c : chans_of[b]) {
            for (std::size_t smp = 0; smp < src_samples[b].size(); ++smp) {
                const auto& ss = src_samples[b][smp];
                const auto& acc = chan_samples[c][smp];
                if (!ss.ok || acc.empty()) continue;
                for (std::size_t it = 0; it < nT; ++it) {
                    const double f = std::exp(-(ss.smin - ref) / o.temperatures[it]);
                    for (std::size_t q = 0; q < nW; ++q) out.S[c][it * nW + q] += f * acc[it * nW + q];
                }
            }
            for (auto& v : out.S[c]) v *= scale;
        }
    }
    out.host_s += since(t_host);
    out.total_s = since(t_all);
    return out;
}

}  // namespace qsa

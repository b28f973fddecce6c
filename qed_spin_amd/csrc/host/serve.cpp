#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <future>
#include <mutex>
#include <cstdio>
#include <thread>
#include <tuple>

#include "algebra.h"
#include "batched.h"
#include "eigsolve.h"
#include "sector.h"
#include "sector_operator.h"
#include "thermo.h"

namespace qsa {

namespace {

using clk = std::chrono::steady_clock;

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/// e^{2 pi i p / L}, exact at the quarter turns (the characters of src/basis/irreps.cpp:161-175).
cplx root_of_unity(long long p, long long L) {
    p %= L;
    if (p < 0) p += L;
    if ((4 * p) % L == 0) {
        switch ((4 * p) / L) {
            case 0: return {1.0, 0.0};
            case 1: return {0.0, 1.0};
            case 2: return {-1.0, 0.0};
            default: return {0.0, -1.0};
        }
    }
    const double th = 6.283185307179586476925286766559 * static_cast<double>(p) / static_cast<double>(L);
    return {std::cos(th), std::sin(th)};
}

/// The chain's Hamiltonian (J1, J2 Heisenberg, scalar chirality on consecutive triangles).
SpinOperator chain_model(int N, double j1, double j2, double chi) {
    SpinOperator H(N);
    auto heis = [&](int i, int j, double J) {
        if (J == 0.0) return;
        H.add(SpinOperator::product(N, "+-", {i, j}, cplx(0.5 * J, 0.0)));
        H.add(SpinOperator::product(N, "-+", {i, j}, cplx(0.5 * J, 0.0)));
        H.add(SpinOperator::product(N, "zz", {i, j}, cplx(J, 0.0)));
    };
    for (int i = 0; i < N; ++i) heis(i, (i + 1) % N, j1);
    for (int i = 0; i < N; ++i) heis(i, (i + 2) % N, j2);
    if (chi != 0.0) {
        static const char* perms[6] = {"xyz", "yzx", "zxy", "xzy", "yxz", "zyx"};
        static const double sign[6] = {1, 1, 1, -1, -1, -1};
        for (int i = 0; i < N; ++i)
            for (int t = 0; t < 6; ++t)
                H.add(SpinOperator::product(N, perms[t], {i, (i + 1) % N, (i + 2) % N}, cplx(sign[t] * chi, 0.0)));
    }
    return H;
}

/// The momentum-m sector of the N translations T^a (T^a[i] = i + a mod N) at n_up.
GroupSpec chain_sector(int N, int n_up, int m) {
    GroupSpec g;
    g.n_sites = N;
    g.n_up = n_up;
    g.G = N;
    g.perms.resize(static_cast<std::size_t>(N) * N);
    for (int a = 0; a < N; ++a)
        for (int i = 0; i < N; ++i) g.perms[static_cast<std::size_t>(a) * N + i] = (i + a) % N;
    g.flips.assign(N, 0ULL);
    g.chars.resize(N);
    for (int a = 0; a < N; ++a) g.chars[a] = root_of_unity(static_cast<long long>(m) * a, N);
    g.slc = build_sublattice(g.perms, g.flips, N, N, -1, true);
    return g;
}

constexpr std::uint64_t kRankBudget = 16ull << 30;

class SectorCache {
public:
    SectorCache(bool enabled, std::uint64_t budget) : enabled_(enabled), budget_(budget) {}

    std::shared_ptr<const DeviceSector> get(int N, int n_up, int m, bool& built) {
        built = false;
        if (!enabled_) {
            built = true;
            return make(N, n_up, m);
        }
        const Key key{N, n_up, m};
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            auto it = map_.find(key);
            if (it == map_.end()) break;
            if (it->second.s) {
                it->second.last = ++tick_;
                return it->second.s;
            }
            cv_.wait(lk);   // another slot is building it
        }
        map_[key] = Entry{};
        lk.unlock();
        std::shared_ptr<const DeviceSector> s;
        try {
            s = make(N, n_up, m);
        } catch (...) {
            lk.lock();
            map_.erase(key);
            cv_.notify_all();
            throw;
        }
        built = true;
        lk.lock();
        auto& e = map_[key];
        e.s = s;
        e.bytes = s->stats().device_bytes;
        e.last = ++tick_;
        bytes_ += e.bytes;
        evict_locked();
        cv_.notify_all();
        return s;
    }

private:
    using Key = std::tuple<int, int, int>;
    struct Entry {
        std::shared_ptr<const DeviceSector> s;
        std::uint64_t bytes = 0, last = 0;
    };
    static std::shared_ptr<const DeviceSector> make(int N, int n_up, int m) {
        auto s = std::make_shared<DeviceSector>(chain_sector(N, n_up, m), kRankBudget);
        QSA_HIP(hipStreamSynchronize(engine_stream()));   // complete before another slot's stream reads it
        return s;
    }
    // Least recently used first, only sectors no request holds.
    void evict_locked() {
        while (bytes_ > budget_) {
            auto victim = map_.end();
            for (auto it = map_.begin(); it != map_.end(); ++it)
                if (it->second.s && it->second.s.use_count() == 1 && (victim == map_.end() || it->second.last < victim->second.last))
                    victim = it;
            if (victim == map_.end()) return;
            bytes_ -= victim->second.bytes;
            map_.erase(victim);
        }
    }
    bool enabled_;
    std::uint64_t budget_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::map<Key, Entry> map_;
    std::uint64_t tick_ = 0, bytes_ = 0;
};

double kbreak(double sH) { return 64.0 * std::numeric_limits<double>::epsilon() * (sH > 0.0 ? sH : 1.0); }

/// One request on the calling thread's engine stream.
void solve(const Request& r, const ServeOptions& o, SectorCache& cache, RequestResult& res, int& builds, int& hits) {
    const int N = r.n_sites, n_up = N / 2;
    auto tb = clk::now();
    auto build_done = [&] {   // build phases end: their device work is finished before the clock stops
        QSA_HIP(hipStreamSynchronize(engine_stream()));
        res.build_s += std::chrono::duration<double>(clk::now() - tb).count();
    };
    bool built = false;
    auto src = cache.get(N, n_up, 0, built);
    (built ? builds : hits) += 1;
    res.dim = src->dim();
    if (src->dim() == 0) throw Error(kInvalid, "serve: the sector is empty");
    const SpinOperator H = chain_model(N, r.j1, r.j2, r.chi);
    const double sH = H.l1_norm();
    auto hprog = std::make_shared<DeviceProgram>(compile_operator(H.dagger()));
    // fixed SpMV width: a request's result must not depend on what else runs
    auto Hop = std::make_shared<SectorOperator>(src, src, hprog, sH, kLaneCsr, ~0ull, /*tune=*/false);
    build_done();
    const std::size_t n = static_cast<std::size_t>(src->dim());
    HipBackend<cplx> be;
    const OperatorMatvec<cplx> applyH{*Hop, true};
    if (r.kind == kReqEigs) {
        const std::size_t k = std::min<std::size_t>(static_cast<std::size_t>(std::max(1, r.k)), n);
        const BlockSolution sol = krylov_schur_lane(be, applyH, n, k, sH, 0, static_cast<cplx*>(nullptr));
        res.values = sol.values;
        res.e0 = sol.values.empty() ? 0.0 : sol.values.front();
    } else if (r.kind == kReqThermal) {
        Curves c;
        if (o.batched) {
            BatchedFtlmOptions fo;
            fo.samples = r.samples;
            fo.krylov = r.depth;
            fo.betas = o.betas;
            fo.block_seeds = {r.seed};
            fo.breakdown_tol = {kbreak(sH)};
            BatchedTimings t;
            c = batched_ftlm({Hop}, fo, t).front();
        } else {
            FtlmOptions fo;
            fo.samples = static_cast<std::size_t>(r.samples);
            fo.krylov = static_cast<std::size_t>(r.depth);
            fo.breakdown_tol = kbreak(sH);
            fo.betas = o.betas;
            fo.seed = r.seed;
            c = ftlm_block(be, applyH, n, fo);
        }
        res.values = c.lnZ;
        res.values.insert(res.values.end(), c.E.begin(), c.E.end());
        res.values.insert(res.values.end(), c.V.begin(), c.V.end());
        res.e0 = c.e_min;
    } else if (r.kind == kReqDynamics) {
        auto gs = be.make_zero_vector(n);
        const BlockSolution sol = krylov_schur_lane(be, applyH, n, 1, sH, 0, gs.get());
        if (sol.values.empty()) throw Error(kConvergence, "serve: no ground state");
        const double e0 = sol.values.front();
        res.e0 = e0;
        // S^z(q) = N^-1/2 sum_j e^{-i q j} S^z_j, q = 2 pi m / N nearest pi; it takes momentum 0 to -m
        const int m = N / 2;
        const int mt = (N - m) % N;
        tb = clk::now();
        bool b2 = false;
        auto tgt = cache.get(N, n_up, mt, b2);
        (b2 ? builds : hits) += 1;
        SpinOperator O(N);
        for (int j = 0; j < N; ++j)
            O.add(SpinOperator::product(N, "z", {j}, root_of_unity(-static_cast<long long>(m) * j, N) / std::sqrt(double(N))));
        const auto& gt = tgt->spec();
        const auto& gsrc = src->spec();
        auto xprog = std::make_shared<DeviceProgram>(compile_projected(O.dagger(), gt.G, gt.perms.data(), gt.flips.data(),
                                                                       gt.chars.data(), gsrc.chars.data(), gt.n_up,
                                                                       gsrc.n_up));
        SectorOperator X(tgt, src, xprog, O.l1_norm(), kLaneGather, 0, false);
        const std::size_t nt = static_cast<std::size_t>(tgt->dim());
        auto phi = be.make_zero_vector(nt);
        X.apply(gs.get(), phi.get(), true);
        SectorOperator Ht(tgt, tgt, hprog, sH, kLaneCsr, ~0ull, false);
        build_done();
        const OperatorMatvec<cplx> applyT{Ht, true};
        const CfResult cf = cf_from_vector(be, applyT, nt, phi.get(), o.omega, o.eta,
                                           static_cast<std::size_t>(std::max(2, r.krylov)), true, e0, kbreak(sH));
        res.values = cf.S;
    } else {
        throw Error(kInvalid, "serve: unknown request kind");
    }
    QSA_HIP(hipStreamSynchronize(engine_stream()));
}

}  // namespace

std::uint64_t request_footprint(const Request& r, const ServeOptions& o) {
    const int N = r.n_sites, n_up = N / 2;
    const double C = binom(N, n_up);
    const double D = C / N + 16.0;
    const double nnz = D * (N + 1.0);
    const double vec = 16.0 * D;
    const double sector = 16.0 * D + (4.0 * C <= static_cast<double>(kRankBudget) ? 4.0 * C : 0.0) + (1 << 20);
    const double csr = nnz * 6.0 + 8.0 * D;
    double b = sector + csr;
    if (r.kind == kReqEigs) b += 64.0 * vec;
    else if (r.kind == kReqThermal) b += o.batched ? nnz * 20.0 + (3.0 * r.samples + 4.0) * vec : 8.0 * vec;
    else b += sector + 2.0 * csr + 52.0 * vec;
    return static_cast<std::uint64_t>(b * 1.1) + (8ull << 20);
}

ServeStats serve(const std::vector<Request>& reqs, const ServeOptions& o, std::vector<RequestResult>& out) {
    require_device();
    if (o.slots < 1) throw Error(kInvalid, "serve: at least one slot");
    for (const auto& r : reqs) {
        if (r.n_sites < 4 || r.n_sites > 40) throw Error(kInvalid, "serve: n_sites must be in [4, 40]");
        if (r.kind == kReqThermal && (o.betas.empty() || r.samples < 1 || r.depth < 1))
            throw Error(kInvalid, "serve: thermal requests need temperatures, samples and depth");
        if (r.kind == kReqDynamics && o.omega.empty()) throw Error(kInvalid, "serve: dynamics requests need omega");
        if (request_footprint(r, o) > o.hbm_budget)
            throw Error(kResourceLimit, "serve: a request needs more HBM than the whole budget");
    }
    out.assign(reqs.size(), RequestResult{});
    std::vector<std::size_t> order(reqs.size());
    for (std::size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](std::size_t a, std::size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    SectorCache cache(o.cache_sectors, o.cache_budget);
    std::mutex mu;
    std::condition_variable cv;
    std::size_t next = 0;   // the next request to admit, in arrival order (first come, first served)
    std::uint64_t reserved = 0;
    int in_flight = 0;
    ServeStats st;
    double last_progress = 0.0;
    const auto t0 = clk::now();
    auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    auto worker = [&](int slot) {
        hipStream_t stream = nullptr;
        try {
            stream = pooled_stream_acquire();
        } catch (const Error&) {
            stream = nullptr;   // the slot runs on the null stream
        }
        StreamScope scope(stream);
        for (;;) {
            std::size_t idx;
            std::uint64_t fp;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (next >= order.size()) {
                        lk.unlock();
                        pooled_stream_release(stream);
                        return;
                    }
                    const Request& r = reqs[order[next]];
                    const double wait = o.realtime ? r.arrival_s - now_s() : 0.0;
                    fp = request_footprint(r, o);
                    if (wait > 0.0) {
                        cv.wait_for(lk, std::chrono::duration<double>(wait));
                        continue;
                    }
                    if (reserved + fp <= o.hbm_budget) break;
                    cv.wait(lk);   // admitted when requests in flight release their memory
                }
                idx = order[next++];
                reserved += fp;
                st.peak_reserved = std::max(st.peak_reserved, reserved);
                st.peak_in_flight = std::max(st.peak_in_flight, ++in_flight);
            }
            RequestResult& res = out[idx];
            res.slot = slot;
            res.reserved_bytes = fp;
            res.start_s = now_s();
            int builds = 0, hits = 0;
            try {
                solve(reqs[idx], o, cache, res, builds, hits);
            } catch (const Error& e) {
                res.status = e.code;
                res.error = e.what();
            } catch (const std::exception& e) {
                res.status = kRuntime;
                res.error = e.what();
            }
            if (res.status != 0 && stream) (void)hipStreamSynchronize(stream);
            res.finish_s = now_s();
            res.latency_s = res.finish_s - (o.realtime ? reqs[idx].arrival_s : 0.0);
            {
                std::lock_guard<std::mutex> lk(mu);
                reserved -= fp;
                --in_flight;
                st.sector_builds += builds;
                st.sector_hits += hits;
                st.build_s += res.build_s;
                const int kind = reqs[idx].kind;
                if (kind >= 0 && kind < 3) st.service_s[kind] += res.finish_s - res.start_s;
                if (res.status == 0) ++st.completed;
                else ++st.failed;
                const double now = now_s();
                if (o.progress_s > 0.0 && now - last_progress >= o.progress_s) {
                    last_progress = now;
                    std::fprintf(stderr, "[serve] %.1f s: %d of %zu done, %d in flight, %.2f GiB reserved\n", now,
                                 st.completed + st.failed, reqs.size(), in_flight,
                                 static_cast<double>(reserved) / (1ull << 30));
                    std::fflush(stderr);
                }
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> pool;
    for (int s = 0; s < o.slots; ++s) pool.emplace_back(worker, s);
    for (auto& t : pool) t.join();
    st.wall_s = now_s();
    std::vector<double> lat;
    for (const auto& r : out)
        if (r.status == 0) lat.push_back(r.latency_s);
    std::sort(lat.begin(), lat.end());
    if (!lat.empty()) {
        auto pct = [&](double p) {
            const std::size_t i = static_cast<std::size_t>(std::ceil(p * static_cast<double>(lat.size()))) - 1;
            return lat[std::min(i, lat.size() - 1)];
        };
        st.p50_s = pct(0.50);
        st.p99_s = pct(0.99);
        st.max_s = lat.back();
        double s = 0.0;
        for (double x : lat) s += x;
        st.mean_s = s / static_cast<double>(lat.size());
    }
    st.solves_per_s = st.wall_s > 0.0 ? st.completed / st.wall_s : 0.0;
    return st;
}

}  // namespace qsa

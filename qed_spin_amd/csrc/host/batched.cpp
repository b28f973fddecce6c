#include "batched.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <thread>

#include "lanczos.h"
#include "tridiag.h"

namespace qsa {

namespace {

using clock_t_ = std::chrono::steady_clock;
double secs(clock_t_::time_point a, clock_t_::time_point b) { return std::chrono::duration<double>(b - a).count(); }

/// Host worker threads for the per-pair host work (seeds, tridiagonals): at most 16.
unsigned host_threads() {
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    return std::min(16u, hc);
}

/// fn(i) for i < n on the host worker threads (dynamic schedule); the first exception is rethrown.
template <class Fn> void parallel_for(std::size_t n, Fn&& fn) {
    const unsigned T = static_cast<unsigned>(std::min<std::size_t>(host_threads(), n));
    if (T <= 1) {
        for (std::size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<std::size_t> next{0};
    std::exception_ptr failure;
    std::atomic<bool> failed{false};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const std::size_t i = next.fetch_add(1);
                if (i >= n || failed.load()) return;
                try {
                    fn(i);
                } catch (...) {
                    if (!failed.exchange(true)) failure = std::current_exception();
                    return;
                }
            }
        });
    for (auto& th : pool) th.join();
    if (failure) std::rethrow_exception(failure);
}

template <class T> void put(DeviceBuffer& b, const std::vector<T>& h, hipStream_t st) {
    b.reset(h.size() * sizeof(T));
    if (!h.empty()) QSA_HIP(hipMemcpyAsync(b.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    QSA_HIP(hipStreamSynchronize(st));
}

/// A pooled stream at the default priority (graph capture needs a stream other than the null one).
struct OwnStream {
    hipStream_t s = nullptr;
    OwnStream() : s(pooled_stream_acquire()) {}
    ~OwnStream() { pooled_stream_release(s); }
};

}  // namespace

void parallel_blocks(std::size_t n, const std::function<void(std::size_t)>& fn) { parallel_for(n, fn); }

BatchedLanczos::BatchedLanczos(std::vector<std::shared_ptr<const SectorOperator>> blocks, int R)
    : ops_(std::move(blocks)), R_(R) {
    if (ops_.empty()) throw Error(kInvalid, "batched lanczos: no blocks");
    if (R_ < 1 || R_ > dev::kBatchMaxR)
        throw Error(kInvalid, "batched lanczos: vectors per block must be 1.." + std::to_string(dev::kBatchMaxR));
    row_off_.assign(1, 0);
    std::vector<std::uint64_t> nnz_off(1, 0);
    for (const auto& op : ops_) {
        if (!op) throw Error(kInvalid, "batched lanczos: null block");
        if (!op->same()) throw Error(kInvalid, "batched lanczos: a block must map its sector to itself");
        if (op->lane() != kLaneCsr)
            throw Error(kDeviceUnsupported, "batched lanczos: every block needs its device CSR (csr lane)");
        if (op->dim() <= 0) throw Error(kInvalid, "batched lanczos: empty block");
        row_off_.push_back(row_off_.back() + static_cast<std::uint64_t>(op->dim()));
        nnz_off.push_back(nnz_off.back() + static_cast<std::uint64_t>(op->csr_info().nnz));
    }
    rows_ = row_off_.back();
    nnz_ = nnz_off.back();
    if (rows_ >= (std::uint64_t{1} << 32)) throw Error(kInvalid, "batched lanczos: more than 2^32 rows in one batch");
    const hipStream_t st = engine_stream();
    row_ptr_.reset((rows_ + 1) * 8);
    col_.reset(std::max<std::uint64_t>(nnz_, 1) * 4);
    // Values: with fewer than 8 vectors per block (the value stream is a large share of the bytes)
    // and when every block stores dictionary ids and the dictionaries hold at most 2^16 values
    // together, 16-bit ids into their concatenation (2 bytes per entry instead of 16); else in full
    // (with 8 or more vectors the gathers dominate and a full value is one broadcast load).
    std::vector<std::uint32_t> dict_off(ops_.size(), 0);
    bool ids = R_ < 8;
    std::uint64_t dict_total = 0;
    for (std::size_t b = 0; ids && b < ops_.size(); ++b) {
        const dev::CsrView A = ops_[b]->device_csr();
        if (A.width < 1 || A.width > 3 || A.dict_n < 1 || !A.dict_re || (!A.values_real && !A.dict_im)) {
            ids = false;
            break;
        }
        dict_off[b] = static_cast<std::uint32_t>(dict_total);
        dict_total += static_cast<std::uint64_t>(A.dict_n);
    }
    if (dict_total > (std::uint64_t{1} << 16)) ids = false;
    if (ids) {
        dict_n_ = static_cast<std::uint32_t>(dict_total);
        vid_.reset(std::max<std::uint64_t>(nnz_, 1) * 2);
        dict_re_.reset(dict_total * 8);
        dict_im_.reset(dict_total * 8);
        for (std::size_t b = 0; b < ops_.size(); ++b) {
            const dev::CsrView A = ops_[b]->device_csr();
            const std::size_t n = static_cast<std::size_t>(A.dict_n) * 8;
            QSA_HIP(hipMemcpyAsync(dict_re_.get<double>() + dict_off[b], A.dict_re, n, hipMemcpyDeviceToDevice, st));
            if (A.values_real)
                dev::fill_zero(false, dict_im_.get<double>() + dict_off[b], static_cast<std::size_t>(A.dict_n), st);
            else
                QSA_HIP(hipMemcpyAsync(dict_im_.get<double>() + dict_off[b], A.dict_im, n, hipMemcpyDeviceToDevice, st));
        }
    } else {
        val_.reset(std::max<std::uint64_t>(nnz_, 1) * sizeof(dev::Zd));
    }
    for (std::size_t b = 0; b < ops_.size(); ++b)
        dev::pack_csr(ops_[b]->device_csr(), row_off_[b], row_off_[b], nnz_off[b], row_ptr_.get<std::uint64_t>(),
                      col_.get<std::uint32_t>(), ids ? nullptr : val_.get<dev::Zd>(),
                      ids ? vid_.get<std::uint16_t>() : nullptr, dict_off[b], st);
    QSA_HIP(hipGetLastError());
    // Tiles: every block cut into runs of at most kBatchTileRows rows, blocks in order.
    tile_begin_.assign(1, 0);
    for (std::size_t b = 0; b < ops_.size(); ++b) {
        for (std::uint64_t r = row_off_[b]; r < row_off_[b + 1]; r += dev::kBatchTileRows) {
            dev::Tile t{};
            t.row_begin = r;
            t.row_end = std::min<std::uint64_t>(r + dev::kBatchTileRows, row_off_[b + 1]);
            t.block = static_cast<std::uint32_t>(b);
            tiles_.push_back(t);
        }
        tile_begin_.push_back(static_cast<std::uint32_t>(tiles_.size()));
    }
    put(tiles_d_, tiles_, st);
    put(tile_begin_d_, tile_begin_, st);
    const std::size_t vec_bytes = rows_ * static_cast<std::size_t>(R_) * sizeof(dev::Zd);
    V0_.reset(vec_bytes);
    V1_.reset(vec_bytes);
    W_.reset(vec_bytes);
    partial_.reset(tiles_.size() * 64 * 2 * sizeof(double));
    partial2_.reset(tiles_.size() * 64 * 2 * sizeof(double));
    fix_.reset(ops_.size() * static_cast<std::size_t>(R_) * sizeof(int));
    put(fix_, std::vector<int>(ops_.size() * static_cast<std::size_t>(R_), 0), st);
    const std::size_t pairs = ops_.size() * static_cast<std::size_t>(R_);
    ov_.reset(pairs * sizeof(dev::Zd));
    cur_a_.reset(pairs * 8);
    cur_b_.reset(pairs * 8);
    inv_b_.reset(pairs * 8);
    active_d_.reset(pairs * sizeof(int));
    steps_d_.reset(pairs * sizeof(int));
    step_d_.reset(sizeof(int));
    cap_d_.reset(ops_.size() * sizeof(int));
    tol_d_.reset(ops_.size() * sizeof(double));
    dev::batch_fill_zero(V1_.get<dev::Zd>(), rows_ * R_, st);
    dev::batch_fill_zero(W_.get<dev::Zd>(), rows_ * R_, st);
    QSA_HIP(hipStreamSynchronize(st));
}

void BatchedLanczos::set_schedule(int kind) {
    sched_kind_ = 0;
    sched_d_.release();
    if (kind != 1 || tiles_.empty()) return;
    // blocks to XCDs, largest first onto the XCD with the fewest rows so far; workgroup w runs on XCD
    // w % 8, so it takes the next tile of that XCD's blocks (of the longest remaining queue when that
    // XCD has none left)
    constexpr int kXcd = 8;
    std::vector<std::size_t> order(ops_.size());
    for (std::size_t b = 0; b < order.size(); ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) { return dim(a) > dim(b); });
    std::uint64_t load[kXcd] = {};
    std::vector<std::uint32_t> queue[kXcd];
    for (const std::size_t b : order) {
        int x = 0;
        for (int y = 1; y < kXcd; ++y)
            if (load[y] < load[x]) x = y;
        load[x] += dim(static_cast<int>(b));
        for (std::uint32_t t = tile_begin_[b]; t < tile_begin_[b + 1]; ++t) queue[x].push_back(t);
    }
    std::size_t head[kXcd] = {};
    std::vector<std::uint32_t> sched(tiles_.size());
    for (std::size_t w = 0; w < sched.size(); ++w) {
        int x = static_cast<int>(w % kXcd);
        if (head[x] == queue[x].size()) {
            for (int y = 0; y < kXcd; ++y)
                if (queue[y].size() - head[y] > queue[x].size() - head[x]) x = y;
        }
        sched[w] = queue[x][head[x]++];
    }
    put(sched_d_, sched, engine_stream());
    QSA_HIP(hipStreamSynchronize(engine_stream()));
    sched_kind_ = 1;
}

int BatchedLanczos::tune_schedule() {
    if (const char* e = std::getenv("QED_SPIN_AMD_BATCH_SCHEDULE")) {   // pinned: 0 or 1
        set_schedule(std::atoi(e) == 1 ? 1 : 0);
        return sched_kind_;
    }
    // small engines finish a step in microseconds either way; above 65536 tiles the two schedules
    // measured alike and a timed apply costs tens of milliseconds. R = 1 keeps the in-order schedule,
    // which its fused step kernel runs in XCD-contiguous runs of tiles (the plain apply timed here
    // is the sector SpMV, which takes no schedule)
    if (R_ == 1 || tiles_.size() < 4096 || tiles_.size() > 65536) {
        set_schedule(0);
        return 0;
    }
    // one warm-up apply, then the two schedules alternately (0, 1, 0, 1), each apply timed
    const hipStream_t st = engine_stream();
    hipEvent_t ev[5];
    for (auto& e : ev) QSA_HIP(hipEventCreate(&e));
    set_schedule(1);
    const std::uint32_t* grouped = sched_d_.get<std::uint32_t>();
    DeviceBuffer keep = std::move(sched_d_);
    dev::BatchView B = packed_view();
    B.sched = nullptr;
    dev::batch_spmv(B, V1_.get<dev::Zd>(), W_.get<dev::Zd>(), st);
    QSA_HIP(hipEventRecord(ev[0], st));
    for (int i = 0; i < 4; ++i) {
        B.sched = (i & 1) ? grouped : nullptr;
        dev::batch_spmv(B, V1_.get<dev::Zd>(), W_.get<dev::Zd>(), st);
        QSA_HIP(hipEventRecord(ev[i + 1], st));
    }
    QSA_HIP(hipEventSynchronize(ev[4]));
    float ms[2] = {0.0f, 0.0f};
    for (int i = 0; i < 4; ++i) {
        float m = 0.0f;
        QSA_HIP(hipEventElapsedTime(&m, ev[i], ev[i + 1]));
        ms[i & 1] += m;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    const int best = ms[1] < ms[0] ? 1 : 0;
    dev::batch_fill_zero(W_.get<dev::Zd>(), rows_ * R_, st);
    QSA_HIP(hipStreamSynchronize(st));
    set_schedule(best);
    return best;
}

dev::BatchView BatchedLanczos::view() const {
    dev::BatchView B;
    B.rows = rows_;
    B.n_blocks = static_cast<std::uint32_t>(ops_.size());
    B.n_tiles = static_cast<std::uint32_t>(tiles_.size());
    B.R = R_;
    B.tiles = tiles_d_.get<dev::Tile>();
    B.sched = sched_d_ ? sched_d_.get<std::uint32_t>() : nullptr;
    B.tile_begin = tile_begin_d_.get<std::uint32_t>();
    B.row_ptr = row_ptr_.get<std::uint64_t>();
    B.col = col_.get<std::uint32_t>();
    B.val = val_.get<dev::Zd>();
    if (dict_n_) {
        B.vid = vid_.get<std::uint16_t>();
        B.dict_re = dict_re_.get<double>();
        B.dict_im = dict_im_.get<double>();
        B.dict_n = dict_n_;
    }
    B.step = step_d_.get<int>();
    B.active = active_d_.get<int>();
    B.steps = steps_d_.get<int>();
    B.cap = cap_d_.get<int>();
    B.fix = fix_.get<int>();
    return B;
}

void BatchedLanczos::set_start(const cplx* host) {
    copy_to_device(V0_.get(), host, rows_ * R_ * sizeof(dev::Zd));
    normalize(nullptr);
}

void BatchedLanczos::clear_start() {
    dev::batch_fill_zero(V0_.get<dev::Zd>(), rows_ * R_, engine_stream());
    QSA_HIP(hipGetLastError());
}

void BatchedLanczos::place(int b, int s, const cplx* device_vector) {
    if (b < 0 || b >= n_blocks() || s < 0 || s >= R_) throw Error(kInvalid, "batched lanczos: no such pair");
    dev::batch_place(V0_.get<dev::Zd>(), R_, row_off_[b], s, reinterpret_cast<const dev::Zd*>(device_vector), dim(b),
                     engine_stream());
    QSA_HIP(hipGetLastError());
}

void BatchedLanczos::normalize(std::vector<double>* norm2) {
    const hipStream_t st = engine_stream();
    dev::BatchView B = view();
    B.active = nullptr;
    const std::size_t pairs = ops_.size() * static_cast<std::size_t>(R_);
    DeviceBuffer n2(norm2 ? pairs * 8 : 0);
    dev::batch_normalize(B, V0_.get<dev::Zd>(), partial_.get<double>(), inv_b_.get<double>(),
                         norm2 ? n2.get<double>() : nullptr, st);
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipStreamSynchronize(st));
    if (norm2) {
        norm2->resize(pairs);
        copy_to_host(norm2->data(), n2.get(), pairs * 8);
    }
}

void BatchedLanczos::sample_start(const std::vector<std::uint32_t>& seeds) {
    const std::size_t nb = ops_.size(), pairs = nb * static_cast<std::size_t>(R_);
    if (seeds.size() != pairs) throw Error(kInvalid, "batched lanczos: one seed per pair");
    const hipStream_t st = engine_stream();
    std::vector<std::uint64_t> dims(nb);
    for (std::size_t b = 0; b < nb; ++b) dims[b] = row_off_[b + 1] - row_off_[b];
    DeviceBuffer rb_d, dim_d, seed_d;
    put(rb_d, row_off_, st);
    put(dim_d, dims, st);
    put(seed_d, seeds, st);
    dev::sample_gauss(V0_.get<dev::Zd>(), R_, rb_d.get<std::uint64_t>(), dim_d.get<std::uint64_t>(),
                      seed_d.get<std::uint32_t>(), static_cast<std::uint32_t>(pairs), st);
    QSA_HIP(hipGetLastError());
    normalize(nullptr);
}

void BatchedLanczos::apply(const cplx* X, cplx* Y) const {
    dev::BatchView B = view();
    B.active = nullptr;
    dev::batch_spmv(B, reinterpret_cast<const dev::Zd*>(X), reinterpret_cast<dev::Zd*>(Y), engine_stream());
    QSA_HIP(hipGetLastError());
}

void BatchedLanczos::run(int M, const std::vector<double>& breakdown_tol, bool local, bool use_graph,
                         std::vector<double>& alpha, std::vector<double>& beta, std::vector<int>& steps,
                         BatchedTimings& t, const std::vector<int>* live) {
    if (M < 1) throw Error(kInvalid, "batched lanczos: at least one step");
    if (breakdown_tol.size() != ops_.size()) throw Error(kInvalid, "batched lanczos: one breakdown tolerance per block");
    const std::size_t nb = ops_.size(), pairs = nb * static_cast<std::size_t>(R_);
    const std::size_t hist = static_cast<std::size_t>(M) * pairs;
    // Graph capture needs a stream other than the null one: a stream of its own when the engine
    // runs on the null stream (everything before it on the null stream is finished first).
    const hipStream_t outer = engine_stream();
    std::unique_ptr<OwnStream> own;
    if (!outer) {
        QSA_HIP(hipStreamSynchronize(nullptr));
        own = std::make_unique<OwnStream>();
    }
    const hipStream_t st = outer ? outer : own->s;
    StreamScope scope(st);
    std::vector<int> cap(nb), ones(pairs, 1), zeros(pairs, 0);
    if (live) {
        if (live->size() != pairs) throw Error(kInvalid, "batched lanczos: one live flag per pair");
        ones = *live;
    }
    for (std::size_t b = 0; b < nb; ++b) cap[b] = static_cast<int>(std::min<std::uint64_t>(M, dim(static_cast<int>(b))));
    put(cap_d_, cap, st);
    put(tol_d_, breakdown_tol, st);
    put(active_d_, ones, st);
    put(steps_d_, zeros, st);
    put(step_d_, std::vector<int>{0}, st);
    alpha_h_.reset(hist * 8);
    beta_h_.reset(hist * 8);
    dev::fill_zero(false, alpha_h_.get(), hist, st);
    dev::fill_zero(false, beta_h_.get(), hist, st);
    dev::fill_zero(false, cur_b_.get(), pairs, st);
    put(scale_, std::vector<double>(3 * pairs, 1.0), st);
    QSA_HIP(hipStreamSynchronize(st));

    const dev::BatchView B = view();
    dev::Zd* V[3] = {V1_.get<dev::Zd>(), V0_.get<dev::Zd>(), W_.get<dev::Zd>()};
    double* partial = partial_.get<double>();
    double* partial2 = partial2_.get<double>();
    double* ah = alpha_h_.get<double>();
    double* bh = beta_h_.get<double>();
    dev::Zd* ov = ov_.get<dev::Zd>();
    double* ca = cur_a_.get<double>();
    double* cb = cur_b_.get<double>();
    double* ib = inv_b_.get<double>();
    const double* tol = tol_d_.get<double>();
    const double thr = 1.49011611938476562e-08;   // sqrt(eps): LocalDGKS3's overlap threshold
    // With the fused SpMV (R = 1 or R >= 8) the next vector is never scaled in place: each of the three
    // buffers keeps a per-pair factor (1 / beta of the step that wrote it; 1 for the normalised
    // start) that every read applies, which rounds as the scaling pass would (same results, one
    // pass over the vectors fewer per step).
    const bool defer = R_ == 1 || R_ >= 8;
    const std::size_t np = pairs;
    double* S[3] = {scale_.get<double>(), scale_.get<double>() + np, scale_.get<double>() + 2 * np};
    auto step = [&](int phase) {
        dev::Zd* p = V[phase % 3];
        dev::Zd* c = V[(phase + 1) % 3];
        dev::Zd* w = V[(phase + 2) % 3];
        const double* sp = defer ? S[phase % 3] : nullptr;
        const double* sc = defer ? S[(phase + 1) % 3] : nullptr;
        double* sw = defer ? S[(phase + 2) % 3] : ib;
        if (!dev::batch_spmv_fused(B, p, c, w, cb, partial, st, sp, sc)) {
            dev::batch_spmv(B, c, w, st);
            dev::batch_vec(B, 0, p, c, w, cb, nullptr, 0.0, partial, st);
        }
        dev::batch_finish(B, 0, partial, ah, bh, ov, ca, cb, ib, tol, st);
        if (local) {
            // W -= alpha V with the local overlap and ||W||^2 in one pass; the correction pass then
            // touches only the pairs whose overlap asks for one (their norms are summed again)
            dev::batch_vec(B, 5, nullptr, c, w, ca, nullptr, 0.0, partial, st, partial2, sc);
            dev::batch_finish(B, 1, partial, ah, bh, ov, ca, cb, ib, tol, st, thr);
            dev::batch_vec(B, 2, nullptr, c, w, nullptr, ov, thr, partial, st, partial2, sc);
            dev::batch_finish(B, 2, partial2, ah, bh, ov, ca, cb, sw, tol, st);
        } else {
            dev::batch_vec(B, 4, nullptr, c, w, ca, nullptr, 0.0, partial, st, nullptr, sc);
            dev::batch_finish(B, 2, partial, ah, bh, ov, ca, cb, sw, tol, st);
        }
        if (!defer) dev::batch_vec(B, 3, nullptr, nullptr, w, ib, nullptr, 0.0, partial, st);
        dev::batch_advance(B, st);
    };
    const int Mrun = *std::max_element(cap.begin(), cap.end());
    hipEvent_t e0, e1;
    QSA_HIP(hipEventCreate(&e0));
    QSA_HIP(hipEventCreate(&e1));
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    auto cleanup = [&] {
        for (int i = 0; i < 3; ++i) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    };
    try {
        if (use_graph) {
            for (int ph = 0; ph < 3 && ph < Mrun; ++ph) {
                {
                    const auto no_fence = capture_guard();   // no fence touches st while it captures
                    QSA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    try {
                        step(ph);
                    } catch (...) {
                        hipGraph_t broken = nullptr;
                        (void)hipStreamEndCapture(st, &broken);
                        if (broken) (void)hipGraphDestroy(broken);
                        throw;
                    }
                    QSA_HIP(hipStreamEndCapture(st, &graph[ph]));
                }
                QSA_HIP(hipGraphInstantiate(&exec[ph], graph[ph], nullptr, nullptr, 0));
            }
        }
        QSA_HIP(hipEventRecord(e0, st));
        for (int j = 0; j < Mrun; ++j) {
            if (use_graph) QSA_HIP(hipGraphLaunch(exec[j % 3], st));
            else step(j % 3);
        }
        QSA_HIP(hipEventRecord(e1, st));
        QSA_HIP(hipGetLastError());
        QSA_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        QSA_HIP(hipEventElapsedTime(&ms, e0, e1));
        t.steps_s += ms * 1e-3;
        t.steps += Mrun;
        t.graph = use_graph;
    } catch (...) {
        (void)hipStreamSynchronize(st);
        cleanup();
        throw;
    }
    cleanup();
    alpha.resize(hist);
    beta.resize(hist);
    steps.resize(pairs);
    copy_to_host(alpha.data(), ah, hist * 8);
    copy_to_host(beta.data(), bh, hist * 8);
    copy_to_host(steps.data(), steps_d_.get(), pairs * sizeof(int));
    for (int s : steps) t.applies += static_cast<std::uint64_t>(s);
    t.rows = rows_;
    t.nnz = nnz_;
}

void BatchedLanczos::run_kept(int M, const std::vector<double>& breakdown_tol, const std::vector<int>& full,
                              cplx* basis_c, std::vector<double>& alpha, std::vector<double>& beta,
                              std::vector<int>& steps, BatchedTimings& t, const std::vector<int>* live) {
    if (M < 1) throw Error(kInvalid, "batched lanczos: at least one step");
    const std::size_t nb = ops_.size(), pairs = nb * static_cast<std::size_t>(R_);
    if (breakdown_tol.size() != nb || full.size() != nb)
        throw Error(kInvalid, "batched lanczos: one breakdown tolerance and one reorthogonalisation flag per block");
    if (live && live->size() != pairs) throw Error(kInvalid, "batched lanczos: one live flag per pair");
    const hipStream_t st = engine_stream();
    const auto t0 = clock_t_::now();
    const std::size_t hist = static_cast<std::size_t>(M) * pairs;
    std::vector<int> cap(nb), zeros(pairs, 0);
    for (std::size_t b = 0; b < nb; ++b) cap[b] = static_cast<int>(std::min<std::uint64_t>(M, dim(static_cast<int>(b))));
    put(cap_d_, cap, st);
    put(tol_d_, breakdown_tol, st);
    put(active_d_, live ? *live : std::vector<int>(pairs, 1), st);
    put(steps_d_, zeros, st);
    put(step_d_, std::vector<int>{0}, st);
    put(fix_, zeros, st);
    DeviceBuffer full_d;
    put(full_d, full, st);
    const bool any_full = std::any_of(full.begin(), full.end(), [](int f) { return f != 0; });
    DeviceBuffer coef(any_full ? nb * static_cast<std::size_t>(M + 1) * R_ * sizeof(dev::Zd) : 8);
    DeviceBuffer gs2(pairs * sizeof(int));   // per pair: the second Gram-Schmidt pass of this step
    alpha_h_.reset(hist * 8);
    beta_h_.reset(hist * 8);
    dev::fill_zero(false, alpha_h_.get(), hist, st);
    dev::fill_zero(false, beta_h_.get(), hist, st);
    dev::fill_zero(false, cur_b_.get(), pairs, st);
    const std::uint64_t layer = rows_ * static_cast<std::uint64_t>(R_);
    auto* basis = reinterpret_cast<dev::Zd*>(basis_c);
    QSA_HIP(hipMemcpyAsync(basis, V0_.get(), layer * sizeof(dev::Zd), hipMemcpyDeviceToDevice, st));
    const dev::BatchView B = view();
    double* partial = partial_.get<double>();
    double* partial2 = partial2_.get<double>();
    double* ah = alpha_h_.get<double>();
    double* bh = beta_h_.get<double>();
    dev::Zd* ov = ov_.get<dev::Zd>();
    double* ca = cur_a_.get<double>();
    double* cb = cur_b_.get<double>();
    double* ib = inv_b_.get<double>();
    const double* tol = tol_d_.get<double>();
    const double thr = 1.49011611938476562e-08;   // sqrt(eps): LocalDGKS3's overlap threshold
    const int Mrun = *std::max_element(cap.begin(), cap.end());
    for (int j = 0; j < Mrun; ++j) {
        const dev::Zd* Vp = basis + static_cast<std::uint64_t>(j > 0 ? j - 1 : 0) * layer;
        const dev::Zd* Vc = basis + static_cast<std::uint64_t>(j) * layer;
        dev::Zd* W = basis + static_cast<std::uint64_t>(j + 1) * layer;
        // the kept layers are stored normalised (the overlaps read them later), so only the first
        // recurrence pass is fused here, not the deferred scaling of run()
        if (!dev::batch_spmv_fused(B, Vp, Vc, W, cb, partial, st)) {
            dev::batch_spmv(B, Vc, W, st);
            dev::batch_vec(B, 0, Vp, Vc, W, cb, nullptr, 0.0, partial, st);
        }
        dev::batch_finish(B, 0, partial, ah, bh, ov, ca, cb, ib, tol, st);
        dev::batch_vec(B, 5, nullptr, Vc, W, ca, nullptr, 0.0, partial, st, partial2);
        dev::batch_finish(B, 1, partial, ah, bh, ov, ca, cb, ib, tol, st, thr);
        dev::batch_vec(B, 2, nullptr, Vc, W, nullptr, ov, thr, partial, st, partial2);
        if (any_full) {
            // CGS2 against every kept layer; the second pass only where the first removed more than
            // 1 - kappa of ||W|| (DGKS, as the reference's FullCGS2 gates it): partial2 holds the
            // norms before the first pass, partial those after it
            dev::batch_gram_schmidt(B, basis, layer, j + 1, M + 1, W, full_d.get<int>(), coef.get<dev::Zd>(), st);
            dev::batch_norm_partial(B, W, partial, st);
            dev::batch_dgks_gate(B, partial2, partial, full_d.get<int>(), kDgksKappa, gs2.get<int>(), st);
            dev::batch_gram_schmidt(B, basis, layer, j + 1, M + 1, W, full_d.get<int>(), coef.get<dev::Zd>(), st,
                                    gs2.get<int>());
            dev::batch_norm_partial(B, W, partial2, st);
        }
        dev::batch_finish(B, 2, partial2, ah, bh, ov, ca, cb, ib, tol, st);
        dev::batch_vec(B, 3, nullptr, nullptr, W, ib, nullptr, 0.0, partial, st);
        dev::batch_advance(B, st);
    }
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipStreamSynchronize(st));
    t.steps_s += secs(t0, clock_t_::now());
    t.steps += Mrun;
    alpha.resize(hist);
    beta.resize(hist);
    steps.resize(pairs);
    copy_to_host(alpha.data(), ah, hist * 8);
    copy_to_host(beta.data(), bh, hist * 8);
    copy_to_host(steps.data(), steps_d_.get(), pairs * sizeof(int));
    for (int s : steps) t.applies += static_cast<std::uint64_t>(s);
    t.rows = rows_;
    t.nnz = nnz_;
}

std::vector<Curves> batched_ftlm(const std::vector<std::shared_ptr<const SectorOperator>>& blocks,
                                 const BatchedFtlmOptions& o, BatchedTimings& t) {
    const auto t_all = clock_t_::now();
    const std::size_t nb = blocks.size();
    if (nb == 0) return {};
    if (o.samples < 1 || o.krylov < 1 || o.betas.empty()) throw Error(kInvalid, "ftlm: bad sizes");
    if (o.block_seeds.size() != nb || o.breakdown_tol.size() != nb)
        throw Error(kInvalid, "ftlm: one seed and one breakdown tolerance per block");
    const auto t_pack = clock_t_::now();
    const int R = std::min(o.samples, dev::kBatchMaxR);
    BatchedLanczos L(blocks, R);
    t.pack_s += secs(t_pack, clock_t_::now());
    const auto t_tune = clock_t_::now();
    t.schedule = L.tune_schedule();
    t.tune_s += secs(t_tune, clock_t_::now());
    std::vector<RitzPool> pools(nb);
    std::vector<double> alpha, beta;
    std::vector<int> steps;
    for (int s0 = 0; s0 < o.samples; s0 += R) {
        const int Rc = std::min(R, o.samples - s0);
        // Samples of this round: pair (block, s) starts from sample s0 + s of the block's stream (the
        // 32-bit Mersenne Twister seed of sample_engine), drawn on the device; the unused columns of
        // a short last round repeat its first sample (never read back).
        const auto t_seed = clock_t_::now();
        std::vector<std::uint32_t> seeds(nb * static_cast<std::size_t>(R));
        for (std::size_t b = 0; b < nb; ++b)
            for (int s = 0; s < R; ++s) {
                const int sample = s0 + (s < Rc ? s : 0);
                seeds[b * R + s] = static_cast<std::uint32_t>(
                    splitmix_seed(o.block_seeds[b], static_cast<std::uint64_t>(sample)));
            }
        L.sample_start(seeds);
        t.seed_s += secs(t_seed, clock_t_::now());
        L.run(o.krylov, o.breakdown_tol, /*local=*/true, o.use_graph, alpha, beta, steps, t);
        const auto t_post = clock_t_::now();
        const std::size_t pairs = nb * static_cast<std::size_t>(R);
        std::vector<std::vector<double>> ritz(nb * static_cast<std::size_t>(Rc)), weight(ritz.size());
        parallel_for(nb * static_cast<std::size_t>(Rc), [&](std::size_t i) {
            const std::size_t b = i / Rc, s = i % Rc;
            const std::size_t pr = b * R + s;
            const int m = steps[pr];
            std::vector<double> a(static_cast<std::size_t>(m)), bn(static_cast<std::size_t>(m));
            for (int j = 0; j < m; ++j) {
                a[j] = alpha[static_cast<std::size_t>(j) * pairs + pr];
                bn[j] = beta[static_cast<std::size_t>(j) * pairs + pr];
            }
            tridiag_weights(a.data(), bn.data(), static_cast<std::size_t>(m), ritz[i], weight[i]);
        });
        for (std::size_t b = 0; b < nb; ++b)   // samples pooled in sample order
            for (int s = 0; s < Rc; ++s)
                if (!ritz[b * Rc + s].empty()) pools[b].add_sample(ritz[b * Rc + s], weight[b * Rc + s]);
        t.post_s += secs(t_post, clock_t_::now());
    }
    std::vector<Curves> out(nb);
    for (std::size_t b = 0; b < nb; ++b) out[b] = pools[b].curves(L.dim(static_cast<int>(b)), o.betas);
    t.total_s += secs(t_all, clock_t_::now());
    return out;
}

}  // namespace qsa

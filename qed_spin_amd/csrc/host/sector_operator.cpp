#include "sector_operator.h"

#include <chrono>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <string>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace qsa {

SectorOperator::SectorOperator(std::shared_ptr<const DeviceSector> rows, std::shared_ptr<const DeviceSector> cols,
                               std::shared_ptr<const DeviceProgram> prog, double norm_bound, int lane,
                               std::uint64_t csr_budget_bytes, bool tune)
    : rows_(std::move(rows)), cols_(cols ? std::move(cols) : rows_), prog_(std::move(prog)), norm_bound_(norm_bound),
      tune_(tune) {
    if (!rows_ || !prog_) throw Error(kInvalid, "SectorOperator: missing sector or program");
    const auto& a = rows_->spec();
    const auto& b = cols_->spec();
    if (a.n_sites != b.n_sites || a.G != b.G || a.perms != b.perms || a.flips != b.flips)
        throw Error(kInvalid, "SectorOperator: the two sectors come from different groups");
    if (a.slc.engaged() != b.slc.engaged() || a.slc.key_bit != b.slc.key_bit)
        throw Error(kInvalid, "SectorOperator: the two sectors use different representative orders");
    if (rows_->irrep_dim() != cols_->irrep_dim())
        throw Error(kInvalid, "SectorOperator: the two sectors carry irreps of different dimensions");
    // between two sectors of an irrep of dimension d > 1 the rows are blocks written for one irrep
    // and an operator that commutes with the group (total S+-, S^2): the reference's
    // CrossSectorMatVec refuses any other pair (src/engine/internal.h:557-563)
    if (rows_->irrep_dim() > 1 && !same() && rows_->irrep().D != cols_->irrep().D)
        throw Error(kInvalid, "SectorOperator: two sectors of different irreps of dimension > 1");
    real_ = prog_->real && rows_->real_characters() && cols_->real_characters();
    lane_ = kLaneGather;
    if (rows_->irrep_dim() > 1) {
        // a representative's states mix under the group: only the CSR lane, complex values
        if (lane == kLaneGather)
            throw Error(kDeviceUnsupported, "SectorOperator: sectors of irreps of dimension > 1 take the CSR lane");
        real_ = false;
        if (dim() > 0) build_csr_irrep(csr_budget_bytes);
        if (lane_ != kLaneCsr && dim() > 0)
            throw Error(kResourceLimit, "SectorOperator: the CSR of an irrep sector does not fit its budget");
        return;
    }
    if (lane != kLaneGather && dim() > 0) build_csr(csr_budget_bytes);
    if (lane == kLaneCsr && lane_ != kLaneCsr)
        throw Error(kResourceLimit, "SectorOperator: the CSR lane was requested but does not fit its budget");
}

namespace {

/// Stage buffers of finished CSR builds, kept for the next builds on the same device: the
/// stream-ordered pool can take a second to map a new block of several GB once its free memory is
/// split, while a kept buffer costs nothing. At most a quarter of the device's memory (and 48 GiB) is
/// kept; the cache lives until the process ends and is never destroyed (no HIP call at exit). A
/// buffer is given back only after the build's stream has finished with it.
class StageCache {
public:
    DeviceBuffer take(std::size_t bytes) {
        int dev = 0;
        QSA_HIP(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(m_);
            std::size_t best = kept_.size();
            for (std::size_t i = 0; i < kept_.size(); ++i)
                if (kept_[i].dev == dev && kept_[i].buf.bytes() >= bytes &&
                    (best == kept_.size() || kept_[i].buf.bytes() < kept_[best].buf.bytes()))
                    best = i;
            if (best < kept_.size()) {
                DeviceBuffer b = std::move(kept_[best].buf);
                held_ -= b.bytes();
                kept_.erase(kept_.begin() + static_cast<std::ptrdiff_t>(best));
                return b;
            }
        }
        return DeviceBuffer(bytes);
    }
    void give(DeviceBuffer b) {
        if (!b) return;
        int dev = 0;
        QSA_HIP(hipGetDevice(&dev));
        std::vector<DeviceBuffer> drop;
        {
            std::lock_guard<std::mutex> lk(m_);
            if (cap_ == 0) {
                std::size_t free_b = 0, total_b = 0;
                QSA_HIP(hipMemGetInfo(&free_b, &total_b));
                cap_ = std::min<std::size_t>(std::size_t{48} << 30, total_b / 4);
            }
            held_ += b.bytes();
            kept_.push_back({dev, std::move(b)});
            while (held_ > cap_ && !kept_.empty()) {   // the oldest go first
                held_ -= kept_.front().buf.bytes();
                drop.push_back(std::move(kept_.front().buf));
                kept_.erase(kept_.begin());
            }
        }
    }

private:
    struct Kept {
        int dev;
        DeviceBuffer buf;
    };
    std::mutex m_;
    std::vector<Kept> kept_;
    std::size_t held_ = 0, cap_ = 0;
};

StageCache& stage_cache() {
    static StageCache* c = new StageCache;
    return *c;
}

}  // namespace

void SectorOperator::build_csr(std::uint64_t budget) {
    const auto t0 = std::chrono::steady_clock::now();
    // QED_SPIN_AMD_BUILD_PROFILE set: the host seconds of each phase on stderr (a diagnostic)
    static const bool phases = std::getenv("QED_SPIN_AMD_BUILD_PROFILE") != nullptr;
    std::vector<std::pair<const char*, double>> marks;
    auto mark = [&](const char* what) {
        if (phases) marks.emplace_back(what, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    };
    const hipStream_t st = engine_stream();
    const std::int64_t rows = dim();
    const dev::SectorView R = rows_->view(), C = cols_->view();
    const dev::ProgramView P = prog_->view();
    DeviceBuffer row_len(static_cast<std::size_t>(rows) * 4);
    DeviceBuffer dfp(dev::DictTable::kSlots * 8), dre(dev::DictTable::kSlots * 8), dim_(dev::DictTable::kSlots * 8);
    DeviceBuffer dslot(dev::DictTable::kSlots * 2), dcnt(8);
    dev::DictTable D;
    D.fp = dfp.get<std::uint64_t>();
    D.re = dre.get<std::uint64_t>();
    D.im = dim_.get<std::uint64_t>();
    D.slot_id = dslot.get<std::uint16_t>();
    D.count = dcnt.get<unsigned>();
    D.flags = dcnt.get<unsigned>() + 1;
    QSA_HIP(hipMemsetAsync(dfp.get(), 0, dfp.bytes(), st));
    QSA_HIP(hipMemsetAsync(dcnt.get(), 0, 8, st));
    // The walk stages (acting groups) slots per row. Rows go in chunks whose stage fits at most
    // 16 GiB and a quarter of the free device memory: with one chunk the CSR is compacted from the
    // stage the walk left; with several, a first pass over the chunks gives the row lengths and the
    // dictionary, a second walks them again and compacts each into place.
    const std::uint64_t slot_bytes = static_cast<std::uint64_t>(std::max<std::uint32_t>(P.n_groups, 1)) *
                                     (4 + sizeof(dev::Zd));
    std::size_t free_b = 0, total_b = 0;
    QSA_HIP(hipMemGetInfo(&free_b, &total_b));
    std::uint64_t cap = std::min<std::uint64_t>(std::uint64_t{16} << 30, free_b / 4);
    if (const char* e = std::getenv("QED_SPIN_AMD_CSR_STAGE_BYTES")) {   // a lower cap (tests, tight memory)
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) cap = std::min<std::uint64_t>(cap, v);
    }
    const std::int64_t chunk = std::min<std::int64_t>(rows, static_cast<std::int64_t>(cap / slot_bytes));
    if (chunk < std::min<std::int64_t>(rows, 256)) return;   // no room to stage: the gather lane
    const std::int64_t n_chunks = (rows + chunk - 1) / chunk;
    mark("set-up");
    // one buffer from the cache: the columns, then (256-byte aligned) the values
    const std::size_t slots = static_cast<std::size_t>(chunk) * std::max<std::uint32_t>(P.n_groups, 1);
    const std::size_t col_bytes = (slots * 4 + 255) / 256 * 256;
    struct Stage {
        DeviceBuffer buf;
        hipStream_t st;
        ~Stage() {
            if (!buf) return;
            (void)hipStreamSynchronize(st);   // the build's kernels are done with it (already, on every path)
            stage_cache().give(std::move(buf));
        }
    } stage{stage_cache().take(col_bytes + slots * sizeof(dev::Zd)), st};
    std::uint32_t* const stage_col = stage.buf.get<std::uint32_t>();
    dev::Zd* const stage_val = reinterpret_cast<dev::Zd*>(stage.buf.get<char>() + col_bytes);
    mark("stage alloc");
    auto walk = [&](std::int64_t c) {
        const std::int64_t r0 = c * chunk, n = std::min<std::int64_t>(chunk, rows - r0);
        return dev::csr_walk(R, C, same(), P, r0, n, row_len.get<std::uint32_t>() + r0, stage_col, stage_val, st);
    };
    for (std::int64_t c = 0; c < n_chunks; ++c) {
        if (!walk(c)) return;   // too many groups
        const std::int64_t r0 = c * chunk, n = std::min<std::int64_t>(chunk, rows - r0);
        dev::csr_stage_dict(n, row_len.get<std::uint32_t>() + r0, stage_val, D, st);
    }
    QSA_HIP(hipGetLastError());
    auto fill = [&](int w) {
        for (std::int64_t c = 0; c < n_chunks; ++c) {
            const std::int64_t r0 = c * chunk, n = std::min<std::int64_t>(chunk, rows - r0);
            if (n_chunks > 1) walk(c);
            dev::csr_compact(n, row_ptr_.get<std::uint64_t>() + r0, stage_col, stage_val, col_.get<std::uint32_t>(),
                             vals_.get(), w, real_, D, st);
        }
    };
    row_ptr_.reset(static_cast<std::size_t>(rows + 1) * 8);
    const std::size_t tb = dev::row_scan_temp_bytes(rows);
    DeviceBuffer temp(tb ? tb : 8);
    dev::row_scan(row_len.get<std::uint32_t>(), rows, row_ptr_.get<std::uint64_t>(), temp.get(), tb, st);
    QSA_HIP(hipGetLastError());
    std::uint64_t nnz = 0;
    unsigned cnt[2] = {0, 0};
    QSA_HIP(hipMemcpyAsync(&nnz, row_ptr_.get<std::uint64_t>() + rows, 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipMemcpyAsync(cnt, dcnt.get(), 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    const bool dict_ok = (cnt[1] & 1u) == 0 && cnt[0] <= 65536 && cnt[0] > 0;
    // The reference's choice (sector_rows.h:365-369): ids into a dictionary unless the full values
    // are the smaller form; u8 ids up to 256 values.
    int width = 0;
    if (dict_ok) {
        const bool narrow = cnt[0] <= 256;
        const bool full = static_cast<std::uint64_t>(cnt[0]) * 16 + nnz * (narrow ? 1 : 2) >= nnz * 16;
        width = full ? 0 : (narrow ? 1 : 2);
        csr_.id_width_ref = full ? 16 : (narrow ? 1 : 2);
    } else {
        csr_.id_width_ref = 16;
    }
    const std::size_t vbytes = width == 0 ? (real_ ? 8 : 16) : static_cast<std::size_t>(width);
    const std::uint64_t bytes = static_cast<std::uint64_t>(rows + 1) * 8 + nnz * 4 + nnz * vbytes;
    if (nnz > 0xFFFFFFFFFFULL || bytes > budget) {
        row_ptr_.release();
        return;
    }
    mark("walk + scan");
    col_.reset(nnz * 4);
    vals_.reset(nnz * vbytes);
    if (width > 0) {
        std::vector<std::uint64_t> fp(dev::DictTable::kSlots), re(dev::DictTable::kSlots), im(dev::DictTable::kSlots);
        copy_to_host(fp.data(), dfp.get(), dfp.bytes());
        copy_to_host(re.data(), dre.get(), dre.bytes());
        copy_to_host(im.data(), dim_.get(), dim_.bytes());
        std::vector<std::uint16_t> slot_id(dev::DictTable::kSlots, 0);
        std::vector<double> dr, di;
        for (int s = 0; s < dev::DictTable::kSlots; ++s) {
            if (!fp[s]) continue;
            slot_id[s] = static_cast<std::uint16_t>(dr.size());
            double a, b;
            std::memcpy(&a, &re[s], 8);
            std::memcpy(&b, &im[s], 8);
            dr.push_back(a);
            di.push_back(b);
        }
        copy_to_device(dslot.get(), slot_id.data(), slot_id.size() * 2);
        dict_re_ = upload(dr.data(), dr.size());
        dict_im_ = upload(di.data(), di.size());
        csr_.dict = static_cast<std::int64_t>(dr.size());
    }
    mark("csr alloc + dictionary");
    fill(width);
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipMemcpyAsync(cnt, dcnt.get(), 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    if (width > 0 && (cnt[1] & 2u)) {   // a fingerprint collision: store the values in full
        width = 0;
        vals_.reset(nnz * (real_ ? 8 : 16));
        dict_re_.release();
        dict_im_.release();
        csr_.dict = 0;
        fill(0);
        QSA_HIP(hipStreamSynchronize(st));
    }
    // u8 ids with columns below 2^24: the id rides in the column word's top byte (one 4-byte load
    // per entry instead of 4 + 1)
    if (width == 1 && col_dim() <= (std::int64_t{1} << 24)) {
        dev::pack_col_ids(col_.get<std::uint32_t>(), vals_.get<std::uint8_t>(), nnz, st);
        QSA_HIP(hipGetLastError());
        QSA_HIP(hipStreamSynchronize(st));
        vals_.release();
        packed_ = true;
    }
    csr_.nnz = static_cast<std::int64_t>(nnz);
    csr_.width = width == 0 ? (real_ ? 8 : 16) : width;
    csr_.bytes = row_ptr_.bytes() + col_.bytes() + vals_.bytes() + dict_re_.bytes() + dict_im_.bytes();
    const double mean = rows ? static_cast<double>(nnz) / static_cast<double>(rows) : 0.0;
    // About four entries per lane (measured on the N = 32 ring). Block applies use the same width:
    // the lanes set the summation order of a row, so a block apply then equals single applies bit
    // for bit.
    int L = 4;
    while (L < 64 && L < mean / 4.0) L *= 2;
    csr_.lanes = L;
    lanes_block_ = L;
    lane_ = kLaneCsr;
    if (tune_ && rows >= (std::int64_t{1} << 18)) {
        mark("compact + pack");
        build_sell();
        mark("sell");
        autotune();
        mark("autotune");
    }
    csr_.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (phases) {
        std::string line = "[csr build] rows " + std::to_string(rows) + " chunks " + std::to_string(n_chunks);
        double prev = 0.0;
        for (const auto& [what, at] : marks) {
            line += std::string(" | ") + what + " " + std::to_string(at - prev);
            prev = at;
        }
        line += " | total " + std::to_string(csr_.build_seconds) + "\n";
        std::fputs(line.c_str(), stderr);
    }
}

void SectorOperator::build_csr_irrep(std::uint64_t budget) {
    const auto t0 = std::chrono::steady_clock::now();
    const hipStream_t st = engine_stream();
    const std::int64_t rows = dim();
    const dev::SectorView R = rows_->view(), C = cols_->view();
    const dev::IrrepView IR = rows_->irrep_view(), IC = cols_->irrep_view();
    const dev::ProgramView P = prog_->view();
    DeviceBuffer row_len(static_cast<std::size_t>(rows) * 4);
    if (!dev::irrep_rows(R, IR, C, IC, same(), P, row_len.get<std::uint32_t>(), nullptr, nullptr, nullptr, st))
        throw Error(kDeviceUnsupported, "SectorOperator: an irrep dimension the device kernels do not hold");
    QSA_HIP(hipGetLastError());
    row_ptr_.reset(static_cast<std::size_t>(rows + 1) * 8);
    const std::size_t tb = dev::row_scan_temp_bytes(rows);
    DeviceBuffer temp(tb ? tb : 8);
    dev::row_scan(row_len.get<std::uint32_t>(), rows, row_ptr_.get<std::uint64_t>(), temp.get(), tb, st);
    QSA_HIP(hipGetLastError());
    std::uint64_t nnz = 0;
    QSA_HIP(hipMemcpyAsync(&nnz, row_ptr_.get<std::uint64_t>() + rows, 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    const std::uint64_t bytes = static_cast<std::uint64_t>(rows + 1) * 8 + nnz * (4 + sizeof(dev::Zd));
    if (nnz > 0xFFFFFFFFFFULL || bytes > budget) {
        row_ptr_.release();
        return;
    }
    col_.reset(std::max<std::uint64_t>(nnz, 1) * 4);
    vals_.reset(std::max<std::uint64_t>(nnz, 1) * sizeof(dev::Zd));
    dev::irrep_rows(R, IR, C, IC, same(), P, nullptr, row_ptr_.get<std::uint64_t>(), col_.get<std::uint32_t>(),
                    vals_.get<dev::Zd>(), st);
    QSA_HIP(hipGetLastError());
    // merged: a column once per row (the blocks of one target reached by several connections summed
    // in walk order), exact zeros dropped
    dev::csr_merge_rows(rows, row_ptr_.get<std::uint64_t>(), col_.get<std::uint32_t>(), vals_.get<dev::Zd>(),
                        row_len.get<std::uint32_t>(), st);
    QSA_HIP(hipGetLastError());
    DeviceBuffer merged_ptr(static_cast<std::size_t>(rows + 1) * 8);
    dev::row_scan(row_len.get<std::uint32_t>(), rows, merged_ptr.get<std::uint64_t>(), temp.get(), tb, st);
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipMemcpyAsync(&nnz, merged_ptr.get<std::uint64_t>() + rows, 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    DeviceBuffer mcol(std::max<std::uint64_t>(nnz, 1) * 4), mval(std::max<std::uint64_t>(nnz, 1) * sizeof(dev::Zd));
    dev::csr_move_rows(rows, row_ptr_.get<std::uint64_t>(), merged_ptr.get<std::uint64_t>(), col_.get<std::uint32_t>(),
                       vals_.get<dev::Zd>(), mcol.get<std::uint32_t>(), mval.get<dev::Zd>(), st);
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipStreamSynchronize(st));
    row_ptr_ = std::move(merged_ptr);
    col_ = std::move(mcol);
    vals_ = std::move(mval);
    csr_.nnz = static_cast<std::int64_t>(nnz);
    csr_.width = 16;
    csr_.dict = 0;
    csr_.id_width_ref = 16;
    csr_.bytes = row_ptr_.bytes() + col_.bytes() + vals_.bytes();
    const double mean = rows ? static_cast<double>(nnz) / static_cast<double>(rows) : 0.0;
    int L = 4;
    while (L < 64 && L < mean / 4.0) L *= 2;
    csr_.lanes = L;
    lanes_block_ = L;
    lane_ = kLaneCsr;
    if (tune_ && rows >= (std::int64_t{1} << 18)) {
        build_sell();
        autotune();
    }
    csr_.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool SectorOperator::build_sell(bool diag_first) {
    if (lane_ != kLaneCsr || dim() == 0) return false;
    if (sell_ && diag_first == sell_diag_first_) return true;
    if (sell_) {   // rebuilt in the other entry order
        csr_.bytes -= sell_ptr_.bytes() + sell_len_.bytes() + sell_col_.bytes() + sell_vals_.bytes();
        sell_ = false;
        sell_col_.release();
        sell_vals_.release();
    }
    sell_diag_first_ = diag_first;
    const hipStream_t st = engine_stream();
    const std::int64_t rows = dim(), slices = (rows + 63) / 64;
    DeviceBuffer words(static_cast<std::size_t>(slices + 1) * 8), flag(4);
    sell_len_.reset(static_cast<std::size_t>(rows) * 2);
    sell_ptr_.reset(static_cast<std::size_t>(slices + 1) * 8);
    QSA_HIP(hipMemsetAsync(words.get(), 0, words.bytes(), st));
    QSA_HIP(hipMemsetAsync(flag.get(), 0, 4, st));
    dev::sell_lengths(row_ptr_.get<std::uint64_t>(), rows, sell_len_.get<std::uint16_t>(), words.get<std::uint64_t>(),
                      flag.get<std::uint32_t>(), st);
    const std::size_t tb = dev::sell_scan_temp_bytes(slices);
    DeviceBuffer temp(tb ? tb : 8);
    dev::sell_scan(words.get<std::uint64_t>(), slices, sell_ptr_.get<std::uint64_t>(), temp.get(), tb, st);
    QSA_HIP(hipGetLastError());
    std::uint64_t total = 0;
    std::uint32_t too_long = 0;
    QSA_HIP(hipMemcpyAsync(&total, sell_ptr_.get<std::uint64_t>() + slices, 8, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipMemcpyAsync(&too_long, flag.get(), 4, hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    if (too_long) {
        sell_len_.release();
        sell_ptr_.release();
        return false;
    }
    // the value stream beside the column words: u8 / u16 ids, or full values (none when packed)
    const int vb = packed_ ? 0 : (csr_.width == 1 || csr_.width == 2) ? csr_.width : (real_ ? 8 : 16);
    sell_col_.reset(std::max<std::uint64_t>(total, 1) * 4);
    if (vb) sell_vals_.reset(std::max<std::uint64_t>(total, 1) * static_cast<std::size_t>(vb));
    QSA_HIP(hipMemsetAsync(sell_col_.get(), 0, sell_col_.bytes(), st));
    if (vb) QSA_HIP(hipMemsetAsync(sell_vals_.get(), 0, sell_vals_.bytes(), st));
    dev::sell_fill(row_ptr_.get<std::uint64_t>(), col_.get<std::uint32_t>(), vals_.get(), vb, rows,
                   sell_ptr_.get<std::uint64_t>(), sell_col_.get<std::uint32_t>(), sell_vals_.get(),
                   sell_diag_first_ && same(), packed_, st);
    QSA_HIP(hipGetLastError());
    QSA_HIP(hipStreamSynchronize(st));
    sell_ = true;
    unroll_ = 8;
    wg_ = 256;
    csr_.bytes += sell_ptr_.bytes() + sell_len_.bytes() + sell_col_.bytes() + sell_vals_.bytes();
    return true;
}

void SectorOperator::autotune() {
    if (lane_ != kLaneCsr || dim() == 0) return;
    const hipStream_t st = engine_stream();
    const bool cv = !real_;
    const std::size_t elem = cv ? 16 : 8;
    const std::size_t rows = static_cast<std::size_t>(dim()), cols = static_cast<std::size_t>(col_dim());
    DeviceBuffer x(cols * elem), y(rows * elem);
    QSA_HIP(hipMemsetAsync(x.get(), 0, x.bytes(), st));
    hipEvent_t a, b;
    QSA_HIP(hipEventCreate(&a));
    QSA_HIP(hipEventCreate(&b));
    auto time_it = [&](auto&& launch) {
        launch();
        QSA_HIP(hipEventRecord(a, st));
        for (int i = 0; i < 3; ++i) launch();
        QSA_HIP(hipEventRecord(b, st));
        QSA_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        QSA_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    };
    // Only choices that leave every result bit unchanged are timed: the gather depth and the
    // workgroup size. The lanes per row set the summation order of a row, so they stay what the
    // mean row length gives and an apply does not depend on timing noise.
    // On the SELL-64 copy one lane owns a row, so no choice changes a sum.
    float best1 = 1e30f;
    int u1 = unroll_, w1 = wg_;
    const std::vector<int> depths = sell() ? std::vector<int>{4, 8, 16} : std::vector<int>{2, 4, 8};
    const std::vector<int> groups = sell() ? std::vector<int>{256, 512, 1024} : std::vector<int>{512, 1024};
    for (int U : depths)
    for (int WG : groups) {
        unroll_ = U;
        wg_ = WG;
        const float t1 = time_it([&] {
            dev::VecSet v;
            v.k = 1;
            v.in[0] = x.get();
            v.out[0] = y.get();
            dev::csr_spmv(csr_view(false, sell()), real_, v, st);
        });
        if (t1 < best1) {
            best1 = t1;
            u1 = U;
            w1 = WG;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    QSA_HIP(hipGetLastError());
    unroll_ = u1;
    wg_ = w1;
}

dev::CsrView SectorOperator::csr_view(bool block, bool sell) const {
    dev::CsrView A;
    A.rows = dim();
    A.row_ptr = row_ptr_.get<std::uint64_t>();
    A.col = col_.get<std::uint32_t>();
    A.vals = vals_.get();
    if (sell && sell_) {
        A.col = sell_col_.get<std::uint32_t>();
        A.vals = sell_vals_.get();
        A.sell_ptr = sell_ptr_.get<std::uint64_t>();
        A.sell_len = sell_len_.get<std::uint16_t>();
        A.sell_acc = sell_acc_;
    }
    A.width = packed_ ? 3 : (csr_.width == 1 || csr_.width == 2) ? csr_.width : 0;
    A.dict_re = dict_re_.get<double>();
    A.dict_im = dict_im_.get<double>();
    A.lanes = block ? lanes_block_ : csr_.lanes;
    A.values_real = real_;
    A.nt = nt_ ? 1 : 0;
    // (block applies keep the dispatcher's workgroup order: their gathers read whole lines already)
    A.xcd = (xcd_ && !block) ? 1 : 0;
    A.unroll = unroll_;
    A.dict_n = static_cast<int>(csr_.dict);
    A.lds_dict = global_dict_ ? 0 : 1;
    A.wg = wg_;
    return A;
}

void SectorOperator::apply(const void* in, void* out, bool complex_vectors) const {
    const void* ins[1] = {in};
    void* outs[1] = {out};
    apply_multi(ins, outs, 1, complex_vectors);
}

bool SectorOperator::lanczos_first_pass(const void* v, const void* vprev, double beta, void* w, bool complex_vectors,
                                        cplx* alpha) const {
    if (lane_ != kLaneCsr || !sell() || !same()) return false;
    const dev::CsrView A = csr_view(false, true);
    const hipStream_t st = engine_stream();
    DeviceBuffer partial(static_cast<std::size_t>(dev::sell_lanczos_blocks(A)) * 2 * sizeof(double) + sizeof(dev::Zd));
    dev::Zd* a = reinterpret_cast<dev::Zd*>(partial.get<char>() + partial.bytes() - sizeof(dev::Zd));
    if (!dev::sell_lanczos(A, real_ && !complex_vectors, v, vprev, beta, w, partial.get<double>(), a, st)) return false;
    QSA_HIP(hipGetLastError());
    dev::Zd h{};
    QSA_HIP(hipMemcpyAsync(&h, a, sizeof(h), hipMemcpyDeviceToHost, st));
    QSA_HIP(hipStreamSynchronize(st));
    *alpha = cplx(h.re, h.im);
    return true;
}

void SectorOperator::apply_multi(const void* const* ins, void* const* outs, int k, bool complex_vectors) const {
    const bool real_vectors = real_ && !complex_vectors;
    const hipStream_t st = engine_stream();
    for (int done = 0; done < k; done += dev::kMaxVec) {
        dev::VecSet v;
        v.k = std::min(dev::kMaxVec, k - done);
        for (int i = 0; i < v.k; ++i) {
            v.in[i] = ins[done + i];
            v.out[i] = outs[done + i];
        }
        if (lane_ == kLaneCsr) dev::csr_spmv(csr_view(false, sell()), real_vectors, v, st);
        else dev::walk_gather(rows_->view(), cols_->view(), same(), prog_->view(), real_vectors, v, st);
    }
    QSA_HIP(hipGetLastError());
}

void SectorOperator::apply_block(int nv, const void* X, void* Y, bool complex_vectors) const {
    if (lane_ != kLaneCsr) throw Error(kDeviceUnsupported, "apply_block needs the CSR lane");
    if (nv != 1 && nv != 2 && nv != 4 && nv != 8) throw Error(kInvalid, "apply_block: nv must be 1, 2, 4 or 8");
    dev::csr_spmv_block(csr_view(true, sell()), real_ && !complex_vectors, nv, X, Y, engine_stream());
    QSA_HIP(hipGetLastError());
}

void SectorOperator::csr_download(std::vector<std::uint64_t>& row_ptr, std::vector<std::uint32_t>& col,
                                  std::vector<double>& val) const {
    csr_download_rows(0, dim(), row_ptr, col, val);
}

void SectorOperator::csr_download_rows(std::int64_t r0, std::int64_t r1, std::vector<std::uint64_t>& row_ptr,
                                       std::vector<std::uint32_t>& col, std::vector<double>& val) const {
    if (lane_ != kLaneCsr) throw Error(kInvalid, "csr_download: no CSR built");
    if (r0 < 0 || r1 < r0 || r1 > dim()) throw Error(kInvalid, "csr_download: row range out of bounds");
    row_ptr.resize(static_cast<std::size_t>(r1 - r0 + 1));
    copy_to_host(row_ptr.data(), row_ptr_.get<std::uint64_t>() + r0, row_ptr.size() * 8);
    const std::size_t e0 = static_cast<std::size_t>(row_ptr.front());
    const std::size_t nnz = static_cast<std::size_t>(row_ptr.back() - row_ptr.front());
    col.resize(nnz);
    if (nnz) copy_to_host(col.data(), col_.get<std::uint32_t>() + e0, nnz * 4);
    val.assign(2 * nnz, 0.0);
    std::vector<std::uint8_t> packed_ids;
    if (packed_) {
        packed_ids.resize(nnz);
        for (std::size_t e = 0; e < nnz; ++e) {
            packed_ids[e] = static_cast<std::uint8_t>(col[e] >> 24);
            col[e] &= 0xFFFFFFu;
        }
    }
    if (csr_.width == 1 || csr_.width == 2) {
        std::vector<double> dr(static_cast<std::size_t>(csr_.dict)), di(static_cast<std::size_t>(csr_.dict));
        copy_to_host(dr.data(), dict_re_.get(), dr.size() * 8);
        copy_to_host(di.data(), dict_im_.get(), di.size() * 8);
        std::vector<std::uint16_t> ids(nnz);
        if (packed_) {
            for (std::size_t e = 0; e < nnz; ++e) ids[e] = packed_ids[e];
        } else if (csr_.width == 1) {
            std::vector<std::uint8_t> b(nnz);
            if (nnz) copy_to_host(b.data(), vals_.get<std::uint8_t>() + e0, nnz);
            for (std::size_t e = 0; e < nnz; ++e) ids[e] = b[e];
        } else if (nnz) {
            copy_to_host(ids.data(), vals_.get<std::uint16_t>() + e0, nnz * 2);
        }
        for (std::size_t e = 0; e < nnz; ++e) {
            val[2 * e] = dr[ids[e]];
            val[2 * e + 1] = di[ids[e]];
        }
    } else if (real_) {
        std::vector<double> v(nnz);
        if (nnz) copy_to_host(v.data(), vals_.get<double>() + e0, nnz * 8);
        for (std::size_t e = 0; e < nnz; ++e) val[2 * e] = v[e];
    } else if (nnz) {
        copy_to_host(val.data(), vals_.get<double>() + 2 * e0, nnz * 16);
    }
}

}  // namespace qsa

// The operator of one symmetry sector (or between two sectors of one group) on device vectors: the
// reference's RepSectorMatVec / CrossSectorMatVec device lanes (src/engine/internal.h:163-609,
// src/gpu/rep_matvec.cu) rebuilt on HIP. Two lanes:
//   gather -- the row walk on the fly, every apply (no storage beyond the sector);
//   csr    -- the merged reduced CSR built once on the device (u8 / u16 value ids into a
//             dictionary, or full values), then a streaming SpMV per apply.
#pragma once

#include <memory>
#include <type_traits>
#include <vector>

#include "../device/launch.h"
#include "algebra.h"
#include "sector.h"

namespace qsa {

enum Lane : int { kLaneAuto = 0, kLaneGather = 1, kLaneCsr = 2 };

struct CsrInfo {
    std::int64_t nnz = 0;
    int width = 0;             ///< stored value bytes: 1 (u8 id), 2 (u16 id), or 8 / 16 (full values)
    std::int64_t dict = 0;     ///< distinct values (0: no dictionary)
    std::uint64_t bytes = 0;   ///< row_ptr + col + values/ids + dictionary
    double build_seconds = 0.0;
    int lanes = 16;
    std::int64_t id_width_ref = 0;   ///< the value width the reference's rule gives (1, 2, 16)
};

class SectorOperator {
public:
    /// tune: time the SpMV's gather depth and workgroup size on a large CSR and keep the fastest;
    /// the lanes per row follow the mean row length either way, so the summation order -- and the
    /// last bits of every apply -- never depend on timing.
    SectorOperator(std::shared_ptr<const DeviceSector> rows, std::shared_ptr<const DeviceSector> cols,
                   std::shared_ptr<const DeviceProgram> prog, double norm_bound, int lane,
                   std::uint64_t csr_budget_bytes, bool tune = true);

    ~SectorOperator() { fence_before_release(); }
    SectorOperator(const SectorOperator&) = delete;
    SectorOperator& operator=(const SectorOperator&) = delete;

    std::int64_t dim() const { return rows_->dim(); }
    std::int64_t col_dim() const { return cols_->dim(); }
    bool same() const { return rows_ == cols_; }
    bool is_real() const { return real_; }
    int lane() const { return lane_; }
    double norm_bound() const { return norm_bound_; }
    const CsrInfo& csr_info() const { return csr_; }
    const DeviceSector& rows() const { return *rows_; }
    const DeviceSector& cols() const { return *cols_; }

    /// out = O in on device vectors (double when is_real() and not complex_vectors, else complex),
    /// on the engine stream.
    void apply(const void* in, void* out, bool complex_vectors = false) const;
    /// The first pass of a Lanczos step fused into the apply (the SELL layout only): w = H v -
    /// beta v_prev (v_prev null: w = H v) and <v, w> into *alpha. false (nothing done) when the
    /// operator cannot (no SELL layout, not square): the caller applies and forms alpha itself.
    bool lanczos_first_pass(const void* v, const void* vprev, double beta, void* w, bool complex_vectors,
                            cplx* alpha) const;
    /// outs[i] = O ins[i], i < k (k-vector applies equal k single applies bit for bit).
    void apply_multi(const void* const* ins, void* const* outs, int k, bool complex_vectors = false) const;
    /// Block-interleaved vectors: X[row * nv + i] (nv in 1, 2, 4, 8); CSR lane only.
    void apply_block(int nv, const void* X, void* Y, bool complex_vectors = false) const;
    /// The CSR on the host (tests): row_ptr, col, values (re, im interleaved).
    void csr_download(std::vector<std::uint64_t>& row_ptr, std::vector<std::uint32_t>& col,
                      std::vector<double>& val) const;
    /// Rows [r0, r1) of the CSR on the host: row_ptr (absolute, r1 - r0 + 1), col, values.
    void csr_download_rows(std::int64_t r0, std::int64_t r1, std::vector<std::uint64_t>& row_ptr,
                           std::vector<std::uint32_t>& col, std::vector<double>& val) const;
    /// Pick the lanes per row of the SpMV (4..64).
    void set_lanes(int lanes) {
        csr_.lanes = lanes;
        lanes_block_ = lanes;
    }
    /// Time the single-vector SpMV at every gather depth (2, 4, 8) and workgroup size (512, 1024) on
    /// this matrix and keep the fastest (both leave the results bit for bit unchanged); a few tens
    /// of milliseconds on a large sector.
    void autotune();
    int lanes_block() const { return lanes_block_; }
    /// The CSR as the kernels read it (CSR lane only).
    dev::CsrView device_csr() const { return csr_view(false); }
    /// Stream the CSR's column / id arrays with non-temporal loads (they are read once per apply).
    void set_nontemporal(bool on) { nt_ = on; }
    bool nontemporal() const { return nt_; }
    /// Give each XCD one contiguous range of rows (workgroup order remapped).
    void set_xcd_rows(bool on) { xcd_ = on; }
    /// Issue the single-vector SpMV's gathers 4 or 8 entries at a time, or 1 (the sums are unchanged).
    void set_unroll(int u) { unroll_ = u; }
    /// Read u8-id dictionaries from global memory instead of staging them in LDS.
    void set_global_dict(bool on) { global_dict_ = on; }
    /// Threads per workgroup of the single-vector SpMV (256, 512 or 1024).
    void set_workgroup(int wg) { wg_ = wg; }
    bool xcd_rows() const { return xcd_; }

    /// Whether applies run on the SELL-64 copy of the CSR (one lane per row; built for the large
    /// operators that time their SpMV).
    bool sell() const { return sell_ && use_sell_; }
    /// Apply through the SELL-64 copy when one exists (default), or through the CSR.
    void set_use_sell(bool on) { use_sell_ = on; }
    /// Build the SELL-64 copy now (tests), or again in the other entry order; false when a row is
    /// too long for it. diag_first: a row's diagonal entry takes its first slot.
    bool build_sell(bool diag_first = false);
    /// SELL-64: partial sums per row, 4 (the default, combined by a fixed tree) or 1; sets the
    /// summation order.
    void set_sell_sums(int acc) { sell_acc_ = acc == 4 ? 4 : 1; }

private:
    void build_csr(std::uint64_t budget);
    /// The CSR of sectors of an irrep of dimension d > 1: counted and filled representative by
    /// representative, full complex values; a target reached by several connections gives several
    /// entries of one column.
    void build_csr_irrep(std::uint64_t budget);
    /// The CSR as a kernel reads it; with sell, the SELL-64 copy (single and multi-vector applies).
    dev::CsrView csr_view(bool block = false, bool sell = false) const;
    int lanes_block_ = 16;

    std::shared_ptr<const DeviceSector> rows_, cols_;
    std::shared_ptr<const DeviceProgram> prog_;
    double norm_bound_;
    bool real_;
    int lane_ = kLaneGather;
    bool nt_ = false;
    bool xcd_ = true;      // single-vector applies: one contiguous row range per XCD
    bool tune_ = true;
    int unroll_ = 4;       // single-vector applies: gathers issued this many entries at a time
    bool global_dict_ = false;
    bool packed_ = false;   // u8 ids in the column words' top byte
    int wg_ = 512;
    CsrInfo csr_;
    DeviceBuffer row_ptr_, col_, vals_, dict_re_, dict_im_;
    bool sell_ = false, use_sell_ = true, sell_diag_first_ = false;
    int sell_acc_ = 4;   // four partial sums per row: as fast single, and the block kernel's 4 lanes per row
    DeviceBuffer sell_ptr_, sell_len_, sell_col_, sell_vals_;
};

/// H v for the Krylov drivers: a sector operator's apply, with its fused first Lanczos pass
/// (lanczos_run takes it when the operator can).
template <class Scalar> struct OperatorMatvec {
    const SectorOperator& op;
    bool complex_vectors;
    void operator()(const Scalar* in, Scalar* out) const { op.apply(in, out, complex_vectors); }
    bool first_pass(const Scalar* v, const Scalar* vprev, double beta, Scalar* w, Scalar& alpha) const {
        cplx a;
        if (!op.lanczos_first_pass(v, vprev, beta, w, complex_vectors, &a)) return false;
        if constexpr (std::is_same_v<Scalar, double>) alpha = a.real();
        else alpha = a;
        return true;
    }
};

}  // namespace qsa

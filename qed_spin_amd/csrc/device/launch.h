// Host-callable launchers of the device kernels (basis.hip, walk.hip, spmv.hip, blas.hip). All
// take an explicit stream and never synchronise; the caller owns every buffer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "dev_common.h"

namespace qsa::dev {

// ---- basis.hip -------------------------------------------------------------------------------
void orbit_flags(const SectorView& S, u64 total, u64* kept, unsigned long long* orbits, hipStream_t st);
std::size_t scan_words_temp_bytes(u64 n_words);
void scan_words(const u64* kept, u64 n_words, u64* counts, u64* offsets, void* temp, std::size_t temp_bytes,
                hipStream_t st);
void orbit_emit(const SectorView& S, u64 total, const u64* kept, const u64* offsets, u64* reps, double* inv_norms,
                hipStream_t st);
void rank_pack(const u64* kept, const u64* offsets, u64 n_words, RankRec* dir, hipStream_t st);
void rank_mark(const SectorView& S, const u64* reps, i64 n, u64* kept, hipStream_t st);
/// psi (zeroed by the caller: C(N, n_up) entries in ascending state order, or 2^N when full) +=
/// the expansion of the sector vector u; *bad = 1 when an image leaves the Sz sector.
void sector_expand(const SectorView& S, const Zd* u, bool full, Zd* psi, unsigned* bad, hipStream_t st);

// ---- walk.hip --------------------------------------------------------------------------------
/// Up to this many vectors per multi-vector launch.
constexpr int kMaxVec = 8;
struct VecSet {
    const void* in[kMaxVec];
    void* out[kMaxVec];
    int k = 0;
};

/// out_v = O in_v for every vector of the set, row by row on the fly (no CSR): rows of sector R,
/// columns of sector C (same: one sector). real: vectors of double (the operator is real), else
/// interleaved complex.
void walk_gather(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, bool real,
                 const VecSet& v, hipStream_t st);

/// Global value dictionary of a CSR build (open addressing on a 64-bit fingerprint, payload bits).
struct DictTable {
    static constexpr int kSlots = 1 << 17;
    u64* fp = nullptr;
    u64* re = nullptr;
    u64* im = nullptr;
    u16* slot_id = nullptr;
    unsigned* count = nullptr;     // distinct values inserted
    unsigned* flags = nullptr;     // bit 0: overflow, bit 1: fingerprint collision
};

/// The merged CSR's rows [row0, row0 + n_rows), walked once: each row's entries in emission order into
/// a stage (entry e of local row i at stage[e * n_rows + i], room for the program's acting groups per
/// row), then each row sorted by (column, emission index), equal columns summed in emission order and
/// exact zeros dropped, in place (sector_rows.h:147-158), with its length in row_len[i]. Then the
/// dictionary from the staged values (csr_stage_dict) and, after the row scan, the CSR from the stage
/// (csr_compact, row_ptr offset to row0). false (nothing launched) when the program has more groups
/// than a row can hold.
bool csr_walk(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, i64 row0, i64 n_rows,
              u32* row_len, u32* stage_col, Zd* stage_val, hipStream_t st);
void csr_stage_dict(i64 rows, const u32* row_len, const Zd* stage_val, const DictTable& D, hipStream_t st);
void csr_compact(i64 rows, const u64* row_ptr, const u32* stage_col, const Zd* stage_val, u32* col, void* vals,
                 int width, bool real, const DictTable& D, hipStream_t st);
std::size_t row_scan_temp_bytes(i64 rows);
void row_scan(const u32* row_len, i64 rows, u64* row_ptr, void* temp, std::size_t temp_bytes, hipStream_t st);

// ---- irrep.hip -------------------------------------------------------------------------------
/// masks[r * words + w] bit b: element 64 w + b fixes representative r (words = ceil(G / 64)).
void stab_masks(const SectorView& S, int words, u64* masks, hipStream_t st);
/// The operator between two sectors of one irrep of dimension d (2..kMaxIrrepDim), representative by
/// representative (the reference's for_each_cross_block, include/ed/matvec/sector_rows.h:72-106):
/// with val null, row_len[state] = the entries of each of a representative's rows (a target reached
/// several times gives several entries); else the entries into the rows' CSR slots (row_ptr from
/// the scan of row_len). Returns false (nothing launched) for an unsupported d.
bool irrep_rows(const SectorView& R, const IrrepView& IR, const SectorView& C, const IrrepView& IC, bool same,
                const ProgramView& P, u32* row_len, const u64* row_ptr, u32* col, Zd* val, hipStream_t st);
/// Each row's entries sorted by column, equal columns summed in their order, exact zeros dropped, in
/// place; the rows' new lengths into row_len.
void csr_merge_rows(i64 rows, const u64* row_ptr, u32* col, Zd* val, u32* row_len, hipStream_t st);
/// The first new_ptr[r + 1] - new_ptr[r] entries of every row moved from old_ptr[r] to new_ptr[r].
void csr_move_rows(i64 rows, const u64* old_ptr, const u64* new_ptr, const u32* col, const Zd* val, u32* ncol, Zd* nval,
                   hipStream_t st);

// ---- observe.hip -----------------------------------------------------------------------------
/// Observables per launch of obs_sweep, times the pairs: the LDS accumulators (16 KB per
/// one-wavefront workgroup).
constexpr int kObsCells = 1024;
/// Workgroups of obs_sweep for a bra sector of `rows` representatives (partial holds 2 x pairs x
/// nob doubles per workgroup): one wavefront each, at most 8192.
int obs_sweep_blocks(i64 rows);
/// out[p][o0 .. o0 + nob) (complex, row length n_obs) = <bras[p]| O_o |kets[p]> for the observables
/// o0 .. o0 + nob of the merged program P (term_obs: each term's observable; gobs[gobs_begin[g] ..
/// gobs_begin[g + 1]): flip group g's observables ascending; R the bra sector, C the ket sector;
/// same: one sector, the diagonal shortcut). n_pairs x nob <= kObsCells. bras / kets: device arrays
/// of device vector pointers.
void obs_sweep(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, const u32* term_obs,
               const u32* gobs_begin, const u32* gobs, int o0, int nob, int n_obs, int n_pairs,
               const Zd* const* bras, const Zd* const* kets, double* partial, double* out, hipStream_t st);

// ---- spmv.hip --------------------------------------------------------------------------------
struct CsrView {
    i64 rows = 0;
    const u64* row_ptr = nullptr;
    const u32* col = nullptr;
    const void* vals = nullptr;    // u8 / u16 ids, or full values
    int width = 0;                 // 1: u8 ids, 2: u16 ids, 3: u8 ids packed in col's top byte, 0: full values
    const double* dict_re = nullptr;
    const double* dict_im = nullptr;
    int lanes = 16;                // lanes per row: 4, 8, 16, 32 or 64
    bool values_real = false;      // the stored values are real (full values stored as double)
    int nt = 0;                    // stream the column / id arrays with non-temporal loads
    int xcd = 0;                   // contiguous row ranges per XCD (grid padded to a multiple of 8)
    int unroll = 1;                // single-vector SpMV: gathers issued 4 or 8 entries at a time (same sums)
    int dict_n = 0;                // dictionary entries (u8 / u16 ids)
    int lds_dict = 1;              // u8 ids: stage the dictionary in LDS
    int wg = 256;                  // single-vector SpMV: threads per workgroup (256, 512, 1024)
    // SELL-64 layout (sell_ptr non-null): col / vals hold the entries slice by slice in groups of
    // four, entry k of row 64 s + j at sell_ptr[s] + 256 (k / 4) + 4 j + k % 4 (slices padded to
    // whole groups); sell_len[row] is the row's entry count. One lane per row.
    const u64* sell_ptr = nullptr;
    const u16* sell_len = nullptr;
    int sell_acc = 1;              // SELL: partial sums per row (1, or 4 combined by a fixed tree)
};
/// out_v = A in_v (real: double vectors; else complex vectors, with real or complex values).
void csr_spmv(const CsrView& A, bool real, const VecSet& v, hipStream_t st);
/// Block-interleaved multi-vector SpMV: X, Y are rows x nv, entry (row, vector) at row * nv + vector.
void csr_spmv_block(const CsrView& A, bool real, int nv, const void* X, void* Y, hipStream_t st);
/// col[e] |= id[e] << 24 (columns below 2^24): the packed u8-id layout.
void pack_col_ids(u32* col, const u8* id, u64 nnz, hipStream_t st);

/// The SELL-64 copy of a CSR (rows in slices of 64, each slice column-major and padded to its
/// longest row). sell_lengths: sell_len[row] and the slice sizes (entries, 64 x the longest row)
/// into slice_words[0 .. n_slices) with slice_words[n_slices] = 0; *too_long (device u32, zeroed
/// by the caller) becomes nonzero when a row has more than 65535 entries. Then an exclusive scan of
/// slice_words gives sell_ptr (sell_scan), and sell_fill copies columns and value ids / values
/// (value_bytes 0: none, 1, 2, 8 or 16) into the slices.
void sell_lengths(const u64* row_ptr, i64 rows, u16* sell_len, u64* slice_words, u32* too_long, hipStream_t st);
std::size_t sell_scan_temp_bytes(i64 n_slices);
void sell_scan(const u64* slice_words, i64 n_slices, u64* sell_ptr, void* temp, std::size_t temp_bytes,
               hipStream_t st);
void sell_fill(const u64* row_ptr, const u32* col, const void* vals, int value_bytes, i64 rows, const u64* sell_ptr,
               u32* sell_col, void* sell_vals, bool diag_first, bool packed, hipStream_t st);

/// Workgroups of sell_lanczos on A (its partial buffer holds two doubles per workgroup).
unsigned sell_lanczos_blocks(const CsrView& A);
/// The first pass of a Lanczos step in the SELL SpMV's epilogue: w = A v - beta v_prev (v_prev null:
/// w = A v), and <v, w> into alpha (device) through partial (2 sell_lanczos_blocks(A) doubles), summed
/// in a fixed order. false (nothing launched) without the SELL layout, or real vectors on complex
/// values.
bool sell_lanczos(const CsrView& A, bool real, const void* v, const void* vprev, double beta, void* w,
                  double* partial, Zd* alpha, hipStream_t st);

// ---- blas.hip (vectors of T = double or Zd; n entries) -------------------------------------------
/// Partial-sum scratch a reduction needs (device), sized by blas_scratch_bytes; never more than
/// kBlasScratchMaxBytes (1024 blocks x 257 columns x one complex partial).
std::size_t blas_scratch_bytes(std::size_t max_cols);
inline constexpr std::size_t kBlasScratchMaxBytes = std::size_t{1024} * 257 * 2 * sizeof(double);
void fill_zero(bool cplx, void* x, std::size_t n, hipStream_t st);
/// x[i] = standard normal entries (complex: both parts) from a counter hash of (seed, i): test and
/// benchmark vectors, not the FTLM samples.
void fill_gaussian(bool cplx, void* x, std::size_t n, u64 seed, hipStream_t st);
void axpy(bool cplx, Zd a, const void* x, void* y, std::size_t n, hipStream_t st);
void scale(bool cplx, Zd a, void* x, std::size_t n, hipStream_t st);
void axpby(bool cplx, Zd a, const void* x, Zd b, void* y, std::size_t n, hipStream_t st);
/// result (device, one Zd) = x^H y.
void dot(bool cplx, const void* x, const void* y, std::size_t n, void* scratch, Zd* result, hipStream_t st);
/// result (device, one double) = ||x||.
void nrm2(bool cplx, const void* x, std::size_t n, void* scratch, double* result, hipStream_t st);
/// y += a x; result = z^H y.
void axpy_dot(bool cplx, Zd a, const void* x, void* y, const void* z, std::size_t n, void* scratch, Zd* result,
              hipStream_t st);
/// y += a x; result = ||y||.
void axpy_nrm2(bool cplx, Zd a, const void* x, void* y, std::size_t n, void* scratch, double* result,
               hipStream_t st);
/// coeffs[j] = V_j^H w for the m columns of the block V (column j at V + j ld); also ||w||^2 into
/// *wnorm2 when non-null.
void dot_many(bool cplx, const void* V, std::size_t ld, std::size_t m, const void* w, std::size_t n, void* scratch,
              Zd* coeffs, double* wnorm2, hipStream_t st);
/// w += sign sum_j coeffs[j] V_j (coeffs in device memory); ||w||^2 after the update into *wnorm2
/// when non-null.
void axpy_many(bool cplx, const Zd* coeffs, const void* V, std::size_t ld, std::size_t m, void* w, std::size_t n,
               void* scratch, double* wnorm2, hipStream_t st, double sign = 1.0);

/// dot_many / axpy_many over columns given as a device array of m device pointers (the reference's
/// array-of-vectors form, BasicBackend::dot_many / axpy_many). axpy_many_table: w += sum_j c_j V_j.
void dot_many_table(bool cplx, const void* const* table, std::size_t m, const void* w, std::size_t n, void* scratch,
                    Zd* coeffs, hipStream_t st);
void axpy_many_table(bool cplx, const Zd* coeffs, const void* const* table, std::size_t m, void* w, std::size_t n,
                     hipStream_t st);

}  // namespace qsa::dev

namespace qsa::dev {
/// out[:, c] = sum_r V[:, r] S[r + c mm] for c < cols (S real, column-major, device memory).
void block_rotate(bool cplx, const void* V, std::size_t ld, std::size_t mm, const double* S, std::size_t cols,
                  void* out, std::size_t n, hipStream_t st);
}  // namespace qsa::dev

namespace qsa::dev {
/// index_and_projection of n states (device arrays): idx[i], proj[2 i], proj[2 i + 1].
void index_proj_batch(const SectorView& S, const u64* states, i64 n, i64* idx, double* proj, hipStream_t st);
}  // namespace qsa::dev

// ---- batch.hip: many blocks x R vectors per Lanczos step ----------------------------------------
namespace qsa::dev {
/// Rows [row_begin, row_end) of one block (at most 256 rows).
struct Tile {
    u64 row_begin, row_end;
    u32 block, pad;
};
struct BatchView {
    u64 rows = 0;                       // packed rows
    u32 n_blocks = 0, n_tiles = 0;
    int R = 0;                          // vectors per block (<= 64), X[row R + s]
    const Tile* tiles = nullptr;        // n_tiles, blocks contiguous
    const u32* sched = nullptr;         // n_tiles: workgroup w runs tile sched[w] (null: tile w)
    const u32* tile_begin = nullptr;    // n_blocks + 1
    const u64* row_ptr = nullptr;       // packed block-diagonal CSR
    const u32* col = nullptr;
    const Zd* val = nullptr;            // the values in full, or (vid set) ids into the dictionary
    const u16* vid = nullptr;
    const double* dict_re = nullptr;
    const double* dict_im = nullptr;
    u32 dict_n = 0;
    int* step = nullptr;                // the current step (device; advanced by batch_advance)
    int* active = nullptr;              // n_blocks x R: the pair is still running
    int* steps = nullptr;               // n_blocks x R: steps recorded
    const int* cap = nullptr;           // n_blocks: Krylov steps of the block (min(M, dim))
    int* fix = nullptr;                 // n_blocks x R: the local overlap asks for a correction (or null)
};
constexpr int kBatchTileRows = 256;
constexpr int kBatchMaxR = 64;
void batch_spmv(const BatchView& B, const Zd* X, Zd* Y, hipStream_t st);
/// W = H V_cur - beta V_prev with the partials of <V_cur, W> (batch_spmv + batch_vec mode 0 in one
/// kernel, the same arithmetic, bit for bit); false (nothing launched) when 1 < R < 8. sprev / scur:
/// per-pair factors
/// of the stored vectors (V = stored times factor), or null (1).
bool batch_spmv_fused(const BatchView& B, const Zd* Vprev, const Zd* Vcur, Zd* W, const double* beta, double* partial,
                      hipStream_t st, const double* sprev = nullptr, const double* scur = nullptr);
/// mode 0: W -= beta V_prev, <V_cur, W>; 1: W -= alpha V_cur, <V_cur, W>; 5: mode 1 and ||W||^2
/// into partial2; 4: W -= alpha V_cur, ||W||^2; 2: local reorthogonalisation, ||W||^2 (with B.fix:
/// only the flagged pairs, into partial2); 3: W *= 1 / beta. Partials per tile and vector. vscale:
/// per-pair factor of the stored V_cur, or null (1).
void batch_vec(const BatchView& B, int mode, const Zd* Vprev, const Zd* Vcur, Zd* W, const double* coef,
               const Zd* zcoef, double threshold, double* partial, hipStream_t st, double* partial2 = nullptr,
               const double* vscale = nullptr);
/// kind 0: alpha; 1: the overlap (and, with B.fix, the flag of the pairs whose |overlap| exceeds
/// threshold from step 1 on); 2: beta (ends breaking-down or capped pairs). The step is read from
/// B.step on the device.
void batch_finish(const BatchView& B, int kind, const double* partial, double* alpha_hist, double* beta_hist, Zd* ov,
                  double* cur_alpha, double* cur_beta, double* inv_beta, const double* breakdown_tol, hipStream_t st,
                  double threshold = 0.0);
/// V scaled to unit norm per (block, vector); inv = 1 / norm, norm2 (may be null) = norm^2.
void batch_normalize(const BatchView& B, Zd* V, double* partial, double* inv, double* norm2, hipStream_t st);
/// V[(row0 + i) R + s] = src[i] for i < n.
void batch_place(Zd* V, int R, u64 row0, int s, const Zd* src, u64 n, hipStream_t st);
void batch_fill_zero(Zd* x, u64 n, hipStream_t st);

// ---- sample.hip ------------------------------------------------------------------------------
/// Gaussian start vectors: vector p (block p / R, slot p % R) of std::mt19937(seeds[p]) ->
/// std::normal_distribution<double>, entry i = (n_2i, n_2i+1) written to
/// V[(row_begin[p / R] + i) R + p % R] for i < dims[p / R]; `vectors` workgroups. Not normalised.
void sample_gauss(Zd* V, int R, const u64* row_begin, const u64* dims, const u32* seeds, u32 vectors, hipStream_t st);
/// One block's CSR into the packed one: rows and columns offset, values in full into val, or (vid
/// set, the block's CSR holding dictionary ids) the ids offset by dict_off into vid.
void pack_csr(const CsrView& A, u64 row_off, u64 col_off, u64 nnz_off, u64* row_ptr, u32* col, Zd* val, u16* vid,
              u32 dict_off, hipStream_t st);
void batch_advance(const BatchView& B, hipStream_t st);
/// One classical Gram-Schmidt pass of the batch vectors W against kept layers 0 .. K-1 (layer k at
/// basis + k * layer entries, the batch layout), only for the blocks whose full[b] is set and, with
/// on (one flag per pair), only for the pairs flagged there:
/// coef[(b * kmax + k) * R + s] = <V_k, W> per pair (fixed-order sums), then W -= sum_k coef_k V_k.
void batch_gram_schmidt(const BatchView& B, const Zd* basis, u64 layer, int K, int kmax, Zd* W, const int* full,
                        Zd* coef, hipStream_t st, const int* on = nullptr);
/// The DGKS gate of the second pass ("twice is enough"): on[pair] = full[block], the pair runs, and
/// ||W|| after the first pass < kappa ||W|| before it, from the tile partials of ||W||^2 before and
/// after (batch_norm_partial's layout).
void batch_dgks_gate(const BatchView& B, const double* before, const double* after, const int* full, double kappa,
                     int* on, hipStream_t st);
/// partial[tile][s] = the tile's part of ||V_s||^2 (what batch_finish kind 2 sums into beta).
void batch_norm_partial(const BatchView& B, const Zd* V, double* partial, hipStream_t st);
/// Layers [K][rows][R] to [R][K][rows] (each pair's K vectors contiguous, for a library GEMM).
void batch_layers_transpose(const Zd* in, u64 rows, int R, int K, Zd* out, hipStream_t st);

/// One chunk of rows of one (channel, sample) pair of the kept-basis overlaps: its rows start at
/// x_base / y_base (element offsets into the panels; layer k of the pair at base + k * ld).
struct OverlapChunk {
    u64 x_base, y_base;
    u32 rows, pair;
};
/// partial[q] (column-major ma x mb of chunk q's pair, at partial + q * kmax * kmax) = the chunk's
/// X^H Y over its rows (FP64 matrix-core products, 16-row slices in order).
void overlap_chunks(const Zd* X, u64 ldx, const Zd* Y, u64 ldy, const OverlapChunk* chunks, u32 n_chunks,
                    const int* ma, const int* mb, int kmax, Zd* partial, hipStream_t st);
/// omega[p] (column-major ma[p] x mb[p], at omega + p * kmax * kmax) = the sum, in chunk order, of
/// the pair's chunk products partial[q] (q = chunk_begin[p] .. chunk_begin[p + 1], each column-major
/// ma[p] x mb[p] at partial + q * kmax * kmax).
void overlap_reduce(const Zd* partial, const u32* chunk_begin, const int* ma, const int* mb, u32 n_pairs, int kmax,
                    Zd* omega, hipStream_t st);
/// A square sector CSR (rows x rows) as a dense column-major complex matrix, zero-filled first.
void csr_to_dense(const CsrView& A, Zd* out, hipStream_t st);
}  // namespace qsa::dev

// Batched Lanczos over many blocks and R vectors per block (device/batch.hip), and the sampled
// thermodynamics built on it: FTLM of every block of a symmetry walk with all blocks and all
// samples advancing through one packed multi-vector matvec per step (the reference runs them block
// by block, samples in lockstep groups of at most 8: include/ed/thermal/ftlm.h:277-480,
// src/engine/thermal.cpp:161-344).
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "sector_operator.h"
#include "thermo.h"

namespace qsa {

struct BatchedTimings {
    double pack_s = 0.0, seed_s = 0.0, steps_s = 0.0, post_s = 0.0, total_s = 0.0;
    std::uint64_t applies = 0;      ///< block-vector applies (sum over steps of the live pairs)
    std::uint64_t rows = 0, nnz = 0;
    int steps = 0;
    bool graph = false;
    int schedule = 0;               ///< the tile schedule the autotuner chose (BatchedLanczos::set_schedule)
    double tune_s = 0.0;
};

/// The packed engine: blocks (square sector operators on the CSR lane) x R complex vectors.
class BatchedLanczos {
public:
    BatchedLanczos(std::vector<std::shared_ptr<const SectorOperator>> blocks, int R);
    ~BatchedLanczos() { fence_before_release(); }
    BatchedLanczos(const BatchedLanczos&) = delete;
    BatchedLanczos& operator=(const BatchedLanczos&) = delete;
    int n_blocks() const { return static_cast<int>(ops_.size()); }
    int R() const { return R_; }
    std::uint64_t rows() const { return rows_; }
    std::uint64_t row_offset(int b) const { return row_off_[b]; }
    std::uint64_t dim(int b) const { return row_off_[b + 1] - row_off_[b]; }

    /// Start vectors from the host, block b sample s at host[(row_offset(b) + i) R + s] (the batch
    /// layout); normalised on the device.
    void set_start(const cplx* host_interleaved);
    /// Start vectors from the device: clear_start(), then place() each vector (dim(b) entries of
    /// a device vector), then normalize(); norm2[b R + s] gets each vector's squared norm.
    void clear_start();
    void place(int b, int s, const cplx* device_vector);
    void normalize(std::vector<double>* norm2);
    /// Start vectors drawn on the device: pair (b, s) gets the Gaussian vector of the 32-bit
    /// Mersenne Twister seed seeds[b R + s] (std::mt19937 -> std::normal_distribution, as the host
    /// draws it), then every vector is normalised.
    void sample_start(const std::vector<std::uint32_t>& seeds);
    /// M Lanczos steps without a kept basis, with breakdown_tol[b] per block: local = LocalDGKS3
    /// with K = 1 (FTLM), else no reorthogonalisation (the continued fraction). The step is captured
    /// as a hipGraph when use_graph. live (optional, B x R): the pairs to run. Returns
    /// alpha[M][B][R], beta[M][B][R] (beta of step j is beta_{j+1}) and steps[B][R].
    void run(int M, const std::vector<double>& breakdown_tol, bool local, bool use_graph, std::vector<double>& alpha,
             std::vector<double>& beta, std::vector<int>& steps, BatchedTimings& t,
             const std::vector<int>* live = nullptr);
    /// M steps keeping every normalised vector: layer j of basis (device, (M + 1) layers of
    /// rows() x R entries, the batch layout) gets V_j, layer 0 the start vectors already placed and
    /// normalised (sample_start, or clear_start / place / normalize). Every pair runs LocalDGKS3
    /// (K = 1); the blocks flagged in full also take two classical Gram-Schmidt passes against all
    /// their kept layers before beta (the reference's FullCGS2 runs). Eager launches: the number of
    /// layers a pass reads grows each step. Outputs as run().
    void run_kept(int M, const std::vector<double>& breakdown_tol, const std::vector<int>& full, cplx* basis,
                  std::vector<double>& alpha, std::vector<double>& beta, std::vector<int>& steps, BatchedTimings& t,
                  const std::vector<int>* live = nullptr);
    /// Y = H X on batch vectors (device pointers, rows() x R complex each).
    void apply(const cplx* X, cplx* Y) const;
    /// The start vectors (rows() x R, device): written by the caller, then normalize().
    cplx* start_buffer() { return V0_.get<cplx>(); }
    /// Which workgroup runs which tile: 0 in order, 1 grouped by XCD (the tiles of a block on one
    /// XCD, blocks spread over the XCDs by rows, so a block's gathers share that XCD's L2). Every
    /// schedule gives the same results bit for bit: a tile's work and partial sums do not depend on
    /// the workgroup that runs it.
    void set_schedule(int kind);
    int schedule() const { return sched_kind_; }
    /// The schedule of the faster packed matvec, timed over a few applies of each (the engine's
    /// autotuner); returns the chosen kind. Engines of fewer than 4096 or more than 65536 tiles keep
    /// schedule 0.
    int tune_schedule();
    /// The packed layout as the kernels read it (tiles, R); every pair live.
    dev::BatchView packed_view() const {
        dev::BatchView B = view();
        B.active = nullptr;
        return B;
    }

private:
    dev::BatchView view() const;
    std::vector<std::shared_ptr<const SectorOperator>> ops_;
    int R_;
    std::uint64_t rows_ = 0, nnz_ = 0;
    std::vector<std::uint64_t> row_off_;
    std::vector<dev::Tile> tiles_;
    std::vector<std::uint32_t> tile_begin_;
    DeviceBuffer row_ptr_, col_, val_, vid_, dict_re_, dict_im_, tiles_d_, tile_begin_d_, cap_d_, active_d_, steps_d_,
        step_d_;
    std::uint32_t dict_n_ = 0;
    int sched_kind_ = 0;
    DeviceBuffer sched_d_;
    DeviceBuffer V0_, V1_, W_, partial_, partial2_, fix_, scale_, alpha_h_, beta_h_, ov_, cur_a_, cur_b_, inv_b_,
        tol_d_;
    friend struct BatchedAccess;
};

struct BatchedFtlmOptions {
    int samples = 40, krylov = 100;
    std::vector<double> betas;
    std::vector<std::uint64_t> block_seeds;   ///< the base seed of every block
    std::vector<double> breakdown_tol;        ///< per block (the reference: 64 eps s_H)
    bool use_graph = true;
};

/// fn(i) for i < n on up to 16 host threads (dynamic schedule); the first exception is rethrown.
void parallel_blocks(std::size_t n, const std::function<void(std::size_t)>& fn);

/// FTLM curves per block (lnZ, E, V per beta, the block's lowest Ritz value in e_min).
std::vector<Curves> batched_ftlm(const std::vector<std::shared_ptr<const SectorOperator>>& blocks,
                                 const BatchedFtlmOptions& o, BatchedTimings& t);

}  // namespace qsa

// The batched Krylov engine: many symmetry blocks of uneven dimension and R vectors per block
// advance one Lanczos step together.
//
// Layout: the blocks' CSRs are packed into one block-diagonal CSR (rows and columns offset by the
// block's first row; values as 16-bit ids into the blocks' dictionaries laid end to end when every
// block has one and they hold at most 65536 values together, else in full); a batch vector holds R
// entries per row, X[row R + s], so the R vectors of a row are one contiguous line. Rows are cut
// into tiles of at most kBatchTileRows rows
// of one block, one workgroup of 256 threads per tile.
//
// Vector kernels: thread t of a tile is (slot t / R, vector t % R); the 256 / R slots take the
// tile's rows in turn, so consecutive threads touch consecutive entries whatever R is. A
// per-(block, vector) reduction is each thread's sum over its rows, then the slots' sums in slot
// order (LDS), then the block's tile sums by a fixed tree: deterministic, and independent of which
// other blocks share the batch.
//
// SpMV: for R >= 8 one thread per (row, vector), the R threads of a row reading the same column and
// value and gathering one contiguous line of R entries; for R < 8, 16 threads per row (R rounded
// up to a power of two, 16 / that many lanes per (row, vector)) splitting the row's entries, summed
// by a fixed xor butterfly.
//
// The Lanczos scalars alpha, beta, the local-reorthogonalisation overlap, the active flags and the
// step counter live on the device; a step needs no host synchronisation, so it is captured once
// as a hipGraph and replayed.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ Zd cmul(Zd a, Zd b) { return Zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Zd cdot(Zd a, Zd b) { return Zd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

/// The tile this workgroup runs: the schedule's entry (a permutation of the tiles), else its own
/// index. Every tile's work and partial sums are the same whichever workgroup runs it.
__device__ __forceinline__ u32 tile_id(const BatchView& B) { return B.sched ? B.sched[blockIdx.x] : blockIdx.x; }

__device__ __forceinline__ bool pair_live(const BatchView& B, u32 block, int s) {
    return !B.active || B.active[static_cast<u64>(block) * B.R + s] != 0;
}

__device__ __forceinline__ Zd entry_value(const BatchView& B, u64 e) {
    if (B.vid) {
        const u32 id = B.vid[e];
        return Zd{B.dict_re[id], B.dict_im[id]};
    }
    return B.val[e];
}

// One thread per (row, vector), R >= 8.
__global__ __launch_bounds__(kThreads) void k_batch_spmv_wide(BatchView B, const Zd* X, Zd* Y) {
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    if (slot >= slots || !pair_live(B, t.block, s)) return;
    for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
        const u64 e0 = B.row_ptr[r], e1 = B.row_ptr[r + 1];
        double ar = 0.0, ai = 0.0;
        for (u64 e = e0; e < e1; ++e) {
            const u64 c = B.col[e];
            const Zd v = entry_value(B, e);
            const Zd x = X[c * R + s];
            ar += v.re * x.re - v.im * x.im;
            ai += v.re * x.im + v.im * x.re;
        }
        Y[r * R + s] = Zd{ar, ai};
    }
}

// R < 8: RP = R rounded up to a power of two, LR = 16 / RP lanes per (row, vector), 16 threads per
// row; lane l of a vector takes entries e0 + l, e0 + l + LR, ... (four lanes per (row, vector) at
// R = 1 measured the same on the N = 28 continued fractions).
template <int RP, int LR>
__global__ __launch_bounds__(kThreads) void k_batch_spmv_narrow(BatchView B, const Zd* X, Zd* Y) {
    constexpr int TPR = RP * LR;
    constexpr int slots = kThreads / TPR;
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R;
    const int slot = threadIdx.x / TPR, w = threadIdx.x % TPR;
    const int s = w / LR, lane = w % LR;
    const bool live = s < R && pair_live(B, t.block, s);
    // every thread of a row group runs the same rounds (the butterfly needs all of them)
    for (u64 base = t.row_begin; base < t.row_end; base += slots) {
        const u64 r = base + slot;
        double ar = 0.0, ai = 0.0;
        if (live && r < t.row_end) {
            const u64 e0 = B.row_ptr[r], e1 = B.row_ptr[r + 1];
            for (u64 e = e0 + lane; e < e1; e += LR) {
                const u64 c = B.col[e];
                const Zd v = entry_value(B, e);
                const Zd x = X[c * R + s];
                ar += v.re * x.re - v.im * x.im;
                ai += v.re * x.im + v.im * x.re;
            }
        }
#pragma unroll
        for (int off = LR / 2; off > 0; off >>= 1) {
            ar += __shfl_xor(ar, off, LR);
            ai += __shfl_xor(ai, off, LR);
        }
        if (live && r < t.row_end && lane == 0) Y[r * R + s] = Zd{ar, ai};
    }
}

// The slots' sums of each vector in slot order into partial[tile][s] (re, im); with only, just the
// vectors whose flag is set. Every thread of the workgroup calls it.
__device__ __forceinline__ void tile_partial(const BatchView& B, double pr, double pi, double* partial,
                                             const int* only = nullptr) {
    __shared__ double sr[kThreads], si[kThreads];
    const int R = B.R, slots = kThreads / R;
    sr[threadIdx.x] = pr;
    si[threadIdx.x] = pi;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < R &&
        (!only || only[static_cast<u64>(B.tiles[tile_id(B)].block) * R + threadIdx.x])) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < slots; ++q) {
            a += sr[q * R + threadIdx.x];
            b += si[q * R + threadIdx.x];
        }
        partial[(static_cast<u64>(tile_id(B)) * 64 + threadIdx.x) * 2] = a;
        partial[(static_cast<u64>(tile_id(B)) * 64 + threadIdx.x) * 2 + 1] = b;
    }
}

// The SpMV of R >= 8 with the first recurrence pass in its epilogue: W = H V_cur - beta V_prev
// (beta of this step; skipped at step 0) and the partial of <V_cur, W>, the same arithmetic as
// batch_spmv followed by batch_vec mode 0, without writing and re-reading W in between.
//
// With sprev / scur the stored vectors are the unnormalised ones and V = U f with the pair's factor
// f (1 / beta of the step that made it): every read multiplies by f, which rounds exactly as the
// stored product of a separate scaling pass would, so the results are the same bit for bit and the
// pass is saved.
__device__ __forceinline__ Zd scaled(Zd v, double f) { return Zd{v.re * f, v.im * f}; }

__global__ __launch_bounds__(kThreads) void k_batch_spmv_wide_fused(BatchView B, const Zd* Vprev, const Zd* Vcur,
                                                                    Zd* W, const double* beta, double* partial,
                                                                    const double* sprev, const double* scur) {
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    const bool live = slot < slots && pair_live(B, t.block, s);
    double pr = 0.0, pi = 0.0;
    if (live) {
        const int step = *B.step;
        const u64 bs = static_cast<u64>(t.block) * R + s;
        const double b = beta[bs];
        const double fp = sprev ? sprev[bs] : 1.0, fc = scur ? scur[bs] : 1.0;
        for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
            const u64 e0 = B.row_ptr[r], e1 = B.row_ptr[r + 1];
            double ar = 0.0, ai = 0.0;
            for (u64 e = e0; e < e1; ++e) {
                const u64 c = B.col[e];
                const Zd v = entry_value(B, e);
                const Zd x = scaled(Vcur[c * R + s], fc);
                ar += v.re * x.re - v.im * x.im;
                ai += v.re * x.im + v.im * x.re;
            }
            const u64 i = r * R + s;
            Zd w{ar, ai};
            if (step > 0) {
                const Zd p = scaled(Vprev[i], fp);
                w = Zd{w.re - b * p.re, w.im - b * p.im};
            }
            W[i] = w;
            const Zd d = cdot(scaled(Vcur[i], fc), w);
            pr += d.re;
            pi += d.im;
        }
    }
    tile_partial(B, pr, pi, partial);
}

// The same for R = 1 (the continued fractions: one vector per block), THREADS threads per tile.
// Each row's entries go to L lanes (lane l: entries e0 + l, e0 + l + L, ... in order) and the
// lanes' sums meet in an xor butterfly; with L = 8 that is the arithmetic of the sector SpMV that
// batch_spmv uses at R = 1, so W equals its product followed by batch_vec mode 0 bit for bit. The
// partial of <V_cur, W> sums the tile's rows in row order, as mode 0 does with one row per
// thread. Workgroups take the tiles in XCD-contiguous runs (workgroup w: tile (w mod 8) * per +
// w / 8) unless a schedule is given, so each XCD's L2 sees neighbouring rows, as the sector SpMV's
// row ranges do.
template <int L, int THREADS>
__global__ __launch_bounds__(THREADS) void k_batch_spmv_one_fused(BatchView B, const Zd* Vprev, const Zd* Vcur,
                                                                  Zd* W, const double* beta, double* partial,
                                                                  const double* sprev, const double* scur) {
    static_assert(THREADS % L == 0 && THREADS >= kBatchTileRows, "whole row groups; one LDS cell per row");
    __shared__ double sr[kBatchTileRows], si[kBatchTileRows];
    u32 tile;
    if (B.sched) {
        tile = B.sched[blockIdx.x];
    } else {
        const u32 per = gridDim.x >> 3;
        tile = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
        if (tile >= B.n_tiles) return;   // the whole workgroup: the grid is rounded up to 8
    }
    const Tile t = B.tiles[tile];
    constexpr int rows_per_pass = THREADS / L;
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    if (threadIdx.x < kBatchTileRows) {
        sr[threadIdx.x] = 0.0;
        si[threadIdx.x] = 0.0;
    }
    __syncthreads();
    if (pair_live(B, t.block, 0)) {
        const int step = *B.step;
        const double b = beta[t.block];
        const double fp = sprev ? sprev[t.block] : 1.0, fc = scur ? scur[t.block] : 1.0;
        // every thread of a row group runs the same rounds (the butterfly needs all of them)
        for (u64 base = t.row_begin; base < t.row_end; base += rows_per_pass) {
            const u64 r = base + g;
            double ar = 0.0, ai = 0.0;
            if (r < t.row_end) {
                const u64 e0 = B.row_ptr[r], e1 = B.row_ptr[r + 1];
                for (u64 e = e0 + lane; e < e1; e += L) {
                    const u64 c = B.col[e];
                    const Zd v = entry_value(B, e);
                    const Zd x = scaled(Vcur[c], fc);
                    ar += v.re * x.re - v.im * x.im;
                    ai += v.re * x.im + v.im * x.re;
                }
            }
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1) {
                ar += __shfl_xor(ar, off, L);
                ai += __shfl_xor(ai, off, L);
            }
            if (r < t.row_end && lane == 0) {
                Zd w{ar, ai};
                if (step > 0) {
                    const Zd p = scaled(Vprev[r], fp);
                    w = Zd{w.re - b * p.re, w.im - b * p.im};
                }
                W[r] = w;
                const Zd d = cdot(scaled(Vcur[r], fc), w);
                sr[r - t.row_begin] = d.re;
                si[r - t.row_begin] = d.im;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int q = 0; q < kBatchTileRows; ++q) {
            a += sr[q];
            c += si[q];
        }
        partial[static_cast<u64>(tile) * 64 * 2] = a;
        partial[static_cast<u64>(tile) * 64 * 2 + 1] = c;
    }
}

// mode 0: W -= beta V_prev (skipped at step 0), then partial of <V_cur, W> (alpha);
// mode 1: W -= alpha V_cur, then partial of <V_cur, W> (the local overlap);
// mode 5: mode 1, and the partial of ||W||^2 into partial2 as well (what mode 2 sums when it does
//         not correct W);
// mode 4: W -= alpha V_cur, then partial of ||W||^2 (no reorthogonalisation);
// mode 2: W -= ov V_cur when |ov| > threshold (from step 1 on), then partial of ||W||^2; with
//         B.fix (after mode 5) only the pairs flagged there, their norms into partial2;
// mode 3: W *= 1 / beta_next (the next basis vector).
// Pairs no longer running are left alone.
__global__ __launch_bounds__(kThreads) void k_batch_vec(BatchView B, int mode, const Zd* Vprev, const Zd* Vcur, Zd* W,
                                                        const double* coef, const Zd* zcoef, double threshold,
                                                        double* partial, double* partial2, const double* vscale) {
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    const u64 bs = static_cast<u64>(t.block) * R + s;
    if (mode == 2 && B.fix) {
        // nothing to correct in this tile's block: the whole workgroup leaves (the usual case)
        bool any = false;
        for (int q = 0; q < R; ++q) any = any || B.fix[static_cast<u64>(t.block) * R + q] != 0;
        if (!any) return;
    }
    const bool live = slot < slots && pair_live(B, t.block, s) && !(mode == 2 && B.fix && !B.fix[bs]);
    double pr = 0.0, pi = 0.0, pn = 0.0;
    if (live) {
        const int step = *B.step;
        const double fc = vscale ? vscale[bs] : 1.0;   // V_cur = stored V_cur times fc
        for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
            const u64 i = r * R + s;
            Zd w = W[i];
            if (mode == 0) {
                if (step > 0) {
                    const double b = coef[bs];
                    const Zd p = Vprev[i];
                    w = Zd{w.re - b * p.re, w.im - b * p.im};
                    W[i] = w;
                }
                const Zd d = cdot(scaled(Vcur[i], fc), w);
                pr += d.re;
                pi += d.im;
            } else if (mode == 1 || mode == 4 || mode == 5) {
                const double a = coef[bs];
                const Zd v = scaled(Vcur[i], fc);
                w = Zd{w.re - a * v.re, w.im - a * v.im};
                W[i] = w;
                if (mode != 4) {
                    const Zd d = cdot(v, w);
                    pr += d.re;
                    pi += d.im;
                    if (mode == 5) pn += w.re * w.re + w.im * w.im;
                } else {
                    pr += w.re * w.re + w.im * w.im;
                }
            } else if (mode == 2) {
                const Zd ov = zcoef[bs];
                if (step > 0 && sqrt(ov.re * ov.re + ov.im * ov.im) > threshold) {
                    const Zd m = cmul(ov, scaled(Vcur[i], fc));
                    w = Zd{w.re - m.re, w.im - m.im};
                    W[i] = w;
                }
                pr += w.re * w.re + w.im * w.im;
            } else {
                const double inv = coef[bs];
                W[i] = Zd{w.re * inv, w.im * inv};
            }
        }
    }
    if (mode == 5) {
        tile_partial(B, pr, pi, partial);
        __syncthreads();   // the LDS of the first sums is read before the second overwrites it
        tile_partial(B, pn, 0.0, partial2);
    } else if (mode == 2 && B.fix) {
        tile_partial(B, pr, pi, partial2, B.fix);
    } else if (mode != 3) {
        tile_partial(B, pr, pi, partial);
    }
}

// The sum over block b's tiles of partial[tile][s] (re, im), one workgroup of 256 threads per
// (block, vector): thread t takes tiles t, t + 256, ... in order, then a fixed pairwise tree in LDS.
// Thread 0 returns the sums. Every thread of the workgroup calls it.
__device__ __forceinline__ void block_sum(const BatchView& B, int b, int s, const double* partial, double& a,
                                          double& c) {
    __shared__ double ra[kThreads], rc[kThreads];
    double x = 0.0, y = 0.0;
    for (u32 t = B.tile_begin[b] + threadIdx.x; t < B.tile_begin[b + 1]; t += kThreads) {
        x += partial[(static_cast<u64>(t) * 64 + s) * 2];
        y += partial[(static_cast<u64>(t) * 64 + s) * 2 + 1];
    }
    ra[threadIdx.x] = x;
    rc[threadIdx.x] = y;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            ra[threadIdx.x] += ra[threadIdx.x + off];
            rc[threadIdx.x] += rc[threadIdx.x + off];
        }
        __syncthreads();
    }
    a = ra[0];
    c = rc[0];
}

// The DGKS gate of the second Gram-Schmidt pass, one workgroup per (block, vector): the pass is
// repeated when the first one left less than kappa of the norm it found.
__global__ __launch_bounds__(kThreads) void k_batch_dgks_gate(BatchView B, const double* before, const double* after,
                                                              const int* full, double kappa, int* on) {
    const u32 b = blockIdx.x / static_cast<u32>(B.R);
    const int s = static_cast<int>(blockIdx.x % static_cast<u32>(B.R));
    double n0 = 0.0, n1 = 0.0, unused = 0.0;
    block_sum(B, static_cast<int>(b), s, before, n0, unused);
    __syncthreads();   // every thread has read the first sum before the LDS is reused
    block_sum(B, static_cast<int>(b), s, after, n1, unused);
    if (threadIdx.x == 0)
        on[blockIdx.x] = (full[b] && pair_live(B, b, s) && sqrt(n1) < kappa * sqrt(n0)) ? 1 : 0;
}

// One workgroup per (block, vector). kind 0: alpha (real part) -> alpha_hist[step]; kind 1: the
// overlap (complex) -> ov; kind 2: beta = sqrt -> beta_hist[step], 1 / beta -> inv_beta; breakdown
// (beta < tol) or the block's Krylov cap ends the pair's run.
__global__ __launch_bounds__(kThreads) void k_batch_finish(BatchView B, int kind, const double* partial,
                                                           double* alpha_hist, double* beta_hist, Zd* ov,
                                                           double* cur_alpha, double* cur_beta, double* inv_beta,
                                                           const double* breakdown_tol, double threshold) {
    const int R = B.R;
    const u64 bs = blockIdx.x;
    const int b = static_cast<int>(bs / R), s = static_cast<int>(bs % R);
    if (!B.active[bs]) return;   // the whole workgroup
    double a, c;
    block_sum(B, b, s, partial, a, c);
    if (threadIdx.x != 0) return;
    const int step = *B.step;
    const u64 hist = static_cast<u64>(step) * B.n_blocks * R + bs;
    if (kind == 0) {
        alpha_hist[hist] = a;
        cur_alpha[bs] = a;
    } else if (kind == 1) {
        ov[bs] = Zd{a, c};
        if (B.fix) B.fix[bs] = (step > 0 && sqrt(a * a + c * c) > threshold) ? 1 : 0;
    } else {
        const double beta = sqrt(a);
        beta_hist[hist] = beta;
        cur_beta[bs] = beta;
        B.steps[bs] = step + 1;
        if (!(beta >= breakdown_tol[b]) || step + 1 >= B.cap[b]) {
            B.active[bs] = 0;   // the pair's run ends here (breakdown, or the block's Krylov space is full)
            inv_beta[bs] = 0.0;
        } else {
            inv_beta[bs] = 1.0 / beta;
        }
    }
}

// ||V||^2 per pair, then V scaled by 1 / ||V||: the normalisation of the start vectors.
__global__ __launch_bounds__(kThreads) void k_batch_norm_partial(BatchView B, const Zd* V, double* partial) {
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    double p = 0.0;
    if (slot < slots)
        for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
            const Zd v = V[r * R + s];
            p += v.re * v.re + v.im * v.im;
        }
    tile_partial(B, p, 0.0, partial);
}

__global__ __launch_bounds__(kThreads) void k_batch_norm_finish(BatchView B, const double* partial, double* inv,
                                                                double* norm2) {
    const u64 bs = blockIdx.x;
    const int b = static_cast<int>(bs / B.R), s = static_cast<int>(bs % B.R);
    double a, c;
    block_sum(B, b, s, partial, a, c);
    if (threadIdx.x != 0) return;
    inv[bs] = a > 0.0 ? 1.0 / sqrt(a) : 0.0;
    if (norm2) norm2[bs] = a;
}

__global__ __launch_bounds__(kThreads) void k_batch_scale(BatchView B, Zd* V, const double* inv) {
    const Tile t = B.tiles[tile_id(B)];
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    if (slot >= slots) return;
    const double f = inv[static_cast<u64>(t.block) * R + s];
    for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
        const Zd v = V[r * R + s];
        V[r * R + s] = Zd{v.re * f, v.im * f};
    }
}

// V[(row0 + i) R + s] = src[i], i < n: one vector placed into the batch layout.
__global__ void k_batch_place(Zd* V, int R, u64 row0, int s, const Zd* src, u64 n) {
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<u64>(gridDim.x) * blockDim.x)
        V[(row0 + i) * R + s] = src[i];
}

__global__ void k_batch_advance(BatchView B) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *B.step += 1;
}

__global__ void k_fill_zd(Zd* x, u64 n) {
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<u64>(gridDim.x) * blockDim.x)
        x[i] = Zd{0.0, 0.0};
}

// Entry e of a sector CSR: its column and its value (dictionary ids of width 1 / 2 / 3, or in full).
__device__ __forceinline__ void decode_entry(const CsrView& A, u64 e, u32& c, Zd& v) {
    const u32 w = A.col[e];
    c = A.width == 3 ? (w & 0xFFFFFFu) : w;
    if (A.width == 3) {
        const u32 id = w >> 24;
        v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
    } else if (A.width == 1) {
        const u32 id = static_cast<const u8*>(A.vals)[e];
        v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
    } else if (A.width == 2) {
        const u32 id = static_cast<const u16*>(A.vals)[e];
        v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
    } else if (A.values_real) {
        v = Zd{static_cast<const double*>(A.vals)[e], 0.0};
    } else {
        v = static_cast<const Zd*>(A.vals)[e];
    }
}

// A square sector CSR as a dense column-major matrix (out zeroed first): out[c n + r] = H[r, c], one
// thread per row (a merged row holds each column once).
__global__ void k_csr_to_dense(CsrView A, Zd* out) {
    const u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
    const u64 n = static_cast<u64>(A.rows);
    if (r >= n) return;
    for (u64 e = A.row_ptr[r]; e < A.row_ptr[r + 1]; ++e) {
        u32 c;
        Zd v;
        decode_entry(A, e, c, v);
        out[static_cast<u64>(c) * n + r] = v;
    }
}

// Pack one block's CSR into the global one: rows shifted by row_off, columns by col_off, values
// expanded from the dictionary (width 1 / 2 / 3) or copied (width 0; double when real) -- or, with a
// packed dictionary, the ids offset by the block's first entry in it.
__global__ void k_pack_csr(CsrView A, u64 row_off, u64 col_off, u64 nnz_off, u64* row_ptr, u32* col, Zd* val,
                           u16* vid, u32 dict_off) {
    const u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
    if (r > static_cast<u64>(A.rows)) return;
    row_ptr[row_off + r] = nnz_off + A.row_ptr[r];
    if (r == static_cast<u64>(A.rows)) return;
    for (u64 e = A.row_ptr[r]; e < A.row_ptr[r + 1]; ++e) {
        const u32 w = A.col[e];
        col[nnz_off + e] = static_cast<u32>(col_off + (A.width == 3 ? (w & 0xFFFFFFu) : w));
        if (vid) {   // the caller checked that the block has ids and the offset ids fit 16 bits
            const u32 id = A.width == 3 ? (w >> 24)
                           : A.width == 1 ? static_cast<u32>(static_cast<const u8*>(A.vals)[e])
                                          : static_cast<u32>(static_cast<const u16*>(A.vals)[e]);
            vid[nnz_off + e] = static_cast<u16>(dict_off + id);
            continue;
        }
        Zd v;
        if (A.width == 3) {
            const u32 id = w >> 24;
            v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
        } else if (A.width == 1) {
            const u32 id = static_cast<const u8*>(A.vals)[e];
            v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
        } else if (A.width == 2) {
            const u32 id = static_cast<const u16*>(A.vals)[e];
            v = Zd{A.dict_re[id], A.values_real ? 0.0 : A.dict_im[id]};
        } else if (A.values_real) {
            v = Zd{static_cast<const double*>(A.vals)[e], 0.0};
        } else {
            v = static_cast<const Zd*>(A.vals)[e];
        }
        val[nnz_off + e] = v;
    }
}

// Whether block b takes a Gram-Schmidt pass: it is fully reorthogonalised and, with per-pair flags,
// at least one of its pairs is flagged.
__device__ __forceinline__ bool gs_block(const BatchView& B, u32 b, const int* full, const int* on) {
    if (!full[b]) return false;
    if (!on) return true;
    bool any = false;
    for (int q = 0; q < B.R; ++q) any = any || on[static_cast<u64>(b) * B.R + q] != 0;
    return any;
}

// Full Gram-Schmidt against kept basis layers (layer k at basis + k * layer, the batch layout), for
// the blocks flagged in full (and, with on, the pairs flagged there). Coefficients: one workgroup
// per (block, k); thread (slot, s) sums <V_k, W> over the block's rows r = first + slot, first +
// slot + slots, ..., then the slots' sums in slot order: coef[(b * kmax + k) * R + s].
// Deterministic, independent of the other blocks.
__global__ __launch_bounds__(kThreads) void k_batch_gs_coef(BatchView B, const Zd* basis, u64 layer, int K, int kmax,
                                                            const Zd* W, const int* full, const int* on, Zd* coef) {
    const u32 b = blockIdx.x / static_cast<u32>(K);
    const int k = static_cast<int>(blockIdx.x % static_cast<u32>(K));
    if (!gs_block(B, b, full, on)) return;   // the whole workgroup
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    const u64 first = B.tiles[B.tile_begin[b]].row_begin, last = B.tiles[B.tile_begin[b + 1] - 1].row_end;
    double pr = 0.0, pi = 0.0;
    if (slot < slots && pair_live(B, b, s) && (!on || on[static_cast<u64>(b) * R + s])) {
        const Zd* V = basis + static_cast<u64>(k) * layer;
        for (u64 r = first + slot; r < last; r += slots) {
            const u64 i = r * R + s;
            const Zd d = cdot(V[i], W[i]);
            pr += d.re;
            pi += d.im;
        }
    }
    __shared__ double sr[kThreads], si[kThreads];
    sr[threadIdx.x] = pr;
    si[threadIdx.x] = pi;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < R) {
        double a = 0.0, c = 0.0;
        for (int q = 0; q < slots; ++q) {
            a += sr[q * R + threadIdx.x];
            c += si[q * R + threadIdx.x];
        }
        coef[(static_cast<u64>(b) * kmax + k) * R + threadIdx.x] = Zd{a, c};
    }
}

// W -= sum_{k < K} coef_k V_k on the tiles of the flagged blocks, k ascending.
__global__ __launch_bounds__(kThreads) void k_batch_gs_update(BatchView B, const Zd* basis, u64 layer, int K, int kmax,
                                                              Zd* W, const int* full, const int* on, const Zd* coef) {
    const Tile t = B.tiles[tile_id(B)];
    if (!gs_block(B, t.block, full, on)) return;
    const int R = B.R, slots = kThreads / R;
    const int slot = threadIdx.x / R, s = threadIdx.x % R;
    if (slot >= slots || !pair_live(B, t.block, s) || (on && !on[static_cast<u64>(t.block) * R + s])) return;
    const Zd* c = coef + static_cast<u64>(t.block) * kmax * R + s;
    for (u64 r = t.row_begin + slot; r < t.row_end; r += slots) {
        const u64 i = r * R + s;
        Zd w = W[i];
        for (int k = 0; k < K; ++k) {
            const Zd m = cmul(c[static_cast<u64>(k) * R], basis[static_cast<u64>(k) * layer + i]);
            w = Zd{w.re - m.re, w.im - m.im};
        }
        W[i] = w;
    }
}

// Layers [K][rows][R] to [R][K][rows]: each pair's vectors contiguous, layer after layer (the
// column-major panels a library GEMM reads).
__global__ void k_batch_layers_transpose(const Zd* in, u64 rows, int R, int K, Zd* out) {
    const u64 n = rows * static_cast<u64>(R) * static_cast<u64>(K);
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<u64>(gridDim.x) * blockDim.x) {
        const u64 s = i % static_cast<u64>(R), kr = i / static_cast<u64>(R);
        const u64 r = kr % rows, k = kr / rows;
        out[(s * static_cast<u64>(K) + k) * rows + r] = in[i];
    }
}

// Kept-basis overlaps X^H Y per chunk of rows on the FP64 matrix cores. A workgroup owns a 64 x 64
// tile of one chunk's (a, b) product; its four waves own 32 x 32 quarters, each four 16 x 16 MFMA
// tiles with a real and an imaginary accumulator (re += xr yr + xi yi, im += xr yi - xi yr: four
// v_mfma_f64_16x16x4 per tile and four rows). 16-row slices of the tile's 64 X and 64 Y vectors are
// staged in LDS with the vector index fastest. Operand layout of v_mfma_f64_16x16x4 (wave64): lane l
// holds A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator entry v of lane l is
// D[4 v + l / 16][l % 16]. A wave skips its 16 x 16 tiles that lie outside the pair's ma x mb product.
// The grid is one-dimensional: workgroup g runs on XCD g % 8, and the tiles of a chunk are
// consecutive on one XCD, so the chunk's vectors, read by every tile of it, stay in that XCD's L2.
typedef double mf_acc __attribute__((ext_vector_type(4)));
constexpr int kOvTile = 64, kOvRows = 16, kXcds = 8;
__global__ __launch_bounds__(256) void k_overlap_mfma(const Zd* X, u64 ldx, const Zd* Y, u64 ldy,
                                                      const OverlapChunk* chunks, u32 n_chunks, int tiles,
                                                      const int* ma, const int* mb, int kmax, Zd* partial) {
    __shared__ Zd xs[kOvRows][kOvTile + 1], ys[kOvRows][kOvTile + 1];   // padded: the column stores spread over banks
    const u32 g = blockIdx.x, xcd = g % kXcds, slot = g / kXcds;
    const u32 per_chunk = static_cast<u32>(tiles * tiles);
    const u32 q = (slot / per_chunk) * kXcds + xcd;   // the chunk
    if (q >= n_chunks) return;
    const u32 tile = slot % per_chunk;
    const OverlapChunk ch = chunks[q];
    const int a0 = static_cast<int>(tile / static_cast<u32>(tiles)) * kOvTile;
    const int b0 = static_cast<int>(tile % static_cast<u32>(tiles)) * kOvTile;
    const int na = ma[ch.pair], nb = mb[ch.pair];
    if (a0 >= na || b0 >= nb) return;   // the whole workgroup: tile outside this pair's product
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wa = (wave & 1) * 32, wb = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    bool on[2][2];
    mf_acc cr[2][2], ci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            on[i][j] = a0 + wa + 16 * i < na && b0 + wb + 16 * j < nb;   // uniform over the wave
            cr[i][j] = mf_acc{0.0, 0.0, 0.0, 0.0};
            ci[i][j] = mf_acc{0.0, 0.0, 0.0, 0.0};
        }
    for (u32 r0 = 0; r0 < ch.rows; r0 += kOvRows) {
        for (int idx = threadIdx.x; idx < kOvTile * kOvRows; idx += 256) {
            const int k = idx / kOvRows, rr = idx % kOvRows;   // consecutive threads: consecutive rows
            const u32 row = r0 + static_cast<u32>(rr);
            const bool in_row = row < ch.rows;
            xs[rr][k] = (in_row && a0 + k < na) ? X[ch.x_base + static_cast<u64>(a0 + k) * ldx + row] : Zd{0.0, 0.0};
            ys[rr][k] = (in_row && b0 + k < nb) ? Y[ch.y_base + static_cast<u64>(b0 + k) * ldy + row] : Zd{0.0, 0.0};
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kOvRows / 4; ++kk) {
            const int rr = kk * 4 + lk;
            Zd xv[2], yv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) xv[i] = xs[rr][wa + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) yv[j] = ys[rr][wb + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (on[i][j]) {
                        cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[i].re, yv[j].re, cr[i][j], 0, 0, 0);
                        ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[i].re, yv[j].im, ci[i][j], 0, 0, 0);
                    }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (on[i][j]) {
                        cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[i].im, yv[j].im, cr[i][j], 0, 0, 0);
                        ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-xv[i].im, yv[j].re, ci[i][j], 0, 0, 0);
                    }
        }
        __syncthreads();
    }
    Zd* out = partial + static_cast<u64>(q) * kmax * kmax;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int a = a0 + wa + 16 * i + 4 * v + lk, b = b0 + wb + 16 * j + li;
                if (a < na && b < nb) out[static_cast<u64>(b) * na + a] = Zd{cr[i][j][v], ci[i][j][v]};
            }
}

// omega[p](a, b) = the sum of the pair's chunk partials in chunk order; one thread per (pair, a, b).
__global__ void k_overlap_reduce(const Zd* partial, const u32* chunk_begin, const int* ma, const int* mb, u32 n_pairs,
                                 int kmax, Zd* omega) {
    const u64 per = static_cast<u64>(kmax) * kmax;
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n_pairs * per;
         i += static_cast<u64>(gridDim.x) * blockDim.x) {
        const u32 p = static_cast<u32>(i / per);
        const u64 e = i % per;
        if (e >= static_cast<u64>(ma[p]) * mb[p]) continue;
        double re = 0.0, im = 0.0;
        for (u32 q = chunk_begin[p]; q < chunk_begin[p + 1]; ++q) {
            const Zd v = partial[static_cast<u64>(q) * per + e];
            re += v.re;
            im += v.im;
        }
        omega[static_cast<u64>(p) * per + e] = Zd{re, im};
    }
}

unsigned grid_for(u64 n) {
    u64 g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return static_cast<unsigned>(g ? g : 1);
}

}  // namespace

void batch_spmv(const BatchView& B, const Zd* X, Zd* Y, hipStream_t st) {
    if (!B.n_tiles) return;
    if (B.R == 1) {
        // one vector per block: the packed matrix is an ordinary CSR, applied by the sector SpMV
        // (8 lanes per row, gathers four entries at a time, a contiguous row range per XCD); pairs
        // that finished still compute their rows, which nothing reads
        CsrView A;
        A.rows = static_cast<i64>(B.rows);
        A.row_ptr = B.row_ptr;
        A.col = B.col;
        if (B.vid) {
            A.vals = B.vid;
            A.width = 2;
            A.dict_re = B.dict_re;
            A.dict_im = B.dict_im;
            A.dict_n = static_cast<int>(B.dict_n);
        } else {
            A.vals = B.val;
            A.width = 0;
        }
        A.values_real = false;
        A.lanes = 8;
        A.xcd = 1;
        A.unroll = 4;
        VecSet v;
        v.k = 1;
        v.in[0] = X;
        v.out[0] = Y;
        csr_spmv(A, false, v, st);
        return;
    }
    const dim3 grid(B.n_tiles), block(kThreads);
    if (B.R >= 8) hipLaunchKernelGGL(k_batch_spmv_wide, grid, block, 0, st, B, X, Y);
    else if (B.R == 1) hipLaunchKernelGGL((k_batch_spmv_narrow<1, 16>), grid, block, 0, st, B, X, Y);
    else if (B.R == 2) hipLaunchKernelGGL((k_batch_spmv_narrow<2, 8>), grid, block, 0, st, B, X, Y);
    else if (B.R <= 4) hipLaunchKernelGGL((k_batch_spmv_narrow<4, 4>), grid, block, 0, st, B, X, Y);
    else hipLaunchKernelGGL((k_batch_spmv_narrow<8, 2>), grid, block, 0, st, B, X, Y);
}

bool batch_spmv_fused(const BatchView& B, const Zd* Vprev, const Zd* Vcur, Zd* W, const double* beta, double* partial,
                      hipStream_t st, const double* sprev, const double* scur) {
    // 1 < R < 8: a fused form of the narrow kernel (the terms of alpha through LDS to keep mode 0's
    // summation order) made the chain24 T > 0 Lanczos runs slower (3.0 -> 3.7 s); the two kernels stay
    if (B.R > 1 && B.R < 8) return false;
    if (!B.n_tiles) return true;
    if (B.R == 1) {
        // 8 lanes per row (the lane count of the sector SpMV that batch_spmv uses at R = 1), 256
        // threads: on the N = 28 continued fractions 1024 threads (two passes per tile) took 0.99 s
        // for the 200 steps against 0.87 s, and 4 lanes per row 0.89 s
        const u32 grid = B.sched ? B.n_tiles : (B.n_tiles + 7u) / 8u * 8u;
        hipLaunchKernelGGL((k_batch_spmv_one_fused<8, kThreads>), dim3(grid), dim3(kThreads), 0, st, B, Vprev, Vcur, W,
                           beta, partial, sprev, scur);
        return true;
    }
    hipLaunchKernelGGL(k_batch_spmv_wide_fused, dim3(B.n_tiles), dim3(kThreads), 0, st, B, Vprev, Vcur, W, beta,
                       partial, sprev, scur);
    return true;
}

void batch_vec(const BatchView& B, int mode, const Zd* Vprev, const Zd* Vcur, Zd* W, const double* coef,
               const Zd* zcoef, double threshold, double* partial, hipStream_t st, double* partial2,
               const double* vscale) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_vec, dim3(B.n_tiles), dim3(kThreads), 0, st, B, mode, Vprev, Vcur, W, coef, zcoef,
                       threshold, partial, partial2, vscale);
}

void batch_finish(const BatchView& B, int kind, const double* partial, double* alpha_hist, double* beta_hist, Zd* ov,
                  double* cur_alpha, double* cur_beta, double* inv_beta, const double* breakdown_tol, hipStream_t st,
                  double threshold) {
    hipLaunchKernelGGL(k_batch_finish, dim3(B.n_blocks * static_cast<u32>(B.R)), dim3(kThreads), 0, st, B, kind, partial,
                       alpha_hist, beta_hist, ov, cur_alpha, cur_beta, inv_beta, breakdown_tol, threshold);
}

void batch_normalize(const BatchView& B, Zd* V, double* partial, double* inv, double* norm2, hipStream_t st) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_norm_partial, dim3(B.n_tiles), dim3(kThreads), 0, st, B, V, partial);
    hipLaunchKernelGGL(k_batch_norm_finish, dim3(B.n_blocks * static_cast<u32>(B.R)), dim3(kThreads), 0, st, B, partial,
                       inv, norm2);
    hipLaunchKernelGGL(k_batch_scale, dim3(B.n_tiles), dim3(kThreads), 0, st, B, V, inv);
}

void batch_place(Zd* V, int R, u64 row0, int s, const Zd* src, u64 n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_batch_place, dim3(grid_for(n)), dim3(256), 0, st, V, R, row0, s, src, n);
}

void csr_to_dense(const CsrView& A, Zd* out, hipStream_t st) {
    const u64 n = static_cast<u64>(A.rows);
    if (!n) return;
    batch_fill_zero(out, n * n, st);
    hipLaunchKernelGGL(k_csr_to_dense, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, A, out);
}

void batch_gram_schmidt(const BatchView& B, const Zd* basis, u64 layer, int K, int kmax, Zd* W, const int* full,
                        Zd* coef, hipStream_t st, const int* on) {
    if (!B.n_tiles || K < 1) return;
    hipLaunchKernelGGL(k_batch_gs_coef, dim3(B.n_blocks * static_cast<u32>(K)), dim3(kThreads), 0, st, B, basis, layer,
                       K, kmax, W, full, on, coef);
    hipLaunchKernelGGL(k_batch_gs_update, dim3(B.n_tiles), dim3(kThreads), 0, st, B, basis, layer, K, kmax, W, full,
                       on, coef);
}

void batch_dgks_gate(const BatchView& B, const double* before, const double* after, const int* full, double kappa,
                     int* on, hipStream_t st) {
    if (!B.n_blocks) return;
    hipLaunchKernelGGL(k_batch_dgks_gate, dim3(B.n_blocks * static_cast<u32>(B.R)), dim3(kThreads), 0, st, B, before,
                       after, full, kappa, on);
}

void batch_norm_partial(const BatchView& B, const Zd* V, double* partial, hipStream_t st) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_norm_partial, dim3(B.n_tiles), dim3(kThreads), 0, st, B, V, partial);
}

void batch_layers_transpose(const Zd* in, u64 rows, int R, int K, Zd* out, hipStream_t st) {
    const u64 n = rows * static_cast<u64>(R) * static_cast<u64>(K);
    if (!n) return;
    hipLaunchKernelGGL(k_batch_layers_transpose, dim3(grid_for(n)), dim3(256), 0, st, in, rows, R, K, out);
}

void overlap_chunks(const Zd* X, u64 ldx, const Zd* Y, u64 ldy, const OverlapChunk* chunks, u32 n_chunks,
                    const int* ma, const int* mb, int kmax, Zd* partial, hipStream_t st) {
    if (!n_chunks) return;
    const int tiles = (kmax + kOvTile - 1) / kOvTile;
    const u64 groups = (static_cast<u64>(n_chunks) + kXcds - 1) / kXcds * kXcds * static_cast<u64>(tiles * tiles);
    hipLaunchKernelGGL(k_overlap_mfma, dim3(static_cast<unsigned>(groups)), dim3(256), 0, st, X, ldx, Y, ldy, chunks,
                       n_chunks, tiles, ma, mb, kmax, partial);
}

void overlap_reduce(const Zd* partial, const u32* chunk_begin, const int* ma, const int* mb, u32 n_pairs, int kmax,
                    Zd* omega, hipStream_t st) {
    if (!n_pairs) return;
    hipLaunchKernelGGL(k_overlap_reduce, dim3(grid_for(static_cast<u64>(n_pairs) * kmax * kmax)), dim3(256), 0, st,
                       partial, chunk_begin, ma, mb, n_pairs, kmax, omega);
}

void batch_advance(const BatchView& B, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, st, B);
}

void batch_fill_zero(Zd* x, u64 n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_zd, dim3(grid_for(n)), dim3(256), 0, st, x, n);
}

void pack_csr(const CsrView& A, u64 row_off, u64 col_off, u64 nnz_off, u64* row_ptr, u32* col, Zd* val, u16* vid,
              u32 dict_off, hipStream_t st) {
    const u64 n = static_cast<u64>(A.rows) + 1;
    hipLaunchKernelGGL(k_pack_csr, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, A, row_off,
                       col_off, nnz_off, row_ptr, col, val, vid, dict_off);
}

}  // namespace qsa::dev

// The batched Krylov engine: many symmetry blocks of uneven dimension and R vectors per block
// advance one Lanczos step together.
//
// Layout: the blocks' CSRs are packed into one block-diagonal CSR (rows and columns offset by the
// block's first row, values in full); a batch vector holds R entries per row, X[row R + s], so the
// R samples of a row are one contiguous line and a wavefront maps lane s to sample s. Rows are cut
// into tiles of at most kTileRows rows of one block; every kernel runs one workgroup per tile and one
// wavefront per row, so a per-(block, sample) reduction is a lane-wise sum over the tile's rows, the
// four wavefronts' sums in order, and then the tiles of the block in order (deterministic, and
// independent of which other blocks share the batch). The Lanczos scalars alpha, beta and the
// local-reorthogonalisation overlap live on the device; a step needs no host synchronisation, so
// it is captured once as a hipGraph and replayed.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kWavesPerTile = 4;

__device__ __forceinline__ Zd cmul(Zd a, Zd b) { return Zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Zd cdot(Zd a, Zd b) { return Zd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

// Y[r, s] = sum_e A[r, e] X[col[e], s]: one wavefront per row, lane s per sample, entries in order
// (pairs no longer running are skipped when B.active is set).
__global__ __launch_bounds__(256) void k_batch_spmv(BatchView B, const Zd* X, Zd* Y) {
    const Tile t = B.tiles[blockIdx.x];
    const int wv = threadIdx.x >> 6, s = threadIdx.x & 63;
    const int R = B.R;
    const bool live = s < R && (!B.active || B.active[t.block * R + s] != 0);
    for (u64 r = t.row_begin + wv; r < t.row_end; r += kWavesPerTile) {
        const u64 e0 = B.row_ptr[r], e1 = B.row_ptr[r + 1];
        double ar = 0.0, ai = 0.0;
        if (live) {
            for (u64 e = e0; e < e1; ++e) {
                const u64 c = B.col[e];  // u32 column
                const Zd v = B.val[e];
                const Zd x = X[c * R + s];
                ar += v.re * x.re - v.im * x.im;
                ai += v.re * x.im + v.im * x.re;
            }
            Y[r * R + s] = Zd{ar, ai};
        }
    }
}

// Lane-wise sums of the tile's rows into partial[tile][s] (the four wavefronts in order).
__device__ __forceinline__ void tile_partial(double pr, double pi, double* partial) {
    __shared__ double sr[kWavesPerTile][64], si[kWavesPerTile][64];
    const int wv = threadIdx.x >> 6, s = threadIdx.x & 63;
    sr[wv][s] = pr;
    si[wv][s] = pi;
    __syncthreads();
    if (wv == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < kWavesPerTile; ++q) {
            a += sr[q][s];
            b += si[q][s];
        }
        partial[(static_cast<u64>(blockIdx.x) * 64 + s) * 2] = a;
        partial[(static_cast<u64>(blockIdx.x) * 64 + s) * 2 + 1] = b;
    }
}

// mode 0: W -= beta V_prev (beta of this step; skipped at step 0), then partial of <V_cur, W>;
// mode 1: W -= alpha V_cur, then partial of <V_cur, W>;
// mode 2: W -= ov V_cur when |ov| > threshold, then partial of ||W||^2;
// mode 3: W *= 1 / beta_next (the next basis vector).
// Inactive (block, sample) pairs are left alone.
__global__ __launch_bounds__(256) void k_batch_vec(BatchView B, int mode, const Zd* Vprev, const Zd* Vcur, Zd* W,
                                                   const double* coef, const Zd* zcoef, double threshold,
                                                   double* partial) {
    const Tile t = B.tiles[blockIdx.x];
    const int wv = threadIdx.x >> 6, s = threadIdx.x & 63;
    const int R = B.R;
    const bool live = s < R && B.active[t.block * R + s] != 0;
    const u64 bs = static_cast<u64>(t.block) * R + s;
    double pr = 0.0, pi = 0.0;
    if (live) {
        for (u64 r = t.row_begin + wv; r < t.row_end; r += kWavesPerTile) {
            const u64 i = r * R + s;
            Zd w = W[i];
            if (mode == 0) {
                if (*B.step > 0) {
                    const double b = coef[bs];
                    const Zd p = Vprev[i];
                    w = Zd{w.re - b * p.re, w.im - b * p.im};
                    W[i] = w;
                }
                const Zd d = cdot(Vcur[i], w);
                pr += d.re;
                pi += d.im;
            } else if (mode == 1) {
                const double a = coef[bs];
                const Zd v = Vcur[i];
                w = Zd{w.re - a * v.re, w.im - a * v.im};
                W[i] = w;
                const Zd d = cdot(v, w);
                pr += d.re;
                pi += d.im;
            } else if (mode == 2) {
                const Zd ov = zcoef[bs];
                if (*B.step > 0 && sqrt(ov.re * ov.re + ov.im * ov.im) > threshold) {
                    const Zd m = cmul(ov, Vcur[i]);
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
    if (mode != 3) tile_partial(pr, pi, partial);
}

// Per block, lane s: the block's tiles in order. kind 0: alpha (real part) -> alpha_hist[step];
// kind 1: the overlap (complex) -> ov; kind 2: beta = sqrt -> beta_hist[step + 1], 1 / beta ->
// inv_beta, breakdown (beta < tol) or the block's Krylov cap ends the pair's run.
__global__ __launch_bounds__(64) void k_batch_finish(BatchView B, int kind, int step, const double* partial,
                                                     double* alpha_hist, double* beta_hist, Zd* ov, double* cur_alpha,
                                                     double* cur_beta, double* inv_beta, const double* breakdown_tol) {
    const int b = blockIdx.x, s = threadIdx.x;
    const int R = B.R;
    if (s >= R) return;
    const u64 bs = static_cast<u64>(b) * R + s;
    if (!B.active[bs]) return;
    if (step < 0) step = *B.step;   // captured graphs read the step from the device
    double a = 0.0, c = 0.0;
    for (u32 t = B.tile_begin[b]; t < B.tile_begin[b + 1]; ++t) {
        a += partial[(static_cast<u64>(t) * 64 + s) * 2];
        c += partial[(static_cast<u64>(t) * 64 + s) * 2 + 1];
    }
    const u64 hist = static_cast<u64>(step) * B.n_blocks * R + bs;
    if (kind == 0) {
        alpha_hist[hist] = a;
        cur_alpha[bs] = a;
    } else if (kind == 1) {
        ov[bs] = Zd{a, c};
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

// 1 / ||V0|| per pair, then V0 scaled: the normalisation of the start vectors.
__global__ __launch_bounds__(256) void k_batch_norm_partial(BatchView B, const Zd* V, double* partial) {
    const Tile t = B.tiles[blockIdx.x];
    const int wv = threadIdx.x >> 6, s = threadIdx.x & 63;
    double p = 0.0;
    if (s < B.R)
        for (u64 r = t.row_begin + wv; r < t.row_end; r += kWavesPerTile) {
            const Zd v = V[r * B.R + s];
            p += v.re * v.re + v.im * v.im;
        }
    tile_partial(p, 0.0, partial);
}

__global__ __launch_bounds__(64) void k_batch_norm_finish(BatchView B, const double* partial, double* inv) {
    const int b = blockIdx.x, s = threadIdx.x;
    if (s >= B.R) return;
    double a = 0.0;
    for (u32 t = B.tile_begin[b]; t < B.tile_begin[b + 1]; ++t) a += partial[(static_cast<u64>(t) * 64 + s) * 2];
    inv[static_cast<u64>(b) * B.R + s] = a > 0.0 ? 1.0 / sqrt(a) : 0.0;
}

__global__ __launch_bounds__(256) void k_batch_scale(BatchView B, Zd* V, const double* inv) {
    const Tile t = B.tiles[blockIdx.x];
    const int wv = threadIdx.x >> 6, s = threadIdx.x & 63;
    if (s >= B.R) return;
    const double f = inv[static_cast<u64>(t.block) * B.R + s];
    for (u64 r = t.row_begin + wv; r < t.row_end; r += kWavesPerTile) {
        const Zd v = V[r * B.R + s];
        V[r * B.R + s] = Zd{v.re * f, v.im * f};
    }
}

__global__ void k_batch_advance(BatchView B) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *B.step += 1;
}

__global__ void k_fill_zd(Zd* x, u64 n) {
    for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<u64>(gridDim.x) * blockDim.x)
        x[i] = Zd{0.0, 0.0};
}

// Pack one block's CSR into the global one: rows shifted by row_off, columns by col_off, values
// expanded from the dictionary (width 1 / 2) or copied (width 0; double when real).
__global__ void k_pack_csr(CsrView A, u64 row_off, u64 col_off, u64 nnz_off, u64* row_ptr, u32* col, Zd* val) {
    const u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
    if (r > static_cast<u64>(A.rows)) return;
    row_ptr[row_off + r] = nnz_off + A.row_ptr[r];
    if (r == static_cast<u64>(A.rows)) return;
    for (u64 e = A.row_ptr[r]; e < A.row_ptr[r + 1]; ++e) {
        col[nnz_off + e] = static_cast<u32>(col_off + A.col[e]);
        Zd v;
        if (A.width == 1) {
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

}  // namespace

void batch_spmv(const BatchView& B, const Zd* X, Zd* Y, hipStream_t st) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_spmv, dim3(B.n_tiles), dim3(256), 0, st, B, X, Y);
}

void batch_vec(const BatchView& B, int mode, const Zd* Vprev, const Zd* Vcur, Zd* W, const double* coef,
               const Zd* zcoef, double threshold, double* partial, hipStream_t st) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_vec, dim3(B.n_tiles), dim3(256), 0, st, B, mode, Vprev, Vcur, W, coef, zcoef, threshold,
                       partial);
}

void batch_finish(const BatchView& B, int kind, int step, const double* partial, double* alpha_hist,
                  double* beta_hist, Zd* ov, double* cur_alpha, double* cur_beta, double* inv_beta,
                  const double* breakdown_tol, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_finish, dim3(B.n_blocks), dim3(64), 0, st, B, kind, step, partial, alpha_hist,
                       beta_hist, ov, cur_alpha, cur_beta, inv_beta, breakdown_tol);
}

void batch_normalize(const BatchView& B, Zd* V, double* partial, double* inv, hipStream_t st) {
    if (!B.n_tiles) return;
    hipLaunchKernelGGL(k_batch_norm_partial, dim3(B.n_tiles), dim3(256), 0, st, B, V, partial);
    hipLaunchKernelGGL(k_batch_norm_finish, dim3(B.n_blocks), dim3(64), 0, st, B, partial, inv);
    hipLaunchKernelGGL(k_batch_scale, dim3(B.n_tiles), dim3(256), 0, st, B, V, inv);
}

void batch_advance(const BatchView& B, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, st, B);
}

void batch_fill_zero(Zd* x, u64 n, hipStream_t st) {
    if (!n) return;
    u64 g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_zd, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, x, n);
}

void pack_csr(const CsrView& A, u64 row_off, u64 col_off, u64 nnz_off, u64* row_ptr, u32* col, Zd* val,
              hipStream_t st) {
    const u64 n = static_cast<u64>(A.rows) + 1;
    hipLaunchKernelGGL(k_pack_csr, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, A, row_off,
                       col_off, nnz_off, row_ptr, col, val);
}

}  // namespace qsa::dev

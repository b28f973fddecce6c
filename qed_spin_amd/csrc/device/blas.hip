// BLAS-1 of the Krylov drivers on device vectors of double or complex<double>, with fused forms
// (axpy + dot, axpy + norm, dot_many / axpy_many for Gram-Schmidt). Every reduction is
// deterministic: a fixed grid for a given n, a fixed xor-butterfly inside the wavefront, fixed
// slots across wavefronts and blocks, and a final tree over the block partials in a fixed order.
// The result stays in device memory; the host reads it when it needs it.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kWaves = kThreads / 64;

inline unsigned grid_for(std::size_t n) {
    const std::size_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
    return static_cast<unsigned>(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

/// Block sum of (a, b): every thread returns the block total (wave sums, then the waves in order).
__device__ __forceinline__ void block_sum2(double& a, double& b) {
    __shared__ double sa[kWaves], sb[kWaves];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int i = 0; i < kWaves; ++i) {
        a += sa[i];
        b += sb[i];
    }
}

template <bool C> struct V;
template <> struct V<false> {
    static __device__ __forceinline__ Zd load(const void* p, std::size_t i) { return Zd{static_cast<const double*>(p)[i], 0.0}; }
    static __device__ __forceinline__ void store(void* p, std::size_t i, Zd z) { static_cast<double*>(p)[i] = z.re; }
};
template <> struct V<true> {
    static __device__ __forceinline__ Zd load(const void* p, std::size_t i) { return static_cast<const Zd*>(p)[i]; }
    static __device__ __forceinline__ void store(void* p, std::size_t i, Zd z) { static_cast<Zd*>(p)[i] = z; }
};

__device__ __forceinline__ Zd cmul(Zd a, Zd b) { return Zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Zd cadd(Zd a, Zd b) { return Zd{a.re + b.re, a.im + b.im}; }
/// conj(a) b
__device__ __forceinline__ Zd cdot(Zd a, Zd b) { return Zd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

template <bool C> __global__ void k_fill_zero(void* x, std::size_t n) {
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
        V<C>::store(x, i, Zd{0.0, 0.0});
}

// Entry i of a Gaussian test vector: two uniforms from splitmix64 of (seed, i), Box-Muller.
__device__ __forceinline__ u64 mix64(u64 z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
template <bool C> __global__ void k_fill_gauss(void* x, std::size_t n, u64 seed) {
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        const u64 a = mix64(seed ^ (2 * i)), b = mix64(seed ^ (2 * i + 1));
        const double u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
        const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(u1));
        const double t = 6.283185307179586 * u2;
        V<C>::store(x, i, Zd{r * cos(t), r * sin(t)});
    }
}

template <bool C> __global__ void k_axpby(Zd a, const void* x, Zd b, void* y, std::size_t n, int mode) {
    // mode 0: y += a x; 1: y *= a; 2: y = a x + b y
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        if (mode == 1) {
            V<C>::store(y, i, C ? cmul(a, V<C>::load(y, i)) : Zd{a.re * V<C>::load(y, i).re, 0.0});
        } else if (mode == 0) {
            const Zd xv = V<C>::load(x, i), yv = V<C>::load(y, i);
            V<C>::store(y, i, C ? cadd(yv, cmul(a, xv)) : Zd{yv.re + a.re * xv.re, 0.0});
        } else {
            const Zd xv = V<C>::load(x, i), yv = V<C>::load(y, i);
            V<C>::store(y, i, C ? cadd(cmul(a, xv), cmul(b, yv)) : Zd{a.re * xv.re + b.re * yv.re, 0.0});
        }
    }
}

// Reduction kinds: 0 dot z^H y; 1 sum |y|^2; both optionally after y += a x (fuse != 0).
template <bool C, int KIND>
__global__ __launch_bounds__(kThreads) void k_reduce(Zd a, const void* x, void* y, const void* z, std::size_t n,
                                                     int fuse, double* partial) {
    double s0 = 0.0, s1 = 0.0;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        Zd yv = V<C>::load(y, i);
        if (fuse) {
            const Zd xv = V<C>::load(x, i);
            yv = C ? cadd(yv, cmul(a, xv)) : Zd{yv.re + a.re * xv.re, 0.0};
            V<C>::store(y, i, yv);
        }
        if constexpr (KIND == 0) {
            const Zd zv = V<C>::load(z, i);
            if constexpr (C) {
                const Zd d = cdot(zv, yv);
                s0 += d.re;
                s1 += d.im;
            } else {
                s0 += zv.re * yv.re;
            }
        } else {
            s0 += yv.re * yv.re;
            if constexpr (C) s0 += yv.im * yv.im;
        }
    }
    block_sum2(s0, s1);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s0;
        partial[2 * blockIdx.x + 1] = s1;
    }
}

/// Sum of nb partial pairs (stride 2) by one block: thread t takes t, t + 256, ... in order, then
/// the block sum. out_kind 0: Zd, 1: sqrt of the first part (a norm), 2: the first part.
__device__ void finish_pairs(const double* partial, int nb, int out_kind, void* out) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        if (out_kind == 0) *static_cast<Zd*>(out) = Zd{a, b};
        else if (out_kind == 1) *static_cast<double*>(out) = sqrt(a);
        else *static_cast<double*>(out) = a;
    }
}

__global__ __launch_bounds__(kThreads) void k_finish(const double* partial, int nb, int out_kind, void* out) {
    finish_pairs(partial, nb, out_kind, out);
}

// The columns of a Gram-Schmidt block: column j at base + j stride (bytes), or table[j] when a
// table of device pointers is given (the reference's array-of-vectors form).
struct Cols {
    const char* base;
    std::size_t stride;
    const void* const* table;
    __device__ __forceinline__ const void* at(int j) const {
        return table ? table[j] : static_cast<const void*>(base + static_cast<std::size_t>(j) * stride);
    }
};

// dot_many: block b owns a row range; the rows come in tiles of kTile staged in LDS (the w tile),
// wavefront v owns the columns j = v, v + 4, ...: its lanes read 64 consecutive rows of column j
// per load (coalesced), sum, and the wave total accumulates into the column's LDS slot (one owner:
// no race). Partials are written column-major [j][b] so the finish reads them contiguously.
constexpr int kManyCols = 256;   // columns per launch
constexpr int kTile = 1024;      // rows per LDS tile

template <bool C>
__global__ __launch_bounds__(kThreads) void k_dot_many(Cols Vb, int m, const void* w,
                                                       std::size_t n, std::size_t rows_per_block, double* partial,
                                                       int nb, int want_norm) {
    __shared__ Zd sw[kTile];
    __shared__ Zd acc[kManyCols + 1];
    const std::size_t r0 = blockIdx.x * rows_per_block;
    const std::size_t r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int j = threadIdx.x; j <= m; j += blockDim.x) acc[j] = Zd{0.0, 0.0};
    double nrm = 0.0;
    __syncthreads();
    for (std::size_t t0 = r0; t0 < r1; t0 += kTile) {
        const int cnt = static_cast<int>((t0 + kTile < r1) ? kTile : (r1 - t0));
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const Zd v = V<C>::load(w, t0 + i);
            sw[i] = v;
            if (want_norm) nrm += v.re * v.re + (C ? v.im * v.im : 0.0);
        }
        __syncthreads();
        for (int j = wv; j < m; j += kWaves) {
            const void* cj = Vb.at(j);
            double pr = 0.0, pi = 0.0;
#pragma unroll 4
            for (int i = ln; i < cnt; i += 64) {
                const Zd vj = V<C>::load(cj, t0 + i);
                if constexpr (C) {
                    const Zd d = cdot(vj, sw[i]);
                    pr += d.re;
                    pi += d.im;
                } else {
                    pr += vj.re * sw[i].re;
                }
            }
            pr = wave_sum(pr);
            if constexpr (C) pi = wave_sum(pi);
            if (ln == 0) acc[j] = Zd{acc[j].re + pr, acc[j].im + pi};
        }
        __syncthreads();
    }
    if (want_norm) {
        double dummy = 0.0;
        block_sum2(nrm, dummy);
        if (threadIdx.x == 0) acc[m] = Zd{nrm, 0.0};
    }
    __syncthreads();
    for (int j = threadIdx.x; j <= m; j += blockDim.x) {
        partial[(static_cast<std::size_t>(j) * nb + blockIdx.x) * 2] = acc[j].re;
        partial[(static_cast<std::size_t>(j) * nb + blockIdx.x) * 2 + 1] = acc[j].im;
    }
}

// One block per column: the sum of its nb block partials (contiguous) in a fixed order.
__global__ __launch_bounds__(kThreads) void k_dot_many_finish(const double* partial, int nb, int m, Zd* coeffs,
                                                              double* wnorm2) {
    const int j = blockIdx.x;
    if (j < m) finish_pairs(partial + static_cast<std::size_t>(j) * nb * 2, nb, 0, coeffs + j);
    else if (wnorm2) finish_pairs(partial + static_cast<std::size_t>(j) * nb * 2, nb, 2, wnorm2);
}

// w += sign sum_j c_j V_j over grid-strided rows, four columns per step; the partial of ||w||^2 after.
template <bool C>
__global__ __launch_bounds__(kThreads) void k_axpy_many(const Zd* coeffs, Cols Vb, int m,
                                                        void* w, std::size_t n, double* partial, int want_norm,
                                                        double sign) {
    __shared__ Zd sc[kManyCols];
    for (int j = threadIdx.x; j < m; j += blockDim.x) sc[j] = Zd{sign * coeffs[j].re, sign * coeffs[j].im};
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        Zd acc = V<C>::load(w, i);
        int j = 0;
        for (; j + 4 <= m; j += 4) {
            const Zd v0 = V<C>::load(Vb.at(j), i);
            const Zd v1 = V<C>::load(Vb.at(j + 1), i);
            const Zd v2 = V<C>::load(Vb.at(j + 2), i);
            const Zd v3 = V<C>::load(Vb.at(j + 3), i);
            if constexpr (C) {
                acc = cadd(acc, cmul(sc[j], v0));
                acc = cadd(acc, cmul(sc[j + 1], v1));
                acc = cadd(acc, cmul(sc[j + 2], v2));
                acc = cadd(acc, cmul(sc[j + 3], v3));
            } else {
                acc.re += sc[j].re * v0.re;
                acc.re += sc[j + 1].re * v1.re;
                acc.re += sc[j + 2].re * v2.re;
                acc.re += sc[j + 3].re * v3.re;
            }
        }
        for (; j < m; ++j) {
            const Zd vj = V<C>::load(Vb.at(j), i);
            acc = C ? cadd(acc, cmul(sc[j], vj)) : Zd{acc.re + sc[j].re * vj.re, 0.0};
        }
        V<C>::store(w, i, acc);
        if (want_norm) s0 += acc.re * acc.re + (C ? acc.im * acc.im : 0.0);
    }
    if (want_norm) {
        block_sum2(s0, s1);
        if (threadIdx.x == 0) {
            partial[2 * blockIdx.x] = s0;
            partial[2 * blockIdx.x + 1] = 0.0;
        }
    }
}

constexpr int kRotCols = 16;
template <bool C>
__global__ __launch_bounds__(256) void k_block_rotate(const void* Vb, std::size_t ld, int mm, const double* S, int c0,
                                                      int nc, void* out, std::size_t n) {
    const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    Zd acc[kRotCols];
#pragma unroll
    for (int c = 0; c < kRotCols; ++c) acc[c] = Zd{0.0, 0.0};
    for (int r = 0; r < mm; ++r) {
        const Zd v = V<C>::load(Vb, r * ld + i);
#pragma unroll
        for (int c = 0; c < kRotCols; ++c)
            if (c < nc) {
                const double s = S[static_cast<std::size_t>(c0 + c) * mm + r];
                acc[c].re += v.re * s;
                if constexpr (C) acc[c].im += v.im * s;
            }
    }
#pragma unroll
    for (int c = 0; c < kRotCols; ++c)
        if (c < nc) V<C>::store(out, static_cast<std::size_t>(c0 + c) * n + i, acc[c]);
}

}  // namespace

std::size_t blas_scratch_bytes(std::size_t max_cols) {
    (void)max_cols;
    const std::size_t red = 2 * kMaxBlocks * sizeof(double);
    const std::size_t many = static_cast<std::size_t>(kMaxBlocks) * (kManyCols + 1) * 2 * sizeof(double);
    return red > many ? red : many;
}

void fill_zero(bool c, void* x, std::size_t n, hipStream_t st) {
    if (!n) return;
    const unsigned g = grid_for(n);
    if (c) hipLaunchKernelGGL(k_fill_zero<true>, dim3(g), dim3(kThreads), 0, st, x, n);
    else hipLaunchKernelGGL(k_fill_zero<false>, dim3(g), dim3(kThreads), 0, st, x, n);
}

void fill_gaussian(bool c, void* x, std::size_t n, u64 seed, hipStream_t st) {
    if (!n) return;
    const unsigned g = grid_for(n);
    if (c) hipLaunchKernelGGL(k_fill_gauss<true>, dim3(g), dim3(kThreads), 0, st, x, n, seed);
    else hipLaunchKernelGGL(k_fill_gauss<false>, dim3(g), dim3(kThreads), 0, st, x, n, seed);
}

static void launch_axpby(bool c, Zd a, const void* x, Zd b, void* y, std::size_t n, int mode, hipStream_t st) {
    if (!n) return;
    const unsigned g = grid_for(n);
    if (c) hipLaunchKernelGGL(k_axpby<true>, dim3(g), dim3(kThreads), 0, st, a, x, b, y, n, mode);
    else hipLaunchKernelGGL(k_axpby<false>, dim3(g), dim3(kThreads), 0, st, a, x, b, y, n, mode);
}

void axpy(bool c, Zd a, const void* x, void* y, std::size_t n, hipStream_t st) { launch_axpby(c, a, x, Zd{0, 0}, y, n, 0, st); }
void scale(bool c, Zd a, void* x, std::size_t n, hipStream_t st) { launch_axpby(c, a, nullptr, Zd{0, 0}, x, n, 1, st); }
void axpby(bool c, Zd a, const void* x, Zd b, void* y, std::size_t n, hipStream_t st) {
    launch_axpby(c, a, x, b, y, n, 2, st);
}

static void reduce(bool c, int kind, Zd a, const void* x, void* y, const void* z, std::size_t n, int fuse,
                   void* scratch, int out_kind, void* out, hipStream_t st) {
    const unsigned g = grid_for(n);
    double* part = static_cast<double*>(scratch);
    if (c) {
        if (kind == 0) hipLaunchKernelGGL((k_reduce<true, 0>), dim3(g), dim3(kThreads), 0, st, a, x, y, z, n, fuse, part);
        else hipLaunchKernelGGL((k_reduce<true, 1>), dim3(g), dim3(kThreads), 0, st, a, x, y, z, n, fuse, part);
    } else {
        if (kind == 0) hipLaunchKernelGGL((k_reduce<false, 0>), dim3(g), dim3(kThreads), 0, st, a, x, y, z, n, fuse, part);
        else hipLaunchKernelGGL((k_reduce<false, 1>), dim3(g), dim3(kThreads), 0, st, a, x, y, z, n, fuse, part);
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part, static_cast<int>(g), out_kind, out);
}

void dot(bool c, const void* x, const void* y, std::size_t n, void* scratch, Zd* result, hipStream_t st) {
    reduce(c, 0, Zd{0, 0}, nullptr, const_cast<void*>(y), x, n, 0, scratch, 0, result, st);
}
void nrm2(bool c, const void* x, std::size_t n, void* scratch, double* result, hipStream_t st) {
    reduce(c, 1, Zd{0, 0}, nullptr, const_cast<void*>(x), nullptr, n, 0, scratch, 1, result, st);
}
void axpy_dot(bool c, Zd a, const void* x, void* y, const void* z, std::size_t n, void* scratch, Zd* result,
              hipStream_t st) {
    reduce(c, 0, a, x, y, z, n, 1, scratch, 0, result, st);
}
void axpy_nrm2(bool c, Zd a, const void* x, void* y, std::size_t n, void* scratch, double* result, hipStream_t st) {
    reduce(c, 1, a, x, y, nullptr, n, 1, scratch, 1, result, st);
}

static void dot_many_cols(bool c, const void* V, std::size_t ld, const void* const* table, std::size_t m,
                          const void* w, std::size_t n, void* scratch, Zd* coeffs, double* wnorm2, hipStream_t st) {
    double* part = static_cast<double*>(scratch);
    const std::size_t elem = c ? sizeof(Zd) : sizeof(double);
    std::size_t done = 0;
    bool first = true;
    std::size_t nb = (n + kTile - 1) / kTile;
    if (nb > static_cast<std::size_t>(kMaxBlocks)) nb = kMaxBlocks;
    if (nb < 1) nb = 1;
    const std::size_t rpb = (n + nb - 1) / nb;
    do {
        const int mc = static_cast<int>((m - done) < static_cast<std::size_t>(kManyCols) ? (m - done) : kManyCols);
        const Cols Vc{table ? nullptr : static_cast<const char*>(V) + done * ld * elem, ld * elem,
                      table ? table + done : nullptr};
        const int want = (first && wnorm2) ? 1 : 0;
        if (c)
            hipLaunchKernelGGL(k_dot_many<true>, dim3(static_cast<unsigned>(nb)), dim3(kThreads), 0, st, Vc, mc, w,
                               n, rpb, part, static_cast<int>(nb), want);
        else
            hipLaunchKernelGGL(k_dot_many<false>, dim3(static_cast<unsigned>(nb)), dim3(kThreads), 0, st, Vc, mc,
                               w, n, rpb, part, static_cast<int>(nb), want);
        hipLaunchKernelGGL(k_dot_many_finish, dim3(static_cast<unsigned>(mc + want)), dim3(kThreads), 0, st, part,
                           static_cast<int>(nb), mc, coeffs + done, want ? wnorm2 : nullptr);
        done += static_cast<std::size_t>(mc);
        first = false;
    } while (done < m);
}

void dot_many(bool c, const void* V, std::size_t ld, std::size_t m, const void* w, std::size_t n, void* scratch,
              Zd* coeffs, double* wnorm2, hipStream_t st) {
    dot_many_cols(c, V, ld, nullptr, m, w, n, scratch, coeffs, wnorm2, st);
}
void dot_many_table(bool c, const void* const* table, std::size_t m, const void* w, std::size_t n, void* scratch,
                    Zd* coeffs, hipStream_t st) {
    if (m) dot_many_cols(c, nullptr, 0, table, m, w, n, scratch, coeffs, nullptr, st);
}

static void axpy_many_cols(bool c, const Zd* coeffs, const void* V, std::size_t ld, const void* const* table,
                           std::size_t m, void* w, std::size_t n, void* scratch, double* wnorm2, hipStream_t st,
                           double sign) {
    double* part = static_cast<double*>(scratch);
    const std::size_t elem = c ? sizeof(Zd) : sizeof(double);
    const unsigned g = grid_for(n);
    std::size_t done = 0;
    do {
        const int mc = static_cast<int>((m - done) < static_cast<std::size_t>(kManyCols) ? (m - done) : kManyCols);
        const bool last = done + static_cast<std::size_t>(mc) >= m;
        const int want = (last && wnorm2) ? 1 : 0;
        const Cols Vc{table ? nullptr : static_cast<const char*>(V) + done * ld * elem, ld * elem,
                      table ? table + done : nullptr};
        if (c)
            hipLaunchKernelGGL(k_axpy_many<true>, dim3(g), dim3(kThreads), 0, st, coeffs + done, Vc, mc, w, n, part,
                               want, sign);
        else
            hipLaunchKernelGGL(k_axpy_many<false>, dim3(g), dim3(kThreads), 0, st, coeffs + done, Vc, mc, w, n,
                               part, want, sign);
        if (want) hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, st, part, static_cast<int>(g), 2, wnorm2);
        done += static_cast<std::size_t>(mc);
    } while (done < m);
}

void axpy_many(bool c, const Zd* coeffs, const void* V, std::size_t ld, std::size_t m, void* w, std::size_t n,
               void* scratch, double* wnorm2, hipStream_t st, double sign) {
    axpy_many_cols(c, coeffs, V, ld, nullptr, m, w, n, scratch, wnorm2, st, sign);
}
void axpy_many_table(bool c, const Zd* coeffs, const void* const* table, std::size_t m, void* w, std::size_t n,
                     hipStream_t st) {
    if (m) axpy_many_cols(c, coeffs, nullptr, 0, table, m, w, n, nullptr, nullptr, st, 1.0);
}

void block_rotate(bool c, const void* Vm, std::size_t ld, std::size_t mm, const double* S, std::size_t cols, void* out,
                  std::size_t n, hipStream_t st) {
    const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
    for (std::size_t c0 = 0; c0 < cols; c0 += kRotCols) {
        const int nc = static_cast<int>(std::min<std::size_t>(kRotCols, cols - c0));
        if (c)
            hipLaunchKernelGGL(k_block_rotate<true>, grid, block, 0, st, Vm, ld, static_cast<int>(mm), S,
                               static_cast<int>(c0), nc, out, n);
        else
            hipLaunchKernelGGL(k_block_rotate<false>, grid, block, 0, st, Vm, ld, static_cast<int>(mm), S,
                               static_cast<int>(c0), nc, out, n);
    }
}

}  // namespace qsa::dev

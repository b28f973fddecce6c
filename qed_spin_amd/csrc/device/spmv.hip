// Reduced-CSR SpMV for the symmetry-sector operator, real or complex, one to eight vectors.
//
// L lanes of a wavefront share a row: lane l takes entries e0 + l, e0 + l + L, ... (coalesced
// column / id streams), and the L partial sums meet through a fixed xor-butterfly, so the result
// is deterministic and a k-vector apply equals k single applies bit for bit (the contract of the
// reference's csr_spmv, src/gpu/rep_matvec.cu:1003-1036). Values come from a dictionary by u8 or
// u16 id (the reference's reduced_csr.h layout) or are stored in full.
#include <hipcub/hipcub.hpp>

#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

template <int NV> struct Ptrs {
    const void* in[NV];
    void* out[NV];
};

// The dictionary an entry's value id points into: u8 ids read it from LDS (staged at kernel start),
// u16 ids from global memory.
struct Dict {
    const double* re;
    const double* im;
};

template <int WIDTH, bool VREAL>
__device__ __forceinline__ Dict stage_dict(const CsrView& A, double* sre, double* sim) {
    if constexpr (WIDTH == 1 || WIDTH == 3) {
        if (A.lds_dict) {
            for (int i = threadIdx.x; i < A.dict_n; i += blockDim.x) {
                sre[i] = A.dict_re[i];
                if constexpr (!VREAL) sim[i] = A.dict_im[i];
            }
            __syncthreads();
            return Dict{sre, sim};
        }
    }
    return Dict{A.dict_re, A.dict_im};
}

template <int WIDTH, bool VREAL>
__device__ __forceinline__ void value_at(const CsrView& A, const Dict& d, u64 e, double& vr, double& vi) {
    if constexpr (WIDTH == 1) {
        const u32 id = A.nt ? __builtin_nontemporal_load(static_cast<const u8*>(A.vals) + e) : static_cast<const u8*>(A.vals)[e];
        vr = d.re[id];
        if constexpr (!VREAL) vi = d.im[id];
    } else if constexpr (WIDTH == 2) {
        const u32 id = A.nt ? __builtin_nontemporal_load(static_cast<const u16*>(A.vals) + e) : static_cast<const u16*>(A.vals)[e];
        vr = A.dict_re[id];
        if constexpr (!VREAL) vi = A.dict_im[id];
    } else {
        if constexpr (VREAL) {
            vr = static_cast<const double*>(A.vals)[e];
        } else {
            const Zd z = static_cast<const Zd*>(A.vals)[e];
            vr = z.re;
            vi = z.im;
        }
    }
}

// The workgroup's place in the row order. With A.xcd the grid is a multiple of 8 and the
// workgroups the dispatcher sends to one XCD (blockIdx = x mod 8, round robin) take one contiguous
// eighth of the rows, so each XCD's L2 sees the columns of neighbouring rows together.
__device__ __forceinline__ i64 row_block(const CsrView& A) {
    if (!A.xcd) return blockIdx.x;
    const u32 per = gridDim.x >> 3;
    return static_cast<i64>(blockIdx.x & 7u) * per + (blockIdx.x >> 3);
}

inline unsigned block_threads(const CsrView& A) { return (A.wg == 512 || A.wg == 1024) ? A.wg : 256; }

inline unsigned grid_blocks(const CsrView& A, u64 threads) {
    const u64 t = block_threads(A);
    u64 b = (threads + t - 1) / t;
    if (A.xcd) b = (b + 7) / 8 * 8;
    return static_cast<unsigned>(b);
}

// Column and value of entry e. WIDTH 3: the u8 id packed in the top byte of the column word
// (columns below 2^24), one 4-byte load per entry.
template <int WIDTH, bool VREAL>
__device__ __forceinline__ void load_entry(const CsrView& A, const Dict& d, u64 e, u32& c, double& vr, double& vi) {
    if constexpr (WIDTH == 3) {
        const u32 w = A.nt ? __builtin_nontemporal_load(A.col + e) : A.col[e];
        c = w & 0xFFFFFFu;
        const u32 id = w >> 24;
        vr = d.re[id];
        if constexpr (!VREAL) vi = d.im[id];
    } else {
        c = A.nt ? __builtin_nontemporal_load(A.col + e) : A.col[e];
        value_at<WIDTH, VREAL>(A, d, e, vr, vi);
    }
}

template <int WIDTH, bool VREAL, bool REAL, int L, int NV, int U>
__global__ __launch_bounds__(1024) void k_spmv(CsrView A, Ptrs<NV> v) {
    __shared__ double sre[(WIDTH == 1 || WIDTH == 3) ? 256 : 1], sim[(WIDTH == 1 || WIDTH == 3) && !VREAL ? 256 : 1];
    const Dict d = stage_dict<WIDTH, VREAL>(A, sre, sim);
    const i64 gid = row_block(A) * blockDim.x + threadIdx.x;
    const i64 row = gid / L;
    const int lane = static_cast<int>(gid % L);
    double ar[NV], ai[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) ar[k] = ai[k] = 0.0;
    auto accumulate = [&](u32 c, double vr, double vi) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if constexpr (REAL) {
                ar[k] += vr * static_cast<const double*>(v.in[k])[c];
            } else {
                const Zd x = static_cast<const Zd*>(v.in[k])[c];
                ar[k] += vr * x.re - vi * x.im;
                ai[k] += vr * x.im + vi * x.re;
            }
        }
    };
    if (row < A.rows) {
        const u64 e0 = A.row_ptr[row], e1 = A.row_ptr[row + 1];
        u64 e = e0 + lane;
        if constexpr (U > 1) {
            // U of the lane's entries per round: their columns and values first, then the U gathers
            // issued together; the sums are taken in entry order, as in the one-at-a-time loop
            for (; e + static_cast<u64>(U - 1) * L < e1; e += static_cast<u64>(U) * L) {
                u32 c[U];
                double vr[U], vi[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    vi[u] = 0.0;
                    load_entry<WIDTH, VREAL>(A, d, e + u * L, c[u], vr[u], vi[u]);
                }
                if constexpr (REAL && NV == 1) {
                    double x[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) x[u] = static_cast<const double*>(v.in[0])[c[u]];
#pragma unroll
                    for (int u = 0; u < U; ++u) ar[0] += vr[u] * x[u];
                } else {
#pragma unroll
                    for (int u = 0; u < U; ++u) accumulate(c[u], vr[u], vi[u]);
                }
            }
        }
        for (; e < e1; e += L) {
            u32 c;
            double vr, vi = 0.0;
            load_entry<WIDTH, VREAL>(A, d, e, c, vr, vi);
            accumulate(c, vr, vi);
        }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            ar[k] += __shfl_xor(ar[k], off, L);
            if constexpr (!REAL) ai[k] += __shfl_xor(ai[k], off, L);
        }
    }
    if (row < A.rows && lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if constexpr (REAL) static_cast<double*>(v.out[k])[row] = ar[k];
            else static_cast<Zd*>(v.out[k])[row] = Zd{ar[k], ai[k]};
        }
    }
}

// Block-interleaved vectors: X[row * NV + k].
template <int WIDTH, bool VREAL, bool REAL, int L, int NV>
__global__ __launch_bounds__(256) void k_spmv_block(CsrView A, const void* X, void* Y) {
    __shared__ double sre[(WIDTH == 1 || WIDTH == 3) ? 256 : 1], sim[(WIDTH == 1 || WIDTH == 3) && !VREAL ? 256 : 1];
    const Dict d = stage_dict<WIDTH, VREAL>(A, sre, sim);
    const i64 gid = row_block(A) * blockDim.x + threadIdx.x;
    const i64 row = gid / L;
    const int lane = static_cast<int>(gid % L);
    double ar[NV], ai[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) ar[k] = ai[k] = 0.0;
    if (row < A.rows) {
        const u64 e0 = A.row_ptr[row], e1 = A.row_ptr[row + 1];
        for (u64 e = e0 + lane; e < e1; e += L) {
            u32 c32;
            double vr, vi = 0.0;
            load_entry<WIDTH, VREAL>(A, d, e, c32, vr, vi);
            const u64 c = c32;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                if constexpr (REAL) {
                    ar[k] += vr * static_cast<const double*>(X)[c * NV + k];
                } else {
                    const Zd x = static_cast<const Zd*>(X)[c * NV + k];
                    ar[k] += vr * x.re - vi * x.im;
                    ai[k] += vr * x.im + vi * x.re;
                }
            }
        }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            ar[k] += __shfl_xor(ar[k], off, L);
            if constexpr (!REAL) ai[k] += __shfl_xor(ai[k], off, L);
        }
    }
    if (row < A.rows && lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if constexpr (REAL) static_cast<double*>(Y)[row * NV + k] = ar[k];
            else static_cast<Zd*>(Y)[row * NV + k] = Zd{ar[k], ai[k]};
        }
    }
}

// ---- SELL-64 ------------------------------------------------------------------------------------
// Slice s holds rows 64 s .. 64 s + 63 in groups of four entries: entry k of row 64 s + j sits at
// sell_ptr[s] + 256 (k / 4) + 4 j + k % 4, so one 16-byte load gives a lane four consecutive
// entries of its row and the group loads of a wavefront's 64 rows are one contiguous kilobyte. A
// slice is padded to its longest row, rounded up to whole groups; the padding is zero and no entry
// beyond a row's length is used. One lane owns a row and sums its entries in CSR order: the sum
// depends neither on the workgroup size nor on how many gathers are in flight, and a k-vector
// apply equals k single applies bit for bit. U entries per round: their groups first, then the U
// gathers together, then the sums in entry order.
// ACC partial sums per row and vector: entry k goes to sum k mod ACC (k in the row's order), and
// the sums meet in a fixed tree, (s0 + s2) + (s1 + s3) for ACC = 4 -- the order a 4-lane butterfly
// gives, so the block kernel's 4 lanes per row produce the same bits.
__device__ __forceinline__ u64 sell_at(u64 slice, u32 j, u32 k) {
    return slice + static_cast<u64>(k >> 2) * 256 + static_cast<u64>(j) * 4 + (k & 3u);
}

typedef u32 u32x4 __attribute__((ext_vector_type(4)));
typedef u16 u16x4 __attribute__((ext_vector_type(4)));
typedef u8 u8x4 __attribute__((ext_vector_type(4)));

template <class V> __device__ __forceinline__ V load_vec(const CsrView& A, const void* p) {
    return A.nt ? __builtin_nontemporal_load(static_cast<const V*>(p)) : *static_cast<const V*>(p);
}

// The four entries of the group at e (a multiple of four): columns and values.
template <int WIDTH, bool VREAL>
__device__ __forceinline__ void load_quad(const CsrView& A, const Dict& d, u64 e, u32* c, double* vr, double* vi) {
    const u32x4 w = load_vec<u32x4>(A, A.col + e);
    if constexpr (WIDTH == 3) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = w[u] & 0xFFFFFFu;
            const u32 id = w[u] >> 24;
            vr[u] = d.re[id];
            if constexpr (!VREAL) vi[u] = d.im[id];
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = w[u];
        if constexpr (WIDTH == 1) {
            const u8x4 id = load_vec<u8x4>(A, static_cast<const u8*>(A.vals) + e);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                vr[u] = d.re[id[u]];
                if constexpr (!VREAL) vi[u] = d.im[id[u]];
            }
        } else if constexpr (WIDTH == 2) {
            const u16x4 id = load_vec<u16x4>(A, static_cast<const u16*>(A.vals) + e);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                vr[u] = A.dict_re[id[u]];
                if constexpr (!VREAL) vi[u] = A.dict_im[id[u]];
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if constexpr (VREAL) {
                    vr[u] = static_cast<const double*>(A.vals)[e + u];
                } else {
                    const Zd z = static_cast<const Zd*>(A.vals)[e + u];
                    vr[u] = z.re;
                    vi[u] = z.im;
                }
            }
        }
    }
}

template <int ACC> __device__ __forceinline__ double tree(const double* s) {
    if constexpr (ACC == 4) return (s[0] + s[2]) + (s[1] + s[3]);
    else return s[0];
}

/// Row `row` of the SELL layout applied to the NV vectors: ACC partial sums per vector, entries in
/// groups of four, U per round.
template <int WIDTH, bool VREAL, bool REAL, int NV, int U, int ACC>
__device__ __forceinline__ void sell_row_sums(const CsrView& A, const Dict& d, i64 row, const Ptrs<NV>& v,
                                              double (&ar)[NV][ACC], double (&ai)[NV][ACC]) {
    static_assert(U % 4 == 0, "whole groups of four entries per round");
    const u32 n = A.nt ? __builtin_nontemporal_load(A.sell_len + row) : A.sell_len[row];
    const u64 first = A.sell_ptr[row >> 6] + static_cast<u64>(row & 63) * 4;   // group g at first + 256 g
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int j = 0; j < ACC; ++j) ar[k][j] = ai[k][j] = 0.0;
    auto accumulate = [&](int j, u32 c, double vr, double vi) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if constexpr (REAL) {
                ar[k][j] += vr * static_cast<const double*>(v.in[k])[c];
            } else {
                const Zd x = static_cast<const Zd*>(v.in[k])[c];
                ar[k][j] += vr * x.re - vi * x.im;
                ai[k][j] += vr * x.im + vi * x.re;
            }
        }
    };
    u32 k = 0;
    for (; k + U <= n; k += U) {
        u32 c[U];
        double vr[U], vi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) vi[u] = 0.0;
#pragma unroll
        for (int g = 0; g < U / 4; ++g)
            load_quad<WIDTH, VREAL>(A, d, first + static_cast<u64>(k / 4 + g) * 256, c + 4 * g, vr + 4 * g, vi + 4 * g);
        if constexpr (REAL && NV == 1) {
            double x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) x[u] = static_cast<const double*>(v.in[0])[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) ar[0][u % ACC] += vr[u] * x[u];
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) accumulate(u % ACC, c[u], vr[u], vi[u]);
        }
    }
    for (; k < n; k += 4) {   // the last groups, entries below n (k stays a multiple of four)
        u32 c[4];
        double vr[4], vi[4] = {0.0, 0.0, 0.0, 0.0};
        load_quad<WIDTH, VREAL>(A, d, first + static_cast<u64>(k / 4) * 256, c, vr, vi);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + static_cast<u32>(u) < n) accumulate(u % ACC, c[u], vr[u], vi[u]);
    }
}

template <int WIDTH, bool VREAL, bool REAL, int NV, int U, int ACC>
__global__ __launch_bounds__(1024) void k_sell(CsrView A, Ptrs<NV> v) {
    __shared__ double sre[(WIDTH == 1 || WIDTH == 3) ? 256 : 1], sim[(WIDTH == 1 || WIDTH == 3) && !VREAL ? 256 : 1];
    const Dict d = stage_dict<WIDTH, VREAL>(A, sre, sim);
    const i64 row = row_block(A) * blockDim.x + threadIdx.x;
    if (row >= A.rows) return;
    double ar[NV][ACC], ai[NV][ACC];
    sell_row_sums<WIDTH, VREAL, REAL, NV, U, ACC>(A, d, row, v, ar, ai);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        if constexpr (REAL) static_cast<double*>(v.out[q])[row] = tree<ACC>(ar[q]);
        else static_cast<Zd*>(v.out[q])[row] = Zd{tree<ACC>(ar[q]), tree<ACC>(ai[q])};
    }
}

// The Lanczos step's first pass in the SpMV's epilogue: w = H v - beta v_prev (v_prev null at the
// first step) written, and the workgroup's partial of <v, w> (the wavefronts' xor trees, then the
// wavefronts in order) into partial[workgroup] -- the apply followed by axpy_dot, without writing
// and reading w in between. k_lanczos_finish sums the partials in workgroup order.
template <int WIDTH, bool VREAL, bool REAL, int U, int ACC>
__global__ __launch_bounds__(1024) void k_sell_lanczos(CsrView A, Ptrs<1> v, const void* vprev, double beta,
                                                       double* partial) {
    __shared__ double sre[(WIDTH == 1 || WIDTH == 3) ? 256 : 1], sim[(WIDTH == 1 || WIDTH == 3) && !VREAL ? 256 : 1];
    __shared__ double wr[16], wi[16];
    const Dict d = stage_dict<WIDTH, VREAL>(A, sre, sim);
    const i64 row = row_block(A) * blockDim.x + threadIdx.x;
    double pr = 0.0, pi = 0.0;
    if (row < A.rows) {
        double ar[1][ACC], ai[1][ACC];
        sell_row_sums<WIDTH, VREAL, REAL, 1, U, ACC>(A, d, row, v, ar, ai);
        if constexpr (REAL) {
            double w = tree<ACC>(ar[0]);
            if (vprev) w = w - beta * static_cast<const double*>(vprev)[row];
            static_cast<double*>(v.out[0])[row] = w;
            pr = static_cast<const double*>(v.in[0])[row] * w;
        } else {
            Zd w{tree<ACC>(ar[0]), tree<ACC>(ai[0])};
            if (vprev) {
                const Zd p = static_cast<const Zd*>(vprev)[row];
                w = Zd{w.re - beta * p.re, w.im - beta * p.im};
            }
            static_cast<Zd*>(v.out[0])[row] = w;
            const Zd x = static_cast<const Zd*>(v.in[0])[row];   // conj(x) w
            pr = x.re * w.re + x.im * w.im;
            pi = x.re * w.im - x.im * w.re;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        pr += __shfl_xor(pr, off, 64);
        pi += __shfl_xor(pi, off, 64);
    }
    const int wave = static_cast<int>(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        wr[wave] = pr;
        wi[wave] = pi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < static_cast<int>(blockDim.x >> 6); ++q) {
            a += wr[q];
            b += wi[q];
        }
        partial[2 * static_cast<u64>(blockIdx.x)] = a;
        partial[2 * static_cast<u64>(blockIdx.x) + 1] = b;
    }
}

__global__ __launch_bounds__(256) void k_lanczos_finish(const double* partial, int nb, Zd* alpha) {
    __shared__ double sr[256], si[256];
    double a = 0.0, b = 0.0;
    for (int i = static_cast<int>(threadIdx.x); i < nb; i += 256) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    sr[threadIdx.x] = a;
    si[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) {
            sr[threadIdx.x] += sr[threadIdx.x + h];
            si[threadIdx.x] += si[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *alpha = Zd{sr[0], si[0]};
}

// Block-interleaved vectors X[row * NV + k] on the SELL layout. ACC lanes share a row, lane l
// taking entries k = l, l + ACC, ... in order (the single-vector kernel's partial sums), combined
// by the xor butterfly whose order is that kernel's tree: the same bits as single applies.
template <int WIDTH, bool VREAL, bool REAL, int NV, int ACC>
__global__ __launch_bounds__(256) void k_sell_block(CsrView A, const void* X, void* Y) {
    __shared__ double sre[(WIDTH == 1 || WIDTH == 3) ? 256 : 1], sim[(WIDTH == 1 || WIDTH == 3) && !VREAL ? 256 : 1];
    const Dict d = stage_dict<WIDTH, VREAL>(A, sre, sim);
    const i64 gid = row_block(A) * blockDim.x + threadIdx.x;
    const i64 row = gid / ACC;
    const int lane = static_cast<int>(gid % ACC);
    double ar[NV], ai[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) ar[q] = ai[q] = 0.0;
    if (row < A.rows) {
        const u32 n = A.sell_len[row];
        const u64 slice = A.sell_ptr[row >> 6];
        const u32 j = static_cast<u32>(row & 63);
        for (u32 k = static_cast<u32>(lane); k < n; k += ACC) {
            u32 c32;
            double vr, vi = 0.0;
            load_entry<WIDTH, VREAL>(A, d, sell_at(slice, j, k), c32, vr, vi);
            const u64 c = c32;
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                if constexpr (REAL) {
                    ar[q] += vr * static_cast<const double*>(X)[c * NV + q];
                } else {
                    const Zd x = static_cast<const Zd*>(X)[c * NV + q];
                    ar[q] += vr * x.re - vi * x.im;
                    ai[q] += vr * x.im + vi * x.re;
                }
            }
        }
    }
#pragma unroll
    for (int off = ACC / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            ar[q] += __shfl_xor(ar[q], off, ACC);
            if constexpr (!REAL) ai[q] += __shfl_xor(ai[q], off, ACC);
        }
    }
    if (row < A.rows && lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            if constexpr (REAL) static_cast<double*>(Y)[row * NV + q] = ar[q];
            else static_cast<Zd*>(Y)[row * NV + q] = Zd{ar[q], ai[q]};
        }
    }
}

template <int WIDTH, bool VREAL, bool REAL> void launch_sell_set(const CsrView& A, const VecSet& v, hipStream_t st) {
    const dim3 grid(grid_blocks(A, static_cast<u64>(A.rows))), block(block_threads(A));
    int done = 0;
    while (done < v.k) {
        const int rest = v.k - done;
        const int nv = rest >= 8 ? 8 : rest >= 4 ? 4 : rest >= 2 ? 2 : 1;
        auto go = [&](auto tag) {
            constexpr int NV = decltype(tag)::value;
            Ptrs<NV> p;
            for (int k = 0; k < NV; ++k) {
                p.in[k] = v.in[done + k];
                p.out[k] = v.out[done + k];
            }
            if (A.sell_acc == 4) {
                if (NV == 1 && A.unroll >= 16) hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 16, 4>), grid, block, 0, st, A, p);
                else if (NV == 1 && A.unroll >= 8) hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 8, 4>), grid, block, 0, st, A, p);
                else hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 4, 4>), grid, block, 0, st, A, p);
            } else {
                if (NV == 1 && A.unroll >= 16) hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 16, 1>), grid, block, 0, st, A, p);
                else if (NV == 1 && A.unroll >= 8) hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 8, 1>), grid, block, 0, st, A, p);
                else hipLaunchKernelGGL((k_sell<WIDTH, VREAL, REAL, NV, 4, 1>), grid, block, 0, st, A, p);
            }
        };
        if (nv == 8) go(std::integral_constant<int, 8>{});
        else if (nv == 4) go(std::integral_constant<int, 4>{});
        else if (nv == 2) go(std::integral_constant<int, 2>{});
        else go(std::integral_constant<int, 1>{});
        done += nv;
    }
}

template <int WIDTH, bool VREAL, bool REAL, int ACC>
void launch_sell_block_acc(const CsrView& A, int nv, const void* X, void* Y, hipStream_t st) {
    CsrView B = A;
    B.wg = 256;
    const dim3 grid(grid_blocks(B, static_cast<u64>(A.rows) * ACC)), block(256);
    switch (nv) {
    case 1: hipLaunchKernelGGL((k_sell_block<WIDTH, VREAL, REAL, 1, ACC>), grid, block, 0, st, B, X, Y); break;
    case 2: hipLaunchKernelGGL((k_sell_block<WIDTH, VREAL, REAL, 2, ACC>), grid, block, 0, st, B, X, Y); break;
    case 4: hipLaunchKernelGGL((k_sell_block<WIDTH, VREAL, REAL, 4, ACC>), grid, block, 0, st, B, X, Y); break;
    case 8: hipLaunchKernelGGL((k_sell_block<WIDTH, VREAL, REAL, 8, ACC>), grid, block, 0, st, B, X, Y); break;
    default: break;
    }
}
template <int WIDTH, bool VREAL, bool REAL>
void launch_sell_block(const CsrView& A, int nv, const void* X, void* Y, hipStream_t st) {
    if (A.sell_acc == 4) launch_sell_block_acc<WIDTH, VREAL, REAL, 4>(A, nv, X, Y, st);
    else launch_sell_block_acc<WIDTH, VREAL, REAL, 1>(A, nv, X, Y, st);
}

template <int WIDTH, bool VREAL, bool REAL, int L> void launch_set(const CsrView& A, const VecSet& v, hipStream_t st) {
    const u64 threads = static_cast<u64>(A.rows) * L;
    const dim3 grid(grid_blocks(A, threads)), block(block_threads(A));
    int done = 0;
    while (done < v.k) {
        const int rest = v.k - done;
        const int nv = rest >= 8 ? 8 : rest >= 4 ? 4 : rest >= 2 ? 2 : 1;
        auto go = [&](auto tag) {
            constexpr int NV = decltype(tag)::value;
            Ptrs<NV> p;
            for (int k = 0; k < NV; ++k) {
                p.in[k] = v.in[done + k];
                p.out[k] = v.out[done + k];
            }
            if (NV == 1 && A.unroll >= 8) hipLaunchKernelGGL((k_spmv<WIDTH, VREAL, REAL, L, NV, 8>), grid, block, 0, st, A, p);
            else if (NV == 1 && A.unroll >= 4) hipLaunchKernelGGL((k_spmv<WIDTH, VREAL, REAL, L, NV, 4>), grid, block, 0, st, A, p);
            else if (NV == 1 && A.unroll > 1) hipLaunchKernelGGL((k_spmv<WIDTH, VREAL, REAL, L, NV, 2>), grid, block, 0, st, A, p);
            else hipLaunchKernelGGL((k_spmv<WIDTH, VREAL, REAL, L, NV, 1>), grid, block, 0, st, A, p);
        };
        if (nv == 8) go(std::integral_constant<int, 8>{});
        else if (nv == 4) go(std::integral_constant<int, 4>{});
        else if (nv == 2) go(std::integral_constant<int, 2>{});
        else go(std::integral_constant<int, 1>{});
        done += nv;
    }
}

template <int WIDTH, bool VREAL, bool REAL> void pick_lanes(const CsrView& A, const VecSet& v, hipStream_t st) {
    switch (A.lanes) {
    case 4: launch_set<WIDTH, VREAL, REAL, 4>(A, v, st); break;
    case 8: launch_set<WIDTH, VREAL, REAL, 8>(A, v, st); break;
    case 32: launch_set<WIDTH, VREAL, REAL, 32>(A, v, st); break;
    case 64: launch_set<WIDTH, VREAL, REAL, 64>(A, v, st); break;
    default: launch_set<WIDTH, VREAL, REAL, 16>(A, v, st); break;
    }
}

template <int WIDTH, bool VREAL, bool REAL, int L>
void launch_block(const CsrView& A, int nv, const void* X, void* Y, hipStream_t st) {
    CsrView B = A;
    B.wg = 256;
    const u64 threads = static_cast<u64>(A.rows) * L;
    const dim3 grid(grid_blocks(B, threads)), block(256);
    switch (nv) {
    case 1: hipLaunchKernelGGL((k_spmv_block<WIDTH, VREAL, REAL, L, 1>), grid, block, 0, st, A, X, Y); break;
    case 2: hipLaunchKernelGGL((k_spmv_block<WIDTH, VREAL, REAL, L, 2>), grid, block, 0, st, A, X, Y); break;
    case 4: hipLaunchKernelGGL((k_spmv_block<WIDTH, VREAL, REAL, L, 4>), grid, block, 0, st, A, X, Y); break;
    case 8: hipLaunchKernelGGL((k_spmv_block<WIDTH, VREAL, REAL, L, 8>), grid, block, 0, st, A, X, Y); break;
    default: break;
    }
}

template <int WIDTH, bool VREAL, bool REAL> void pick_lanes_block(const CsrView& A, int nv, const void* X, void* Y, hipStream_t st) {
    switch (A.lanes) {
    case 4: launch_block<WIDTH, VREAL, REAL, 4>(A, nv, X, Y, st); break;
    case 8: launch_block<WIDTH, VREAL, REAL, 8>(A, nv, X, Y, st); break;
    case 32: launch_block<WIDTH, VREAL, REAL, 32>(A, nv, X, Y, st); break;
    case 64: launch_block<WIDTH, VREAL, REAL, 64>(A, nv, X, Y, st); break;
    default: launch_block<WIDTH, VREAL, REAL, 16>(A, nv, X, Y, st); break;
    }
}

template <bool BLOCK>
void dispatch(const CsrView& A, bool vreal, bool real, const VecSet* v, int nv, const void* X, void* Y, hipStream_t st) {
    // real vectors need real values; complex vectors take either
    auto go = [&](auto w, auto vr, auto xr) {
        constexpr int W = decltype(w)::value;
        constexpr bool VR = decltype(vr)::value, XR = decltype(xr)::value;
        if (A.sell_ptr) {
            if constexpr (BLOCK) launch_sell_block<W, VR, XR>(A, nv, X, Y, st);
            else launch_sell_set<W, VR, XR>(A, *v, st);
        } else {
            if constexpr (BLOCK) pick_lanes_block<W, VR, XR>(A, nv, X, Y, st);
            else pick_lanes<W, VR, XR>(A, *v, st);
        }
    };
    using T = std::true_type;
    using F = std::false_type;
    auto by_width = [&](auto vr, auto xr) {
        if (A.width == 1) go(std::integral_constant<int, 1>{}, vr, xr);
        else if (A.width == 2) go(std::integral_constant<int, 2>{}, vr, xr);
        else if (A.width == 3) go(std::integral_constant<int, 3>{}, vr, xr);
        else go(std::integral_constant<int, 0>{}, vr, xr);
    };
    if (real) by_width(T{}, T{});
    else if (vreal) by_width(T{}, F{});
    else by_width(F{}, F{});
}

__global__ void k_pack_col_ids(u32* col, const u8* id, u64 nnz) {
    for (u64 e = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; e < nnz; e += static_cast<u64>(gridDim.x) * blockDim.x)
        col[e] |= static_cast<u32>(id[e]) << 24;
}

// One thread per row; a wavefront is one slice (the grid's blocks are multiples of 64 threads).
__global__ __launch_bounds__(256) void k_sell_lengths(const u64* row_ptr, i64 rows, u16* sell_len, u64* slice_words,
                                                      u32* too_long) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    u32 n = 0;
    if (r < rows) {
        const u64 len = row_ptr[r + 1] - row_ptr[r];
        if (len > 0xFFFFu) atomicOr(too_long, 1u);
        n = len > 0xFFFFu ? 0xFFFFu : static_cast<u32>(len);
        sell_len[r] = static_cast<u16>(n);
    }
    u32 m = n;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const u32 o = static_cast<u32>(__shfl_xor(static_cast<int>(m), off, 64));
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && r < rows) slice_words[r >> 6] = static_cast<u64>((m + 3u) & ~3u) * 64;
}

// diag_first (a square operator): a row's diagonal entry, when it has one, moves to its first slot,
// the others keeping their order -- the k = 0 loads of a wavefront then read x[row] for 64
// consecutive rows, one coalesced line instead of 64 gathers.
__global__ __launch_bounds__(256) void k_sell_fill(const u64* row_ptr, const u32* col, const void* vals, int vb,
                                                   i64 rows, const u64* sell_ptr, u32* sell_col, void* sell_vals,
                                                   int diag_first, int packed) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const u64 e0 = row_ptr[r], n = row_ptr[r + 1] - e0;
    const u64 slice = sell_ptr[r >> 6];
    const u32 j = static_cast<u32>(r & 63);
    u64 diag = n;   // the diagonal entry's position in the row (n: none)
    if (diag_first)
        for (u64 k = 0; k < n && diag == n; ++k)
            if ((packed ? (col[e0 + k] & 0xFFFFFFu) : col[e0 + k]) == static_cast<u32>(r)) diag = k;
    for (u64 k = 0; k < n; ++k) {
        // slot k takes the diagonal first, then the row's other entries in order
        const u64 from = diag == n ? k : (k == 0 ? diag : (k <= diag ? k - 1 : k));
        const u64 dst = sell_at(slice, j, static_cast<u32>(k)), src = e0 + from;
        sell_col[dst] = col[src];
        if (vb == 1) static_cast<u8*>(sell_vals)[dst] = static_cast<const u8*>(vals)[src];
        else if (vb == 2) static_cast<u16*>(sell_vals)[dst] = static_cast<const u16*>(vals)[src];
        else if (vb == 8) static_cast<double*>(sell_vals)[dst] = static_cast<const double*>(vals)[src];
        else if (vb == 16) static_cast<Zd*>(sell_vals)[dst] = static_cast<const Zd*>(vals)[src];
    }
}

}  // namespace

void sell_lengths(const u64* row_ptr, i64 rows, u16* sell_len, u64* slice_words, u32* too_long, hipStream_t st) {
    if (rows <= 0) return;
    const u64 g = (static_cast<u64>(rows) + 255) / 256;
    hipLaunchKernelGGL(k_sell_lengths, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, row_ptr, rows, sell_len,
                       slice_words, too_long);
}

std::size_t sell_scan_temp_bytes(i64 n_slices) {
    std::size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const u64*>(nullptr), static_cast<u64*>(nullptr),
                                           static_cast<int>(n_slices + 1));
    return bytes;
}

void sell_scan(const u64* slice_words, i64 n_slices, u64* sell_ptr, void* temp, std::size_t temp_bytes, hipStream_t st) {
    (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, slice_words, sell_ptr, static_cast<int>(n_slices + 1), st);
}

void sell_fill(const u64* row_ptr, const u32* col, const void* vals, int value_bytes, i64 rows, const u64* sell_ptr,
               u32* sell_col, void* sell_vals, bool diag_first, bool packed, hipStream_t st) {
    if (rows <= 0) return;
    const u64 g = (static_cast<u64>(rows) + 255) / 256;
    hipLaunchKernelGGL(k_sell_fill, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, row_ptr, col, vals, value_bytes,
                       rows, sell_ptr, sell_col, sell_vals, diag_first ? 1 : 0, packed ? 1 : 0);
}

void pack_col_ids(u32* col, const u8* id, u64 nnz, hipStream_t st) {
    if (!nnz) return;
    u64 g = (nnz + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_pack_col_ids, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, col, id, nnz);
}

void csr_spmv(const CsrView& A, bool real, const VecSet& v, hipStream_t st) {
    if (A.rows == 0 || v.k == 0) return;
    dispatch<false>(A, A.values_real, real, &v, 0, nullptr, nullptr, st);
}

void csr_spmv_block(const CsrView& A, bool real, int nv, const void* X, void* Y, hipStream_t st) {
    if (A.rows == 0) return;
    dispatch<true>(A, A.values_real, real, nullptr, nv, X, Y, st);
}

unsigned sell_lanczos_blocks(const CsrView& A) { return grid_blocks(A, static_cast<u64>(A.rows)); }

bool sell_lanczos(const CsrView& A, bool real, const void* v, const void* vprev, double beta, void* w,
                  double* partial, Zd* alpha, hipStream_t st) {
    if (!A.sell_ptr || A.rows == 0 || (real && !A.values_real)) return false;
    const unsigned nb = sell_lanczos_blocks(A);
    const dim3 grid(nb), block(block_threads(A));
    Ptrs<1> p;
    p.in[0] = v;
    p.out[0] = w;
    auto go = [&](auto wd, auto vr, auto xr) {
        constexpr int W = decltype(wd)::value;
        constexpr bool VR = decltype(vr)::value, XR = decltype(xr)::value;
        if (A.sell_acc == 4) {
            if (A.unroll >= 16) hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 16, 4>), grid, block, 0, st, A, p, vprev, beta, partial);
            else if (A.unroll >= 8) hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 8, 4>), grid, block, 0, st, A, p, vprev, beta, partial);
            else hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 4, 4>), grid, block, 0, st, A, p, vprev, beta, partial);
        } else {
            if (A.unroll >= 16) hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 16, 1>), grid, block, 0, st, A, p, vprev, beta, partial);
            else if (A.unroll >= 8) hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 8, 1>), grid, block, 0, st, A, p, vprev, beta, partial);
            else hipLaunchKernelGGL((k_sell_lanczos<W, VR, XR, 4, 1>), grid, block, 0, st, A, p, vprev, beta, partial);
        }
    };
    using T = std::true_type;
    using F = std::false_type;
    auto by_width = [&](auto vr, auto xr) {
        if (A.width == 1) go(std::integral_constant<int, 1>{}, vr, xr);
        else if (A.width == 2) go(std::integral_constant<int, 2>{}, vr, xr);
        else if (A.width == 3) go(std::integral_constant<int, 3>{}, vr, xr);
        else go(std::integral_constant<int, 0>{}, vr, xr);
    };
    if (real) by_width(T{}, T{});
    else if (A.values_real) by_width(T{}, F{});
    else by_width(F{}, F{});
    hipLaunchKernelGGL(k_lanczos_finish, dim3(1), dim3(256), 0, st, partial, static_cast<int>(nb), alpha);
    return true;
}

}  // namespace qsa::dev

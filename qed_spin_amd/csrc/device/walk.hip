// Term application in a symmetry sector: the on-the-fly row walk (gather lane) and the merged
// reduced-CSR build (a staged walk, a wavefront-per-row sort and merge, the dictionary, compaction).
//
// A row r of O between sector R (rows) and sector C (columns) comes from the walk of the program
// of O^dagger from R's representative s_r: each acting group gives its target t and h, the target
// is reduced to its orbit representative with the character phase (index_and_projection), and the
// entry inv_norm[r] conj(h proj) lands at the target's index (sector_rows.h:41-56). The gather lane
// sums entries in emission order; the CSR build merges each row first -- columns sorted (stable),
// equal columns summed in emission order, exact zeros dropped (sector_rows.h:147-158) -- so its
// nnz and values are the reference's merged CSR entry for entry.
#include <hipcub/hipcub.hpp>

#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

constexpr int kMaxRow = 128;     // entries of one row before merging (= acting groups)
constexpr int kLdsSlots = 512;   // per-block value dictionary

inline unsigned blocks_for(u64 n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

template <int NV> struct Ptrs {
    const void* in[NV];
    void* out[NV];
};

template <bool SLC, bool REAL, int NV>
__global__ __launch_bounds__(256) void k_gather(SectorView R, SectorView C, int same, ProgramView P, Ptrs<NV> v) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R.dim) return;
    const u64 s = R.reps[r];
    const double w = R.inv_norms[r];
    double ar[NV], ai[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) ar[k] = ai[k] = 0.0;
    for (u32 g = 0; g < P.n_groups; ++g) {
        u64 t;
        double hr, hi;
        if (!connection(P, s, g, t, hr, hi)) continue;
        double vr, vi;
        const i64 c = entry_of<SLC>(C, same != 0, r, s, w, t, hr, hi, vr, vi);
        if (c < 0) continue;
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
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if constexpr (REAL) static_cast<double*>(v.out[k])[r] = ar[k];
        else static_cast<Zd*>(v.out[k])[r] = Zd{ar[k], ai[k]};
    }
}

// ---- the gather lane with the column sector's action in LDS --------------------------------------
//
// The walk above reads the group action from the 256-entry byte tables (G x bpw x 256 words: 512 KB
// at N = 32 with 64 elements), the key tables and the Pascal table in global memory: about 21 L1
// line accesses per entry, which bound it. Here every workgroup holds the column sector's action by
// nibbles (G x nq x 16 words, 32-bit when N <= 32: 32 KB at N = 32, G = 64), the keyed flips, the
// key tables, the characters and the rows of Pascal's triangle the rank needs in LDS (37 KB at N = 32,
// G = 64), and walks rows in a grid-stride loop so each workgroup of 16 waves loads them once. What stays in global memory per target:
// the sublattice lead and candidate entries, the rank directory record, the column's 1/norm and x.
// The arithmetic is the walk above, operation for operation, so the results are its bits.

template <class W> struct LdsAction {
    const double* chr;
    const double* chi;
    const W* img;     // G x nq x 16: images of nibbles (keyed when the sublattice order is engaged)
    const W* flip;    // G
    const W* key;     // nq x 16
    const W* unkey;   // nq x 16
    const W* binom;   // n_sites x nb: C(b, j)
    int nq, G, nb;
};

/// LDS bytes of the action of sector C for words of type W.
template <class W> std::size_t lds_action_bytes(const SectorView& C) {
    const std::size_t q16 = static_cast<std::size_t>(C.nq) * 16;
    std::size_t words = static_cast<std::size_t>(C.G) * q16 + static_cast<std::size_t>(C.G);
    if (C.slc_L > 0) words += 2 * q16;
    if (C.n_up >= 0) words += static_cast<std::size_t>(C.n_sites) * static_cast<std::size_t>(C.n_up + 1);
    return 16 * static_cast<std::size_t>(C.G) + sizeof(W) * words;
}

template <bool SLC, class W> __device__ LdsAction<W> stage_action(const SectorView& C, unsigned char* raw) {
    LdsAction<W> A;
    A.nq = C.nq;
    A.G = C.G;
    A.nb = C.n_up + 1;
    double* chr = reinterpret_cast<double*>(raw);
    double* chi = chr + C.G;
    W* img = reinterpret_cast<W*>(chi + C.G);
    const int n_img = C.G * C.nq * 16;
    W* flip = img + n_img;
    W* key = flip + C.G;
    W* unkey = key + (SLC ? C.nq * 16 : 0);
    W* bn = unkey + (SLC ? C.nq * 16 : 0);
    for (int i = threadIdx.x; i < C.G; i += blockDim.x) {
        chr[i] = C.chi_re[i];
        chi[i] = C.chi_im[i];
        flip[i] = static_cast<W>(C.nib_flip[i]);
    }
    for (int i = threadIdx.x; i < n_img; i += blockDim.x) img[i] = static_cast<W>(C.nib[i]);
    if constexpr (SLC) {
        for (int i = threadIdx.x; i < C.nq * 16; i += blockDim.x) {
            key[i] = static_cast<W>(C.nib_key[i]);
            unkey[i] = static_cast<W>(C.nib_unkey[i]);
        }
    }
    if (C.n_up >= 0)
        for (int i = threadIdx.x; i < C.n_sites * A.nb; i += blockDim.x)
            bn[i] = static_cast<W>(C.binom[(i / A.nb) * 65 + i % A.nb]);
    A.chr = chr;
    A.chi = chi;
    A.img = img;
    A.flip = flip;
    A.key = key;
    A.unkey = unkey;
    A.binom = bn;
    return A;
}

template <class W> __device__ __forceinline__ W nib_map(const W* tab, int nq, u64 s) {
    W o = 0;
    for (int q = 0; q < nq; ++q) o |= tab[(q << 4) | static_cast<int>((s >> (4 * q)) & 15u)];
    return o;
}

/// index_and_projection over the LDS action: the same running minimum, ties and character sums.
template <bool SLC, class W>
__device__ __forceinline__ i64 lds_index_and_projection(const SectorView& S, const LdsAction<W>& A, u64 t,
                                                        double& pr, double& pi) {
    if (S.n_up >= 0 && popc(t) != S.n_up) return -1;
    W best = ~W(0);
    double ar = 0.0, ai = 0.0;
    auto visit = [&](int g) {
        const W img = nib_map(A.img + g * A.nq * 16, A.nq, t) ^ A.flip[g];
        if (img < best) {
            best = img;
            ar = 0.0 + A.chr[g];
            ai = 0.0 - A.chi[g];
        } else if (img == best) {
            ar += A.chr[g];
            ai -= A.chi[g];
        }
    };
    u64 rb;
    if constexpr (!SLC) {
        for (int g = 0; g < A.G; ++g) visit(g);
        rb = static_cast<u64>(best);
    } else {
        for_each_candidate(S, static_cast<u64>(nib_map(A.key, A.nq, t)), visit);
        rb = static_cast<u64>(nib_map(A.unkey, A.nq, static_cast<u64>(best)));
    }
    if (S.n_up >= 0 && popc(rb) != S.n_up) return -1;
    i64 idx;
    if (S.rank_dir) {
        u64 r = rb;
        if (S.n_up >= 0) {
            r = 0;
            int j = 0;
            for (u64 x = rb; x; x &= x - 1) {
                ++j;
                r += static_cast<u64>(A.binom[(__ffsll(static_cast<long long>(x)) - 1) * A.nb + j]);
            }
        }
        const RankRec e = S.rank_dir[r >> 6];
        const u64 bit = 1ULL << (r & 63);
        if (!(e.bits & bit)) return -1;
        idx = static_cast<i64>(e.base) + popc(e.bits & (bit - 1ULL));
    } else {
        idx = index_of_rep(S, rb);
        if (idx < 0) return -1;
    }
    const double w = S.inv_norms[idx];
    pr = ar * w;
    pi = ai * w;
    return idx;
}

/// entry_of over the LDS action: the same operations, so the same bits.
template <bool SLC, class W>
__device__ __forceinline__ i64 lds_entry_of(const SectorView& C, const LdsAction<W>& A, bool same, i64 r, u64 s,
                                            double w, u64 t, double hr, double hi, double& vr, double& vi) {
    if (same && t == s) {
        vr = hr;
        vi = -hi;
        return r;
    }
    double pr, pi;
    const i64 c = lds_index_and_projection<SLC, W>(C, A, t, pr, pi);
    if (c < 0) return -1;
    const double xr = hr * pr - hi * pi;
    const double xi = hr * pi + hi * pr;
    vr = w * xr;
    vi = w * (-xi);
    return c;
}

template <bool SLC, bool REAL, int NV, class W>
__global__ __launch_bounds__(1024) void k_gather_lds(SectorView R, SectorView C, int same, ProgramView P,
                                                    Ptrs<NV> v) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const LdsAction<W> A = stage_action<SLC, W>(C, lds_raw);
    __syncthreads();
    const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
    for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < R.dim; r += stride) {
        const u64 s = R.reps[r];
        const double w = R.inv_norms[r];
        double ar[NV], ai[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) ar[k] = ai[k] = 0.0;
        for (u32 g = 0; g < P.n_groups; ++g) {
            u64 t;
            double hr, hi;
            if (!connection(P, s, g, t, hr, hi)) continue;
            double vr, vi;
            const i64 c = lds_entry_of<SLC, W>(C, A, same != 0, r, s, w, t, hr, hi, vr, vi);
            if (c < 0) continue;
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
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if constexpr (REAL) static_cast<double*>(v.out[k])[r] = ar[k];
            else static_cast<Zd*>(v.out[k])[r] = Zd{ar[k], ai[k]};
        }
    }
}

/// LDS budget of the gather workgroup's action. Workgroups of 1024 threads (16 waves) share one copy:
/// two of them per CU (32 waves, eight per SIMD) fit in 160 KB.
constexpr std::size_t kLdsActionMax = 64 * 1024;
/// Workgroups of the grid-stride gather: four per CU of a 256-CU device, each loading the action once.
constexpr u64 kGatherLdsBlocks = 1024;

/// Launch the LDS-action gather when the action fits kLdsActionMax; false when it does not.
template <bool SLC, bool REAL>
bool launch_gather_lds(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, const VecSet& v,
                       hipStream_t st) {
    if (!C.nib || C.nq <= 0) return false;
    const bool narrow = C.n_sites <= 32;
    const std::size_t bytes = narrow ? lds_action_bytes<u32>(C) : lds_action_bytes<u64>(C);
    if (bytes > kLdsActionMax) return false;
    const u64 need = blocks_for(static_cast<u64>(R.dim), 1024);
    const dim3 grid(static_cast<unsigned>(need < kGatherLdsBlocks ? need : kGatherLdsBlocks)), block(1024);
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
            if (narrow)
                hipLaunchKernelGGL((k_gather_lds<SLC, REAL, NV, u32>), grid, block, bytes, st, R, C, same ? 1 : 0, P,
                                   p);
            else
                hipLaunchKernelGGL((k_gather_lds<SLC, REAL, NV, u64>), grid, block, bytes, st, R, C, same ? 1 : 0, P,
                                   p);
        };
        if (nv == 8) go(std::integral_constant<int, 8>{});
        else if (nv == 4) go(std::integral_constant<int, 4>{});
        else if (nv == 2) go(std::integral_constant<int, 2>{});
        else go(std::integral_constant<int, 1>{});
        done += nv;
    }
    return true;
}

template <bool SLC, bool REAL>
void launch_gather(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, const VecSet& v,
                   hipStream_t st) {
    if (launch_gather_lds<SLC, REAL>(R, C, same, P, v, st)) return;
    const dim3 grid(blocks_for(static_cast<u64>(R.dim), 256)), block(256);
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
            hipLaunchKernelGGL((k_gather<SLC, REAL, NV>), grid, block, 0, st, R, C, same ? 1 : 0, P, p);
        };
        if (nv == 8) go(std::integral_constant<int, 8>{});
        else if (nv == 4) go(std::integral_constant<int, 4>{});
        else if (nv == 2) go(std::integral_constant<int, 2>{});
        else go(std::integral_constant<int, 1>{});
        done += nv;
    }
}

__device__ __forceinline__ u64 bits_of(double x) { return static_cast<u64>(__double_as_longlong(x)); }

__device__ __forceinline__ u64 mix64(u64 z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// The dictionary key of a value: each word through its own bijective mix before they are combined,
// so structured pairs (v and -v: both sign bits flipped) do not cancel; 0 marks an empty slot.
__device__ __forceinline__ u64 fingerprint(u64 re, u64 im) {
    const u64 a = mix64(re ^ 0x9E3779B97F4A7C15ULL);
    const u64 b = mix64(im ^ 0xD1B54A32D192ED03ULL);
    const u64 z = mix64(a ^ ((b << 32) | (b >> 32)));
    return z ? z : 1ULL;
}

// ---- one walk: the rows staged, sorted and merged, then the dictionary and the CSR -------------
// Entry e of local row i of a chunk of rows sits at stage[e * rows + i] (a column per entry slot:
// the threads of a wave write and read consecutive rows together). The walk (k_csr_emit, one thread per row, the column
// sector's action in LDS when it fits) writes each row's entries in emission order; k_csr_sort then
// takes one wavefront per row: a bitonic sort of the keys (column, emission index) across the lanes,
// equal columns summed in emission order, exact zeros dropped, the row written back in place -- the
// merged row of the per-thread walk above, entry for entry and bit for bit, without its private
// arrays.

template <bool SLC, bool LDS, class W>
__global__ __launch_bounds__(1024) void k_csr_emit(SectorView R, SectorView C, int same, ProgramView P, i64 row0,
                                                   i64 n_rows, u32* row_len, u32* stage_col, Zd* stage_val) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LdsAction<W> A{};
    if constexpr (LDS) {
        A = stage_action<SLC, W>(C, lds_raw);
        __syncthreads();
    }
    const u64 rows = static_cast<u64>(n_rows);
    const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_rows; i += stride) {
        const i64 r = row0 + i;
        const u64 s = R.reps[r];
        const double w = R.inv_norms[r];
        u32 n = 0;
        for (u32 g = 0; g < P.n_groups; ++g) {
            u64 t;
            double hr, hi;
            if (!connection(P, s, g, t, hr, hi)) continue;
            double vr, vi;
            i64 c;
            if constexpr (LDS) c = lds_entry_of<SLC, W>(C, A, same != 0, r, s, w, t, hr, hi, vr, vi);
            else c = entry_of<SLC>(C, same != 0, r, s, w, t, hr, hi, vr, vi);
            if (c < 0) continue;
            stage_col[static_cast<u64>(n) * rows + static_cast<u64>(i)] = static_cast<u32>(c);
            stage_val[static_cast<u64>(n) * rows + static_cast<u64>(i)] = Zd{vr, vi};
            ++n;
        }
        row_len[i] = n;
    }
}

/// The value of element x (of the row's 64 E sorted elements, element x in lane x % 64, slot x / 64)
/// read by every lane from wherever it sits.
template <int E, class T> __device__ __forceinline__ T elem_at(const T (&v)[E], int x) {
    const int src = x & 63;
    T out = __shfl(v[0], src, 64);
    if constexpr (E == 2) {
        const T hi = __shfl(v[1], src, 64);
        if (x >= 64) out = hi;
    }
    return out;
}

/// One row (n <= 64 E entries at stage[e * rows + r]) sorted by (column, emission index), equal
/// columns summed in emission order, exact zeros dropped, written back in place; by one wavefront.
/// K: the key word, (column << 7) | e (32-bit when the columns are below 2^25).
template <int E, class K>
__device__ __forceinline__ void sort_merge_row(u64 rows, u64 r, int n, int lane, u32* row_len, u32* stage_col,
                                               Zd* stage_val) {
    constexpr K kPad = ~K(0);
    K key[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int e = lane + 64 * q;
        key[q] = e < n ? static_cast<K>((static_cast<K>(stage_col[static_cast<u64>(e) * rows + r]) << 7) |
                                        static_cast<K>(e))
                       : kPad;
    }
    // bitonic sort of the 64 E keys, element x = 64 q + lane
#pragma unroll
    for (int k = 2; k <= 64 * E; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                // E == 2, k == 128: element x and x + 64 sit in this lane's two slots, ascending
                const K a = key[0] < key[E - 1] ? key[0] : key[E - 1];
                const K b = key[0] < key[E - 1] ? key[E - 1] : key[0];
                key[0] = a;
                key[E - 1] = b;
                continue;
            }
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int x = 64 * q + lane;
                const K other = __shfl_xor(key[q], j, 64);
                const bool up = (x & k) == 0, lower = (x & j) == 0;
                key[q] = (lower == up) ? (other < key[q] ? other : key[q]) : (other > key[q] ? other : key[q]);
            }
        }
    }
    u32 col[E];
    double vr[E], vi[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int x = 64 * q + lane;
        col[q] = static_cast<u32>(key[q] >> 7);
        vr[q] = vi[q] = 0.0;
        if (x < n) {
            const Zd v = stage_val[static_cast<u64>(key[q] & K(127)) * rows + r];
            vr[q] = v.re;
            vi[q] = v.im;
        }
    }
    // a run of equal columns is summed by its first element, left to right (emission order)
    bool head[E], cont[E];
    double sr[E], si[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int x = 64 * q + lane;
        const u32 prev = elem_at<E>(col, x > 0 ? x - 1 : 0);
        const u32 next = elem_at<E>(col, x + 1 < 64 * E ? x + 1 : x);
        head[q] = x < n && (x == 0 || prev != col[q]);
        cont[q] = head[q] && x + 1 < n && next == col[q];
        sr[q] = vr[q];
        si[q] = vi[q];
    }
    bool any_run = false;
#pragma unroll
    for (int q = 0; q < E; ++q) any_run = any_run || cont[q];
    if (__any(any_run)) {
        for (int d = 1; d < n; ++d) {
            bool more = false;
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int x = 64 * q + lane;
                const int y = x + d < 64 * E ? x + d : x;
                const u32 cy = elem_at<E>(col, y);
                const double ry = elem_at<E>(vr, y), iy = elem_at<E>(vi, y);
                if (head[q] && x + d < n && cy == col[q]) {
                    sr[q] += ry;
                    si[q] += iy;
                    more = true;
                }
            }
            if (!__any(more)) break;
        }
    }
    bool keep[E];
    u64 mask[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        keep[q] = head[q] && !(sr[q] == 0.0 && si[q] == 0.0);
        mask[q] = __ballot(keep[q]);
    }
    const u64 below = (1ULL << lane) - 1ULL;
    u32 before = 0;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        if (keep[q]) {
            const u64 pos = before + static_cast<u32>(__popcll(mask[q] & below));
            stage_col[pos * rows + r] = col[q];
            stage_val[pos * rows + r] = Zd{sr[q], si[q]};
        }
        before += static_cast<u32>(__popcll(mask[q]));
    }
    if (lane == 0) row_len[r] = before;
}

/// One wavefront per row; rows of at most 64 entries take the one-key-per-lane network.
template <int E, class K>
__global__ __launch_bounds__(256) void k_csr_sort(i64 rows_n, u32* row_len, u32* stage_col, Zd* stage_val) {
    const i64 r = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (r >= rows_n) return;   // the whole wave
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int n = static_cast<int>(row_len[r]);
    if (E == 2 && n > 64)
        sort_merge_row<E, K>(static_cast<u64>(rows_n), static_cast<u64>(r), n, lane, row_len, stage_col, stage_val);
    else
        sort_merge_row<1, K>(static_cast<u64>(rows_n), static_cast<u64>(r), n, lane, row_len, stage_col, stage_val);
}

// The dictionary from the staged values: per workgroup an LDS table of its rows' distinct values,
// then into the global table (the same set of values as the walking count pass inserts).
__global__ __launch_bounds__(256) void k_csr_stage_dict(i64 rows_n, const u32* row_len, const Zd* stage_val,
                                                        DictTable D) {
    __shared__ unsigned long long s_fp[kLdsSlots];
    __shared__ u64 s_re[kLdsSlots], s_im[kLdsSlots];
    __shared__ int s_full;
    for (int i = threadIdx.x; i < kLdsSlots; i += blockDim.x) s_fp[i] = 0ULL;
    if (threadIdx.x == 0) s_full = 0;
    __syncthreads();
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < rows_n) {
        const int n = static_cast<int>(row_len[r]);
        const u64 rows = static_cast<u64>(rows_n);
        for (int e = 0; e < n; ++e) {
            const Zd v = stage_val[static_cast<u64>(e) * rows + r];
            const u64 re = bits_of(v.re), im = bits_of(v.im);
            const u64 fp = fingerprint(re, im);
            const int h = static_cast<int>((fp >> 17) & (kLdsSlots - 1));
            bool placed = false;
            for (int p = 0; p < kLdsSlots && !placed; ++p) {
                const int i = (h + p) & (kLdsSlots - 1);
                const unsigned long long old = atomicCAS(&s_fp[i], 0ULL, static_cast<unsigned long long>(fp));
                if (old == 0ULL) {
                    s_re[i] = re;
                    s_im[i] = im;
                    placed = true;
                } else if (old == fp) {
                    placed = true;
                }
            }
            if (!placed) s_full = 1;
        }
    }
    __syncthreads();
    if (s_full) {
        if (threadIdx.x == 0) atomicOr(D.flags, 1u);
        return;
    }
    for (int i = threadIdx.x; i < kLdsSlots; i += blockDim.x) {
        const u64 fp = s_fp[i];
        if (!fp) continue;
        const int h = static_cast<int>((fp >> 17) & (DictTable::kSlots - 1));
        bool placed = false;
        for (int p = 0; p < 4096 && !placed; ++p) {
            const int j = (h + p) & (DictTable::kSlots - 1);
            const unsigned long long old =
                atomicCAS(reinterpret_cast<unsigned long long*>(&D.fp[j]), 0ULL, static_cast<unsigned long long>(fp));
            if (old == 0ULL) {
                D.re[j] = s_re[i];
                D.im[j] = s_im[i];
                atomicAdd(D.count, 1u);
                placed = true;
            } else if (old == fp) {
                placed = true;
            }
        }
        if (!placed) atomicOr(D.flags, 1u);
    }
}

// The CSR from the stage: columns, and value ids (width 1 / 2) or full values (width 0).
template <int WIDTH, bool REAL>
__global__ __launch_bounds__(256) void k_csr_compact(i64 rows_n, const u64* row_ptr, const u32* stage_col,
                                                     const Zd* stage_val, u32* col, void* vals, DictTable D) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows_n) return;
    const u64 rows = static_cast<u64>(rows_n);
    const u64 e0 = row_ptr[r], n = row_ptr[r + 1] - e0;
    for (u64 e = 0; e < n; ++e) {
        col[e0 + e] = stage_col[e * rows + r];
        const Zd v = stage_val[e * rows + r];
        if constexpr (WIDTH == 0) {
            if constexpr (REAL) static_cast<double*>(vals)[e0 + e] = v.re;
            else static_cast<Zd*>(vals)[e0 + e] = v;
        } else {
            const u64 re = bits_of(v.re), im = bits_of(v.im);
            const u64 fp = fingerprint(re, im);
            const int h = static_cast<int>((fp >> 17) & (DictTable::kSlots - 1));
            int slot = -1;
            for (int p = 0; p < 4096; ++p) {
                const int j = (h + p) & (DictTable::kSlots - 1);
                const u64 f = D.fp[j];
                if (f == fp) {
                    slot = j;
                    break;
                }
                if (f == 0) break;
            }
            u32 id = 0;
            if (slot < 0 || D.re[slot] != re || D.im[slot] != im) atomicOr(D.flags, 2u);
            else id = D.slot_id[slot];
            if constexpr (WIDTH == 1) static_cast<u8*>(vals)[e0 + e] = static_cast<u8>(id);
            else static_cast<u16*>(vals)[e0 + e] = static_cast<u16>(id);
        }
    }
}

}  // namespace

bool csr_walk(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, i64 row0, i64 n_rows,
              u32* row_len, u32* stage_col, Zd* stage_val, hipStream_t st) {
    if (P.n_groups > static_cast<u32>(kMaxRow)) return false;
    if (n_rows <= 0) return true;
    const u64 need = blocks_for(static_cast<u64>(n_rows), 1024);
    const dim3 grid(static_cast<unsigned>(need < kGatherLdsBlocks ? need : kGatherLdsBlocks)), block(1024);
    const bool slc = C.slc_L > 0, narrow = C.n_sites <= 32;
    const std::size_t bytes = !C.nib ? kLdsActionMax + 1 : narrow ? lds_action_bytes<u32>(C) : lds_action_bytes<u64>(C);
#define QSA_EMIT(S, L, W, B)                                                                                           \
    hipLaunchKernelGGL((k_csr_emit<S, L, W>), grid, block, B, st, R, C, same ? 1 : 0, P, row0, n_rows, row_len,     \
                       stage_col, stage_val)
    if (bytes <= kLdsActionMax) {
        if (slc && narrow) QSA_EMIT(true, true, u32, bytes);
        else if (slc) QSA_EMIT(true, true, u64, bytes);
        else if (narrow) QSA_EMIT(false, true, u32, bytes);
        else QSA_EMIT(false, true, u64, bytes);
    } else if (slc) {
        QSA_EMIT(true, false, u64, 0);
    } else {
        QSA_EMIT(false, false, u64, 0);
    }
#undef QSA_EMIT
    const dim3 sgrid(blocks_for(static_cast<u64>(n_rows) * 64, 256)), sblock(256);
    const bool short_cols = C.dim < (i64{1} << 25);
#define QSA_SORT(E, K) \
    hipLaunchKernelGGL((k_csr_sort<E, K>), sgrid, sblock, 0, st, n_rows, row_len, stage_col, stage_val)
    if (P.n_groups <= 64 && short_cols) QSA_SORT(1, u32);
    else if (P.n_groups <= 64) QSA_SORT(1, u64);
    else if (short_cols) QSA_SORT(2, u32);
    else QSA_SORT(2, u64);
#undef QSA_SORT
    return true;
}

void csr_stage_dict(i64 rows, const u32* row_len, const Zd* stage_val, const DictTable& D, hipStream_t st) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_csr_stage_dict, dim3(blocks_for(static_cast<u64>(rows), 256)), dim3(256), 0, st, rows, row_len,
                       stage_val, D);
}

void csr_compact(i64 rows, const u64* row_ptr, const u32* stage_col, const Zd* stage_val, u32* col, void* vals,
                 int width, bool real, const DictTable& D, hipStream_t st) {
    if (rows <= 0) return;
    const dim3 grid(blocks_for(static_cast<u64>(rows), 256)), block(256);
#define QSA_COMPACT(W, RE) \
    hipLaunchKernelGGL((k_csr_compact<W, RE>), grid, block, 0, st, rows, row_ptr, stage_col, stage_val, col, vals, D)
    if (width == 1) QSA_COMPACT(1, false);
    else if (width == 2) QSA_COMPACT(2, false);
    else if (real) QSA_COMPACT(0, true);
    else QSA_COMPACT(0, false);
#undef QSA_COMPACT
}

void walk_gather(const SectorView& R, const SectorView& C, bool same, const ProgramView& P, bool real,
                 const VecSet& v, hipStream_t st) {
    if (R.dim == 0 || v.k == 0) return;
    const bool slc = C.slc_L > 0;
    if (slc && real) launch_gather<true, true>(R, C, same, P, v, st);
    else if (slc) launch_gather<true, false>(R, C, same, P, v, st);
    else if (real) launch_gather<false, true>(R, C, same, P, v, st);
    else launch_gather<false, false>(R, C, same, P, v, st);
}


namespace {
struct ToU64 {
    __host__ __device__ u64 operator()(u32 x) const { return static_cast<u64>(x); }
};
}  // namespace

std::size_t row_scan_temp_bytes(i64 rows) {
    std::size_t bytes = 0;
    hipcub::TransformInputIterator<u64, ToU64, const u32*> it(nullptr, ToU64{});
    hipcub::DeviceScan::InclusiveSum(nullptr, bytes, it, static_cast<u64*>(nullptr), static_cast<int>(rows));
    return bytes;
}

// row_ptr[0] = 0, row_ptr[r + 1] = sum of row_len[0..r].
void row_scan(const u32* row_len, i64 rows, u64* row_ptr, void* temp, std::size_t temp_bytes, hipStream_t st) {
    hipMemsetAsync(row_ptr, 0, sizeof(u64), st);
    if (rows == 0) return;
    hipcub::TransformInputIterator<u64, ToU64, const u32*> it(row_len, ToU64{});
    hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, it, row_ptr + 1, static_cast<int>(rows), st);
}

}  // namespace qsa::dev

namespace qsa::dev {

namespace {
template <bool SLC>
__global__ void k_index_proj(SectorView S, const u64* states, i64 n, i64* idx, double* proj) {
    const i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pr = 0.0, pi = 0.0;
    idx[i] = index_and_projection<SLC>(S, states[i], pr, pi);
    proj[2 * i] = pr;
    proj[2 * i + 1] = pi;
}
}  // namespace

void index_proj_batch(const SectorView& S, const u64* states, i64 n, i64* idx, double* proj, hipStream_t st) {
    if (n <= 0) return;
    const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
    if (S.slc_L > 0) hipLaunchKernelGGL(k_index_proj<true>, grid, block, 0, st, S, states, n, idx, proj);
    else hipLaunchKernelGGL(k_index_proj<false>, grid, block, 0, st, S, states, n, idx, proj);
}

}  // namespace qsa::dev

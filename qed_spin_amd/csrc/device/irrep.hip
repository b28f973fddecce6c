// Sectors of an irrep of dimension d > 1: the stabiliser of every representative, and the operator
// between two such sectors built representative by representative.
//
// Representative r holds rank(r) <= d states |r; a> = sqrt(d / |G|) sum_j C_r[j][a] Ptilde_0j |r>,
// Ptilde_ij = sum_g conj(D(g)_ij) U(g), with C_r its stabiliser class's matrix (C^dag M C = I for
// M = sum over the stabiliser of conj(D(s))). A connection (t, h) of the row representative's state
// s gives the block conj(h) C_r^dag A(t) C_c between r's states and those of t's representative c,
// A(t) = sum_{g: g t = rep_c} D(g)^T (include/ed/matvec/rep_symmetry_basis_policy.h:245-275,
// include/ed/matvec/sector_rows.h:72-106); the diagonal connection (t = s) gives conj(h) I.
#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

__global__ __launch_bounds__(256) void k_stab_masks(SectorView S, int words, u64* masks) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= S.dim) return;
    const u64 s = S.reps[r];
    for (int w = 0; w < words; ++w) {
        u64 m = 0;
        for (int b = 0; b < 64; ++b) {
            const int g = w * 64 + b;
            if (g >= S.G) break;
            if (apply_element(S, g, s) == s) m |= 1ULL << b;
        }
        masks[static_cast<u64>(r) * words + w] = m;
    }
}

__device__ __forceinline__ Zd zadd(Zd a, Zd b) { return Zd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Zd zmul(Zd a, Zd b) { return Zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Zd zconj(Zd a) { return Zd{a.re, -a.im}; }

// The representative index of t in the column sector (or -1) and A = sum_{g: g t = rep} D(g)^T: one
// pass over the images with a running minimum, A restarting at a new minimum.
template <bool SLC, int D>
__device__ __forceinline__ i64 index_and_matrix(const SectorView& S, const IrrepView& I, u64 t, Zd (&A)[D][D]) {
    if (S.n_up >= 0 && popc(t) != S.n_up) return -1;
    u64 rb = ~0ULL, best = ~0ULL;
    auto visit = [&](int g, u64 img, u64 key) {
        if (key > best) return;
        if (key < best) {
            best = key;
            rb = img;
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) A[i][j] = Zd{0.0, 0.0};
        }
        const Zd* Dg = I.D + static_cast<size_t>(g) * D * D;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) A[i][j] = zadd(A[i][j], Dg[j * D + i]);
    };
    if constexpr (!SLC) {
        for (int g = 0; g < S.G; ++g) {
            const u64 img = apply_element(S, g, t);
            visit(g, img, img);
        }
    } else {
        const u64 k = key_of(S, t);
        for_each_candidate(S, k, [&](int g) {
            const u64 img = apply_element(S, g, t);
            visit(g, img, key_of(S, img));
        });
    }
    if (S.n_up >= 0 && popc(rb) != S.n_up) return -1;
    return index_of_rep(S, rb);
}

// One thread per row representative. Count pass (val null): each of its rank rows gets the number
// of entries the walk emits; fill pass: the entries, in walk order, into the rows' slots.
template <bool SLC, int D>
__global__ __launch_bounds__(256) void k_irrep_rows(SectorView R, IrrepView IR, SectorView C, IrrepView IC, int same,
                                                    ProgramView P, u32* row_len, const u64* row_ptr, u32* col, Zd* val) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R.dim) return;
    const u64 s = R.reps[r];
    const int cls = IR.rep_class[r];
    const int rr = IR.class_rank[cls];
    const Zd* Cr = IR.C + static_cast<size_t>(cls) * D * D;
    const u64 o_r = IR.offset[r];
    u32 count = 0;
    for (u32 g = 0; g < P.n_groups; ++g) {
        u64 t;
        double hr, hi;
        if (!connection(P, s, g, t, hr, hi)) continue;
        const Zd ch{hr, -hi};   // conj(h)
        if (same && t == s) {
            if (val)
                for (int a = 0; a < rr; ++a) {
                    const u64 pos = row_ptr[o_r + a] + count;
                    col[pos] = static_cast<u32>(o_r + a);
                    val[pos] = ch;
                }
            count += 1;
            continue;
        }
        Zd A[D][D];
        const i64 c = index_and_matrix<SLC, D>(C, IC, t, A);
        if (c < 0) continue;   // t's orbit holds no state of the irrep
        const int cc = IC.rep_class[c];
        const int rc = IC.class_rank[cc];
        if (val) {
            const Zd* Cc = IC.C + static_cast<size_t>(cc) * D * D;
            const u64 o_c = IC.offset[c];
            // column b of T = A C_c, then entry (a, b) = conj(h) (C_r^dag T)[a][b] into slot b of
            // rank row a: one column of T in registers at a time besides A
            for (int b = 0; b < rc; ++b) {
                Zd t[D];
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    Zd acc{0.0, 0.0};
#pragma unroll
                    for (int j = 0; j < D; ++j) acc = zadd(acc, zmul(A[i][j], Cc[j * D + b]));
                    t[i] = acc;
                }
                for (int a = 0; a < rr; ++a) {
                    Zd acc{0.0, 0.0};
#pragma unroll
                    for (int i = 0; i < D; ++i) acc = zadd(acc, zmul(zconj(Cr[i * D + a]), t[i]));
                    const u64 pos = row_ptr[o_r + a] + count + static_cast<u64>(b);
                    col[pos] = static_cast<u32>(o_c + b);
                    val[pos] = zmul(ch, acc);
                }
            }
        }
        count += static_cast<u32>(rc);
    }
    if (!val)
        for (int a = 0; a < rr; ++a) row_len[o_r + a] = count;
}

// Per row (one thread; the rows of an irrep operator are short): the entries sorted by column with a
// stable insertion sort, equal columns summed in walk order, exact zeros dropped; the new length
// into row_len.
__global__ __launch_bounds__(256) void k_row_merge(i64 rows, const u64* row_ptr, u32* col, Zd* val, u32* row_len) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const u64 b = row_ptr[r], e = row_ptr[r + 1];
    for (u64 i = b + 1; i < e; ++i) {
        const u32 c = col[i];
        const Zd v = val[i];
        u64 j = i;
        while (j > b && col[j - 1] > c) {
            col[j] = col[j - 1];
            val[j] = val[j - 1];
            --j;
        }
        col[j] = c;
        val[j] = v;
    }
    u64 w = b;
    for (u64 i = b; i < e;) {
        const u32 c = col[i];
        Zd acc = val[i];
        u64 k = i + 1;
        while (k < e && col[k] == c) {
            acc = zadd(acc, val[k]);
            ++k;
        }
        if (acc.re != 0.0 || acc.im != 0.0) {
            col[w] = c;
            val[w] = acc;
            ++w;
        }
        i = k;
    }
    row_len[r] = static_cast<u32>(w - b);
}

__global__ __launch_bounds__(256) void k_row_move(i64 rows, const u64* old_ptr, const u64* new_ptr, const u32* col,
                                                  const Zd* val, u32* ncol, Zd* nval) {
    const i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const u64 o = old_ptr[r], n0 = new_ptr[r], n = new_ptr[r + 1] - n0;
    for (u64 k = 0; k < n; ++k) {
        ncol[n0 + k] = col[o + k];
        nval[n0 + k] = val[o + k];
    }
}

template <int D>
void launch_rows(const SectorView& R, const IrrepView& IR, const SectorView& C, const IrrepView& IC, bool same,
                 const ProgramView& P, u32* row_len, const u64* row_ptr, u32* col, Zd* val, hipStream_t st) {
    const unsigned grid = static_cast<unsigned>((R.dim + 255) / 256);
    if (C.slc_L > 0)
        hipLaunchKernelGGL((k_irrep_rows<true, D>), dim3(grid), dim3(256), 0, st, R, IR, C, IC, same ? 1 : 0, P,
                           row_len, row_ptr, col, val);
    else
        hipLaunchKernelGGL((k_irrep_rows<false, D>), dim3(grid), dim3(256), 0, st, R, IR, C, IC, same ? 1 : 0, P,
                           row_len, row_ptr, col, val);
}

}  // namespace

void csr_merge_rows(i64 rows, const u64* row_ptr, u32* col, Zd* val, u32* row_len, hipStream_t st) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_row_merge, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, st, rows, row_ptr, col,
                       val, row_len);
}

void csr_move_rows(i64 rows, const u64* old_ptr, const u64* new_ptr, const u32* col, const Zd* val, u32* ncol, Zd* nval,
                   hipStream_t st) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_row_move, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, st, rows, old_ptr,
                       new_ptr, col, val, ncol, nval);
}

void stab_masks(const SectorView& S, int words, u64* masks, hipStream_t st) {
    if (S.dim <= 0) return;
    hipLaunchKernelGGL(k_stab_masks, dim3(static_cast<unsigned>((S.dim + 255) / 256)), dim3(256), 0, st, S, words,
                       masks);
}

bool irrep_rows(const SectorView& R, const IrrepView& IR, const SectorView& C, const IrrepView& IC, bool same,
                const ProgramView& P, u32* row_len, const u64* row_ptr, u32* col, Zd* val, hipStream_t st) {
    if (IR.d != IC.d) return false;
    if (R.dim <= 0) return true;
    switch (IR.d) {
    case 2: launch_rows<2>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 3: launch_rows<3>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 4: launch_rows<4>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 5: launch_rows<5>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 6: launch_rows<6>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 7: launch_rows<7>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    case 8: launch_rows<8>(R, IR, C, IC, same, P, row_len, row_ptr, col, val, st); return true;
    default: return false;
    }
}

}  // namespace qsa::dev

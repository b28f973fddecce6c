// Device-side view of a symmetry sector and of a compiled program, and the per-state operations
// every kernel shares: the group action (byte lookup tables), the sublattice key order, the
// representative search with its character phase (index_and_projection), the reverse lookup and
// the row walk's connection. Compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qsa::dev {

using u64 = std::uint64_t;
using i64 = std::int64_t;
using u32 = std::uint32_t;
using i32 = std::int32_t;
using u16 = std::uint16_t;
using u8 = std::uint8_t;

struct Zd {
    double re, im;
};

/// One record of a sector's rank directory: which of 64 consecutive ranks (colex ranks at fixed
/// n_up, the states themselves otherwise) are representatives, a bit each, and the index of the
/// first of them. Representatives ascend with their rank, so the index of a representative is the
/// record's base plus the set bits below its own. 16 bytes per 64 ranks: one aligned load.
struct alignas(16) RankRec {
    u64 bits;
    u32 base;
    u32 spare;
};

/// A sector as the kernels read it (all pointers device memory). The representative of an orbit
/// is its least member (in key order when slc_L > 0); reps ascend by state value.
struct SectorView {
    int n_sites = 0, n_up = -1, G = 0, bpw = 0;
    const u64* lut = nullptr;        // G x bpw x 256: image of byte b = v under element g
    const u64* flips = nullptr;      // G
    const double* chi_re = nullptr;  // G
    const double* chi_im = nullptr;  // G
    int slc_L = 0, slc_m = 0;
    const u64* to_key = nullptr;     // bpw x 256
    const u64* cand_rec = nullptr;   // m x 2^L: lead | count << 16 | (count == 1 ? element : offset) << 32
    const u16* cand = nullptr;       // the candidate lists, entry after entry
    const u64* reps = nullptr;       // dim, ascending
    const double* inv_norms = nullptr;
    i64 dim = 0;
    const RankRec* rank_dir = nullptr; // C(N, n_up) / 64 records (2^N / 64 when n_up < 0), or null
    const u64* binom = nullptr;      // 65 x 65 Pascal table, binom[n * 65 + k] = C(n, k)
    // The action by nibbles, for kernels that hold it in LDS: nq = ceil(N / 4) positions; nib[(g nq +
    // q) 16 + v] is the image of nibble q = v under element g's permutation and nib_flip[g] its flip,
    // both in key positions when slc_L > 0 (the key is linear in the state, so key(g s) is the OR of
    // the nibbles' keyed images XOR the keyed flip); nib_key / nib_unkey (nq x 16) move a state's
    // nibbles to key positions and back.
    int nq = 0;
    const u64* nib = nullptr;
    const u64* nib_flip = nullptr;
    const u64* nib_key = nullptr;
    const u64* nib_unkey = nullptr;
};

/// An irrep of dimension d > 1 on a sector (the reference's rep_sector.h d > 1 fields): D(g) per
/// element, and per representative its stabiliser class -- the class's rank (the representative's
/// number of states, at most d) and its d x d matrix C (columns from the rank on zero) -- and the
/// index of its first state. d = 1 leaves every pointer null.
struct IrrepView {
    int d = 1;
    const Zd* D = nullptr;          // G x d x d, row-major
    const Zd* C = nullptr;          // per class: d x d, row-major
    const u8* class_rank = nullptr; // per class
    const u16* rep_class = nullptr; // per representative
    const u64* offset = nullptr;    // per representative, and one past the last
};

/// The largest irrep dimension of a sector (the reference's kMaxIrrepDim, include/ed/matvec/
/// sector_rows.h:66). The row kernel holds a d x d block per thread.
constexpr int kMaxIrrepDim = 8;

/// A compiled program (program.h layout) in device memory.
struct ProgramView {
    u32 n_groups = 0;
    const u64* group_flip = nullptr;
    const i32* group_setbits = nullptr;
    const u32* group_vbegin = nullptr;
    const u64* vsub_val = nullptr;
    const u32* vsub_tbegin = nullptr;
    const u64* term_sign = nullptr;
    const double* coeff_re = nullptr;
    const double* coeff_im = nullptr;
};

__device__ __forceinline__ int popc(u64 x) { return __popcll(x); }

/// Element g on state s: OR of the byte tables, then the XOR flip (compiled_group.h:73-78).
__device__ __forceinline__ u64 apply_element(const SectorView& S, int g, u64 s) {
    const u64* l = S.lut + static_cast<size_t>(g) * S.bpw * 256;
    u64 o = 0;
    for (int b = 0; b < S.bpw; ++b) o |= l[(b << 8) + static_cast<int>((s >> (8 * b)) & 0xFFu)];
    return o ^ S.flips[g];
}

__device__ __forceinline__ u64 key_of(const SectorView& S, u64 s) {
    u64 k = 0;
    for (int b = 0; b < S.bpw; ++b) k |= S.to_key[(b << 8) + static_cast<int>((s >> (8 * b)) & 0xFFu)];
    return k;
}

__device__ __forceinline__ u32 pattern_of(const SectorView& S, u64 k, int j) {
    return static_cast<u32>(k >> (S.n_sites - (j + 1) * S.slc_L)) & ((1u << S.slc_L) - 1u);
}

/// The candidate record of block j of key k: the least leading block reachable from that block's
/// pattern (bits 0-15), how many elements reach it (bits 16-31) and the element itself when it is
/// the only one, else its list's offset in cand (bits 32-63). One load per block.
__device__ __forceinline__ u64 cand_record(const SectorView& S, u64 k, int j) {
    return S.cand_rec[(static_cast<u32>(j) << S.slc_L) + pattern_of(S, k, j)];
}

/// fn(g) for the elements that can map the state (key k) to its representative, ascending: those
/// whose image has the least leading block (sublattice_code.h:115-147).
template <class Fn> __device__ __forceinline__ void for_each_candidate(const SectorView& S, u64 k, Fn&& fn) {
    u64 rec[8];
    u32 best = 0xFFFFFFFFu;
    for (int j = 0; j < S.slc_m; ++j) {
        rec[j] = cand_record(S, k, j);
        const u32 l = static_cast<u32>(rec[j] & 0xFFFFu);
        best = l < best ? l : best;
    }
    // the blocks reaching the least leading block: a single element, or a list cand[pos, end)
    u32 pos[8], end[8], one[8];
    int n = 0;
    for (int j = 0; j < S.slc_m; ++j)
        if (static_cast<u32>(rec[j] & 0xFFFFu) == best) {
            const u32 cnt = static_cast<u32>((rec[j] >> 16) & 0xFFFFu), pay = static_cast<u32>(rec[j] >> 32);
            one[n] = cnt == 1 ? pay : 0xFFFFFFFFu;
            pos[n] = cnt == 1 ? 0 : pay;
            end[n] = cnt == 1 ? 1 : pay + cnt;
            ++n;
        }
    auto at = [&](int i) { return one[i] != 0xFFFFFFFFu ? one[i] : static_cast<u32>(S.cand[pos[i]]); };
    if (n == 1) {
        if (one[0] != 0xFFFFFFFFu) {
            fn(static_cast<int>(one[0]));
            return;
        }
        for (u32 p = pos[0]; p < end[0]; ++p) fn(static_cast<int>(S.cand[p]));
        return;
    }
    for (;;) {
        int pick = -1;
        u32 g = 0xFFFFFFFFu;
        for (int i = 0; i < n; ++i)
            if (pos[i] < end[i] && at(i) < g) {
                g = at(i);
                pick = i;
            }
        if (pick < 0) return;
        ++pos[pick];
        fn(static_cast<int>(g));
    }
}

/// Colex combinadic rank of a state with n_up set bits: the j-th set bit at position b adds C(b, j).
__device__ __forceinline__ i64 rank_of(const SectorView& S, u64 s) {
    i64 r = 0;
    int j = 0;
    while (s) {
        const int b = __ffsll(static_cast<long long>(s)) - 1;
        ++j;
        r += static_cast<i64>(S.binom[b * 65 + j]);
        s &= s - 1;
    }
    return r;
}

/// Index of a representative, or -1: the rank directory when present, else a binary search of reps.
__device__ __forceinline__ i64 index_of_rep(const SectorView& S, u64 rb) {
    if (S.rank_dir) {
        const u64 r = S.n_up >= 0 ? static_cast<u64>(rank_of(S, rb)) : rb;
        const RankRec e = S.rank_dir[r >> 6];
        const u64 bit = 1ULL << (r & 63);
        if (!(e.bits & bit)) return -1;
        return static_cast<i64>(e.base) + popc(e.bits & (bit - 1ULL));
    }
    i64 lo = 0, hi = S.dim;
    while (lo < hi) {
        const i64 mid = lo + (hi - lo) / 2;
        if (S.reps[mid] < rb) lo = mid + 1;
        else hi = mid;
    }
    return (lo < S.dim && S.reps[lo] == rb) ? lo : -1;
}

/// The orbit index of t and its projection conj(sum_{g: g t = rep} chi(g)) / norm, or -1
/// (rep_symmetry_basis_policy.h:221-241): a running minimum over the images in ascending element
/// order; the character sum restarts at a new minimum (0.0 + re, 0.0 - im) and grows on ties.
template <bool SLC>
__device__ __forceinline__ i64 index_and_projection(const SectorView& S, u64 t, double& pr, double& pi) {
    if (S.n_up >= 0 && popc(t) != S.n_up) return -1;
    u64 rb = ~0ULL, best = ~0ULL;
    double ar = 0.0, ai = 0.0;
    if constexpr (!SLC) {
        for (int g = 0; g < S.G; ++g) {
            const u64 img = apply_element(S, g, t);
            if (img < best) {
                best = img;
                rb = img;
                ar = 0.0 + S.chi_re[g];
                ai = 0.0 - S.chi_im[g];
            } else if (img == best) {
                ar += S.chi_re[g];
                ai -= S.chi_im[g];
            }
        }
    } else {
        const u64 k = key_of(S, t);
        for_each_candidate(S, k, [&](int g) {
            const u64 img = apply_element(S, g, t);
            const u64 ik = key_of(S, img);
            if (ik < best) {
                best = ik;
                rb = img;
                ar = 0.0 + S.chi_re[g];
                ai = 0.0 - S.chi_im[g];
            } else if (ik == best) {
                ar += S.chi_re[g];
                ai -= S.chi_im[g];
            }
        });
    }
    if (S.n_up >= 0 && popc(rb) != S.n_up) return -1;   // the group preserves popcount; never index past the table
    const i64 idx = index_of_rep(S, rb);
    if (idx < 0) return -1;
    const double w = S.inv_norms[idx];
    pr = ar * w;
    pi = ai * w;
    return idx;
}

/// Group g of the program on state s (row_walk.h:28-53): false when it does not act; else its
/// target t = s ^ F and h = sum of +-coeff of the subgroup whose value is s & F, in key order.
__device__ __forceinline__ bool connection(const ProgramView& P, u64 s, u32 g, u64& t, double& hr, double& hi) {
    const u64 F = P.group_flip[g];
    const u64 v = s & F;
    const i32 sb = P.group_setbits[g];
    if (sb >= 0 && popc(v) != sb) return false;
    u32 lo = P.group_vbegin[g], hi_ = P.group_vbegin[g + 1];
    const u32 stop = hi_;
    while (lo < hi_) {
        const u32 mid = lo + (hi_ - lo) / 2;
        if (P.vsub_val[mid] < v) lo = mid + 1;
        else hi_ = mid;
    }
    if (lo == stop || P.vsub_val[lo] != v) return false;
    hr = 0.0;
    hi = 0.0;
    for (u32 k = P.vsub_tbegin[lo]; k < P.vsub_tbegin[lo + 1]; ++k) {
        if (popc(s & P.term_sign[k]) & 1) {
            hr += -P.coeff_re[k];
            hi += -P.coeff_im[k];
        } else {
            hr += P.coeff_re[k];
            hi += P.coeff_im[k];
        }
    }
    t = s ^ F;
    return true;
}

/// Matrix entry of row r (state s, 1/norm w) for a connection (t, h) into the column sector:
/// diagonal conj(h) when the sectors are one and t == s, else w conj(h proj(t)) at index(t)
/// (sector_rows.h:41-56). Returns the column or -1. Written without fused multiply-adds so the
/// values carry the reference's bits.
template <bool SLC>
__device__ __forceinline__ i64 entry_of(const SectorView& C, bool same, i64 r, u64 s, double w, u64 t, double hr,
                                        double hi, double& vr, double& vi) {
    if (same && t == s) {
        vr = hr;
        vi = -hi;
        return r;
    }
    double pr, pi;
    const i64 c = index_and_projection<SLC>(C, t, pr, pi);
    if (c < 0) return -1;
    const double xr = hr * pr - hi * pi;
    const double xi = hr * pi + hi * pr;
    vr = w * xr;
    vi = w * (-xi);
    return c;
}

}  // namespace qsa::dev

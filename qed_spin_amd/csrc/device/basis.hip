// GPU-resident sector enumeration: the orbit scan of a fixed-Sz (or full) subspace into ascending
// representatives with closed-form norms, and the rank directory (rank -> index).
//
// The subspace is cut into 64-state words in rank (= ascending state) order. Pass 1 tests every
// state for being its orbit's representative and keeps those whose character sum over the
// stabiliser survives (a bit per state); an exclusive scan of the words' popcounts places them;
// pass 2 re-walks the words with a set bit and writes reps and 1/norm in order. The result is the
// reference's sector bit for bit: same representatives, same order, same norms
// (orbit_table.h:147-239, group_sector.cpp:159-205).
#include <hipcub/hipcub.hpp>

#include "dev_common.h"
#include "launch.h"

namespace qsa::dev {

namespace {

__device__ __forceinline__ u64 binom_at(const SectorView& S, int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    return S.binom[n * 65 + k];
}

/// The state of colex rank r with n_up set bits (combinadic.h:86-96); the state itself for the
/// full space.
__device__ u64 state_of_rank(const SectorView& S, u64 rank) {
    if (S.n_up < 0) return rank;
    // colex unranking: the i-th highest set bit is the largest p with C(p, i) <= the remaining rank;
    // the positions decrease, so p only moves down (at most n_sites + n_up table reads)
    u64 s = 0;
    int p = S.n_sites - 1;
    for (int i = S.n_up; i >= 1; --i) {
        while (binom_at(S, p, i) > rank) --p;
        s |= 1ULL << p;
        rank -= binom_at(S, p, i);
        --p;
    }
    return s;
}

__device__ __forceinline__ u64 next_state(const SectorView& S, u64 v) {
    if (S.n_up < 0) return v + 1;
    const u64 c = v & (~v + 1ULL);
    const u64 r = v + c;
    return (((r ^ v) >> 2) / c) | r;
}

/// Representative test (orbit_table.h:147-171): no image smaller, in the plain or key order.
template <bool SLC> __device__ bool is_rep(const SectorView& S, u64 s) {
    if constexpr (!SLC) {
        for (int g = 0; g < S.G; ++g)
            if (apply_element(S, g, s) < s) return false;
        return true;
    } else {
        const u64 k = key_of(S, s);
        u32 least = 0xFFFFFFFFu;
        for (int j = 0; j < S.slc_m; ++j) {
            const u32 l = static_cast<u32>(cand_record(S, k, j) & 0xFFFFu);
            least = l < least ? l : least;
        }
        if (least < pattern_of(S, k, 0)) return false;
        bool rep = true;
        for_each_candidate(S, k, [&](int g) {
            if (rep && key_of(S, apply_element(S, g, s)) < k) rep = false;
        });
        return rep;
    }
}

/// norm^2 = |sum_{h in Stab(s), ascending} chi(h)|^2 / |Stab|, 1 when |Stab| = 1 (orbit_table.h:93-99).
__device__ double norm_sq(const SectorView& S, u64 s) {
    double sr = 0.0, si = 0.0;
    int cnt = 0;
    for (int g = 0; g < S.G; ++g)
        if (apply_element(S, g, s) == s) {
            sr += S.chi_re[g];
            si += S.chi_im[g];
            ++cnt;
        }
    if (cnt == 1) return 1.0;
    const double a = sr * sr;
    const double b = si * si;
    return (a + b) / static_cast<double>(cnt);
}

template <bool SLC>
__global__ __launch_bounds__(256) void k_orbit_flags(SectorView S, u64 total, u64 n_words, u64* kept,
                                                     unsigned long long* orbits) {
    const u64 w = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long my_orbits = 0;
    if (w < n_words) {
        const u64 r0 = w * 64;
        const u64 r1 = (r0 + 64 < total) ? r0 + 64 : total;
        u64 s = state_of_rank(S, r0);
        u64 f = 0;
        for (u64 r = r0; r < r1; ++r) {
            if (is_rep<SLC>(S, s)) {
                ++my_orbits;
                if (norm_sq(S, s) > 1e-12) f |= 1ULL << (r - r0);
            }
            if (r + 1 < r1) s = next_state(S, s);
        }
        kept[w] = f;
    }
    // integer sum: any order gives the same count
    for (int off = 32; off > 0; off >>= 1) my_orbits += __shfl_down(my_orbits, off, 64);
    if ((threadIdx.x & 63) == 0 && my_orbits) atomicAdd(orbits, my_orbits);
}

// The same bitmap with one lane per state (a wave's ballot is one word): for subspaces small enough
// that one thread per 64 states leaves most of the device idle. The orbit count is summed per
// workgroup in LDS, one atomic per workgroup.
template <bool SLC>
__global__ __launch_bounds__(256) void k_orbit_flags_lane(SectorView S, u64 total, u64 n_words, u64* kept,
                                                          unsigned long long* orbits) {
    __shared__ unsigned long long wg_orbits;
    if (threadIdx.x == 0) wg_orbits = 0;
    __syncthreads();
    const u64 r = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool rep = false, keep = false;
    if (r < total) {
        const u64 s = state_of_rank(S, r);
        rep = is_rep<SLC>(S, s);
        keep = rep && norm_sq(S, s) > 1e-12;
    }
    const u64 fk = __ballot(keep);
    const u64 fr = __ballot(rep);
    const u64 w = r >> 6;
    if ((threadIdx.x & 63) == 0) {
        if (w < n_words) kept[w] = fk;
        if (fr) atomicAdd(&wg_orbits, static_cast<unsigned long long>(__popcll(fr)));
    }
    __syncthreads();
    if (threadIdx.x == 0 && wg_orbits) atomicAdd(orbits, wg_orbits);
}

__global__ void k_popcount_words(const u64* kept, u64 n_words, u64* counts) {
    const u64 w = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w < n_words) counts[w] = static_cast<u64>(__popcll(kept[w]));
}

__global__ __launch_bounds__(256) void k_orbit_emit(SectorView S, u64 total, u64 n_words, const u64* kept,
                                                    const u64* offsets, u64* reps, double* inv_norms) {
    const u64 w = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    const u64 f = kept[w];
    if (!f) return;
    u64 off = offsets[w];
    const u64 r0 = w * 64;
    const u64 r1 = (r0 + 64 < total) ? r0 + 64 : total;
    u64 s = state_of_rank(S, r0);
    for (u64 r = r0; r < r1; ++r) {
        if ((f >> (r - r0)) & 1ULL) {
            reps[off] = s;
            inv_norms[off] = 1.0 / sqrt(norm_sq(S, s));
            ++off;
        }
        if (r + 1 < r1) s = next_state(S, s);
    }
}

// One lane per state: a kept state's index is its word's offset plus the kept bits below it.
__global__ __launch_bounds__(256) void k_orbit_emit_lane(SectorView S, u64 total, const u64* kept, const u64* offsets,
                                                         u64* reps, double* inv_norms) {
    const u64 r = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= total) return;
    const u64 w = r >> 6;
    const int bit = static_cast<int>(r & 63);
    const u64 f = kept[w];
    if (!((f >> bit) & 1ULL)) return;
    const u64 off = offsets[w] + static_cast<u64>(__popcll(f & ((1ULL << bit) - 1ULL)));
    const u64 s = state_of_rank(S, r);
    reps[off] = s;
    inv_norms[off] = 1.0 / sqrt(norm_sq(S, s));
}

// The rank directory from the kept bitmap and its exclusive scan: record w = (word w, kept before it).
__global__ void k_rank_pack(const u64* kept, const u64* offsets, u64 n_words, RankRec* dir) {
    const u64 w = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    dir[w] = RankRec{kept[w], static_cast<u32>(offsets[w]), 0u};
}

// The bitmap of given representatives (an imported sector): the bit of each one's rank.
__global__ void k_rank_mark(SectorView S, const u64* reps, i64 n, unsigned long long* kept) {
    const i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 s = reps[i];
    const u64 r = S.n_up >= 0 ? static_cast<u64>(rank_of(S, s)) : s;
    atomicOr(kept + (r >> 6), 1ULL << (r & 63));
}

// A sector vector in the Sz sector (full: the 2^N space), before normalisation
// (src/engine/eigs.cpp:699-740): psi[index(g r_a)] += (u_a inv_norm_a) conj(chi(g)), g ascending.
// One thread per representative; orbits are disjoint, so each entry of psi has a single writer and
// the stabiliser's repeated images add up in element order. An image outside the Sz sector (a flip
// that changes the magnetisation) is not written; it raises *bad.
__global__ __launch_bounds__(256) void k_expand(SectorView S, const Zd* u, int full, Zd* psi, unsigned* bad) {
    const i64 a = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= S.dim) return;
    const Zd ua = u[a];
    if (ua.re == 0.0 && ua.im == 0.0) return;
    const double w = S.inv_norms[a];
    const double yr = ua.re * w, yi = ua.im * w;
    const u64 r = S.reps[a];
    for (int g = 0; g < S.G; ++g) {
        const u64 t = apply_element(S, g, r);
        if (!full && popc(t) != S.n_up) {
            *bad = 1u;
            continue;
        }
        const i64 at = full ? static_cast<i64>(t) : rank_of(S, t);
        const double cr = S.chi_re[g], ci = -S.chi_im[g];
        psi[at].re += yr * cr - yi * ci;
        psi[at].im += yr * ci + yi * cr;
    }
}

inline unsigned blocks_for(u64 n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

// Up to these many states of the subspace the orbit scan (flags) and the emission take one lane per
// state, above them one thread per 64 states stepping by Gosper's rule: an unranking per state costs
// more than a step once a thread per word fills the device (measured: at 4e7 states the flag pass
// takes 1.6 ms a word per thread and 2.5 ms a state per lane, the emission, which unranks only the
// kept states, 1.4 ms and 0.6 ms).
constexpr u64 kLanePerStateMax = u64{1} << 26;
constexpr u64 kLaneFlagsMax = u64{1} << 24;

}  // namespace

// The kept-representative bitmap and the orbit count of the subspace (before the character filter).
void orbit_flags(const SectorView& S, u64 total, u64* kept, unsigned long long* orbits, hipStream_t st) {
    const u64 n_words = (total + 63) / 64;
    if (total <= kLaneFlagsMax) {
        const dim3 grid(blocks_for(total, 256));
        if (S.slc_L > 0)
            hipLaunchKernelGGL(k_orbit_flags_lane<true>, grid, dim3(256), 0, st, S, total, n_words, kept, orbits);
        else
            hipLaunchKernelGGL(k_orbit_flags_lane<false>, grid, dim3(256), 0, st, S, total, n_words, kept, orbits);
        return;
    }
    if (S.slc_L > 0)
        hipLaunchKernelGGL(k_orbit_flags<true>, dim3(blocks_for(n_words, 256)), dim3(256), 0, st, S, total, n_words,
                           kept, orbits);
    else
        hipLaunchKernelGGL(k_orbit_flags<false>, dim3(blocks_for(n_words, 256)), dim3(256), 0, st, S, total,
                           n_words, kept, orbits);
}

std::size_t scan_words_temp_bytes(u64 n_words) {
    std::size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<u64*>(nullptr), static_cast<u64*>(nullptr),
                                     static_cast<int>(n_words));
    return bytes;
}

// offsets[w] = number of kept states in words before w; returns via *total_kept (device) the sum.
void scan_words(const u64* kept, u64 n_words, u64* counts, u64* offsets, void* temp, std::size_t temp_bytes,
                hipStream_t st) {
    hipLaunchKernelGGL(k_popcount_words, dim3(blocks_for(n_words, 256)), dim3(256), 0, st, kept, n_words, counts);
    hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, static_cast<int>(n_words), st);
}

void orbit_emit(const SectorView& S, u64 total, const u64* kept, const u64* offsets, u64* reps, double* inv_norms,
                hipStream_t st) {
    const u64 n_words = (total + 63) / 64;
    if (total <= kLanePerStateMax) {
        hipLaunchKernelGGL(k_orbit_emit_lane, dim3(blocks_for(total, 256)), dim3(256), 0, st, S, total, kept, offsets,
                           reps, inv_norms);
        return;
    }
    hipLaunchKernelGGL(k_orbit_emit, dim3(blocks_for(n_words, 256)), dim3(256), 0, st, S, total, n_words, kept,
                       offsets, reps, inv_norms);
}

void rank_pack(const u64* kept, const u64* offsets, u64 n_words, RankRec* dir, hipStream_t st) {
    if (n_words == 0) return;
    hipLaunchKernelGGL(k_rank_pack, dim3(blocks_for(n_words, 256)), dim3(256), 0, st, kept, offsets, n_words, dir);
}

void rank_mark(const SectorView& S, const u64* reps, i64 n, u64* kept, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rank_mark, dim3(blocks_for(static_cast<u64>(n), 256)), dim3(256), 0, st, S, reps, n,
                       reinterpret_cast<unsigned long long*>(kept));
}

void sector_expand(const SectorView& S, const Zd* u, bool full, Zd* psi, unsigned* bad, hipStream_t st) {
    if (S.dim <= 0) return;
    hipLaunchKernelGGL(k_expand, dim3(blocks_for(static_cast<u64>(S.dim), 256)), dim3(256), 0, st, S, u,
                       full ? 1 : 0, psi, bad);
}

}  // namespace qsa::dev

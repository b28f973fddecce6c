// =============================================================================
// oracle/oracle.cpp -- CPU restatement of QED_Spin's symmetry-sector matvec and Krylov path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load liboracle.so; nothing in the product (qed_spin_amd/) links,
// imports or calls it. Every function restates one reference function by its definition --
// scalar loops, brute force where the reference uses tables -- and cites the reference
// file:line it follows (paths relative to the reference checkout).
//
// Pins (what each part is checked against; see tests/test_oracle_*.py):
//   basis (rank/unrank, orbit scan, norms, key order)  -> oracle/_ref fixtures built from the
//        reference's own headers (bit-exact), brute force over 2^N, the Burnside sum rule.
//   index_and_projection, sector rows, merged CSR      -> oracle/_ref fixtures (bit-exact),
//        E^dag H E of the dense matrix built from term lists in numpy.
//   Lanczos / tridiagonal                             -> numpy eigh of the dense block,
//        the reference's golden eigs records.
//   FTLM moments + combine, seeds                     -> dense exact thermodynamics, the
//        reference's golden sampled-FTLM record (one 4096-state block, seed 11).
//   continued fraction                                -> closed form of a 1x1 / 2x2 tridiagonal,
//        the T=0 Lehmann sum of a dense ground state.
//
// Parity unpinned: orc_block_seed is pinned on the first sampled block of a recorded walk, but the
// ORDER in which a multi-block thermal walk hands seeds to its blocks, and the keys of the sources
// of a multi-source T > 0 walk, are restated from the reference's source text only: no recorded
// value of the reference pins the seed of a sampled block beyond the first.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <vector>

using u64 = std::uint64_t;
using i64 = std::int64_t;
using cplx = std::complex<double>;

namespace {

// ---- combinadic (include/ed/basis/combinadic.h:28-96) ---------------------------------------
u64 binom(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    u64 r = 1;
    for (int i = 1; i <= k; ++i) r = r * static_cast<u64>(n - k + i) / static_cast<u64>(i);
    return r;
}

// colex rank: the j-th set bit (1-based, low to high) at position b adds C(b, j)  (combinadic.h:70-83)
i64 rank_state(u64 s, int n_bits, int k) {
    i64 r = 0;
    int seen = 0;
    for (int b = 0; b < n_bits && seen < k; ++b)
        if ((s >> b) & 1ULL) {
            ++seen;
            r += static_cast<i64>(binom(b, seen));
        }
    return r;
}

// inverse of rank_state (combinadic.h:86-96)
u64 unrank_state(u64 rank, int n_bits, int k) {
    u64 s = 0;
    for (int i = k - 1; i >= 0; --i) {
        int p = i;
        while (p + 1 < n_bits && binom(p + 1, i + 1) <= rank) ++p;
        s |= 1ULL << p;
        rank -= binom(p, i + 1);
    }
    return s;
}

// Gosper (include/ed/basis/gosper.h:24-28)
u64 next_perm(u64 v) {
    const u64 c = v & (~v + 1ULL);
    const u64 r = v + c;
    return (((r ^ v) >> 2) / c) | r;
}

// element action by its definition: bit i of g s = bit perm[i] of s, then XOR flip
// (include/ed/basis/bits.h:32-36, compiled_group.h:73-78)
u64 act(const int* perm, u64 flip, int n, u64 s) {
    u64 r = 0;
    for (int i = 0; i < n; ++i) r |= ((s >> perm[i]) & 1ULL) << i;
    return r ^ flip;
}

// key order of a sublattice code: site s moves to key bit key_bit[s]
// (include/ed/basis/sublattice_code.h:98-103, 353-356)
u64 to_key(const int* key_bit, int n, u64 s) {
    if (!key_bit) return s;
    u64 k = 0;
    for (int i = 0; i < n; ++i)
        if ((s >> i) & 1ULL) k |= 1ULL << key_bit[i];
    return k;
}

struct Group {
    const int* perms;
    const u64* flips;
    int G, n;
    const int* key_bit;   // null: the plain order
    u64 img(int g, u64 s) const { return act(perms + static_cast<std::size_t>(g) * n, flips ? flips[g] : 0ULL, n, s); }
    u64 key(u64 s) const { return to_key(key_bit, n, s); }
};

}  // namespace

extern "C" {

u64 orc_binom(int n, int k) { return binom(n, k); }
i64 orc_rank_state(u64 s, int n_bits, int k) { return rank_state(s, n_bits, k); }
u64 orc_unrank_state(u64 r, int n_bits, int k) { return unrank_state(r, n_bits, k); }
u64 orc_next_bit_permutation(u64 v) { return next_perm(v); }
u64 orc_apply_element(const int* perm, u64 flip, int n, u64 s) { return act(perm, flip, n, s); }

// The block system and key order of the sublattice representative search
// (include/ed/basis/sublattice_code.h:234-283 block_system_, :286-312 build_ rule and key order).
// mode: 0 never, 1 whenever a block system exists, -1 the size rule (relaxed: 16 distinct
// permutations, else 64). Writes key_bit[N] and returns the block size L, or 0 (no code).
int orc_sublattice_key_bits(const int* perms, const u64* flips, int G, int N, int mode, int relaxed, int* key_bit) {
    (void)flips;
    if (G <= 1 || N < 4 || N > 64 || G > 65535 || mode == 0) return 0;
    std::vector<std::vector<int>> sp;
    for (int g = 0; g < G; ++g) {
        std::vector<int> p(perms + static_cast<std::size_t>(g) * N, perms + static_cast<std::size_t>(g + 1) * N);
        if (std::find(sp.begin(), sp.end(), p) == sp.end()) sp.push_back(p);
    }
    if (mode < 0) {
        const std::size_t min_perms = relaxed ? 16 : 64;
        if (N < 24 || sp.size() < min_perms) return 0;
    }
    int L = 0;
    std::vector<int> best;
    for (int b = 1; b < N; ++b) {
        // finest invariant partition joining sites 0 and b: union-find closed under every element
        std::vector<int> parent(N);
        std::iota(parent.begin(), parent.end(), 0);
        auto find = [&](int x) {
            while (parent[x] != x) x = parent[x];
            return x;
        };
        parent[b] = 0;
        std::vector<std::pair<int, int>> todo{{0, b}};
        while (!todo.empty()) {
            auto [x, y] = todo.back();
            todo.pop_back();
            for (const auto& p : sp) {
                const int px = p[x], py = p[y];
                const int u = find(px), v = find(py);
                if (u == v) continue;
                parent[std::max(u, v)] = std::min(u, v);
                todo.emplace_back(px, py);
            }
        }
        std::vector<int> size(N, 0);
        for (int s = 0; s < N; ++s) ++size[find(s)];
        const int l = size[0];
        if (l > 16 || l <= L || N % l != 0 || N / l < 2 || N / l > 8) continue;
        bool equal = true;
        for (int s = 0; s < N && equal; ++s) equal = size[find(s)] == l;
        if (!equal) continue;
        std::vector<int> root_block(N, -1), block(N);
        int nb = 0;
        for (int s = 0; s < N; ++s) {
            int& rb = root_block[find(s)];
            if (rb < 0) rb = nb++;
            block[s] = rb;
        }
        L = l;
        best = block;
    }
    if (best.empty()) return 0;
    const int m = N / L;
    std::vector<std::vector<int>> sites(m);
    for (int s = 0; s < N; ++s) sites[best[s]].push_back(s);
    for (int j = 0; j < m; ++j)
        for (int t = 0; t < L; ++t) key_bit[sites[j][t]] = N - (j + 1) * L + t;
    return L;
}

// ---- orbit scan + closed-form norms ----------------------------------------------------------
// For every state of the subspace, ascending: the representative test (no image smaller in the
// order) and the stabiliser {g : g s = s} (orbit_table.h:147-171 visit_state, :179-239 scan),
// then norm^2 = |sum_{h in Stab, ascending} chi(h)|^2 / |Stab| and 1/sqrt, kept if > 1e-12
// (orbit_table.h:93-99 projected_norm_sq_stab, src/engine/group_sector.cpp:159-205 filter_reps).
// n_up < 0: the full 2^N space. Returns the sector dimension (reps written up to `cap`);
// *n_orbits gets the number of orbits before the character filter.
i64 orc_sector_build(int N, int n_up, const int* perms, const u64* flips, int G, const double* chi_re,
                     const double* chi_im, const int* key_bit, u64* reps, double* inv_norms, i64 cap, i64* n_orbits) {
    Group grp{perms, flips, G, N, key_bit};
    i64 dim = 0, orbits = 0;
    const u64 total = n_up >= 0 ? binom(N, n_up) : (1ULL << N);
    u64 s = n_up >= 0 ? (n_up == 0 ? 0ULL : unrank_state(0, N, n_up)) : 0ULL;
    std::vector<int> stab;
    for (u64 i = 0; i < total; ++i) {
        const u64 ks = grp.key(s);
        bool rep = true;
        stab.clear();
        for (int g = 0; g < G && rep; ++g) {
            const u64 im = grp.img(g, s);
            const u64 ki = grp.key(im);
            if (ki < ks) rep = false;
            else if (im == s) stab.push_back(g);
        }
        if (G == 0) stab.push_back(0);
        if (rep) {
            ++orbits;
            double nsq = 1.0;
            if (stab.size() != 1 && G > 0) {
                cplx sum(0.0, 0.0);
                for (int h : stab) sum += cplx(chi_re[h], chi_im[h]);
                nsq = std::norm(sum) / static_cast<double>(stab.size());
            }
            const double inv = nsq <= 1e-12 ? 0.0 : 1.0 / std::sqrt(nsq);
            if (inv != 0.0) {
                if (dim < cap) {
                    reps[dim] = s;
                    inv_norms[dim] = inv;
                }
                ++dim;
            }
        }
        if (i + 1 < total) s = (n_up >= 0) ? next_perm(s) : s + 1;
    }
    if (n_orbits) *n_orbits = orbits;
    return dim;
}

// index_and_projection by its definition (include/ed/matvec/rep_symmetry_basis_policy.h:221-241):
// a running minimum over every image in ascending g (in key order with a sublattice code); on a
// new minimum the conjugate-character sum restarts (0.0 + re, 0.0 - im), on a tie it grows; the
// rep's index by binary search of the ascending reps; proj = acc * inv_norm. -1 if the popcount
// is off or the orbit is not in the sector.
i64 orc_index_and_projection(int N, int n_up, const int* perms, const u64* flips, int G, const double* chi_re,
                             const double* chi_im, const int* key_bit, const u64* reps, const double* inv_norms,
                             i64 dim, u64 t, double* proj_re, double* proj_im) {
    if (n_up >= 0 && __builtin_popcountll(t) != n_up) return -1;
    Group grp{perms, flips, G, N, key_bit};
    u64 rb = ~0ULL, best = ~0ULL;
    double ar = 0.0, ai = 0.0;
    for (int g = 0; g < G; ++g) {
        const u64 im = grp.img(g, t);
        const u64 k = grp.key(im);
        if (k < best) {
            best = k;
            rb = im;
            ar = 0.0 + chi_re[g];
            ai = 0.0 - chi_im[g];
        } else if (k == best) {
            ar += chi_re[g];
            ai -= chi_im[g];
        }
    }
    const u64* it = std::lower_bound(reps, reps + dim, rb);
    if (it == reps + dim || *it != rb) return -1;
    const i64 k = it - reps;
    *proj_re = ar * inv_norms[k];
    *proj_im = ai * inv_norms[k];
    return k;
}

// ---- row walk and sector rows ------------------------------------------------------------------
// A compiled program (include/ed/ops/program.h:51-94): groups by flip mask, value subgroups,
// terms. connection() (include/ed/ops/row_walk.h:28-53): v = s & F; popcount gate; binary search
// of v among the group's values; h = sum of +-coeff in key order from 0; t = s ^ F.
struct Program {
    int n_groups;
    const u64* group_flip;
    const int* group_setbits;
    const unsigned* group_vbegin;
    const u64* vsub_val;
    const unsigned* vsub_tbegin;
    const u64* term_sign;
    const double* coeff_re;
    const double* coeff_im;
};

static bool connection(const Program& P, u64 s, int g, u64& t, cplx& h) {
    const u64 F = P.group_flip[g];
    const u64 v = s & F;
    if (P.group_setbits[g] >= 0 && __builtin_popcountll(v) != P.group_setbits[g]) return false;
    unsigned lo = P.group_vbegin[g], hi = P.group_vbegin[g + 1];
    while (lo < hi) {
        const unsigned mid = lo + (hi - lo) / 2;
        if (P.vsub_val[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    if (lo == P.group_vbegin[g + 1] || P.vsub_val[lo] != v) return false;
    h = cplx(0.0, 0.0);
    for (unsigned k = P.vsub_tbegin[lo]; k < P.vsub_tbegin[lo + 1]; ++k) {
        const cplx c(P.coeff_re[k], P.coeff_im[k]);
        if (__builtin_popcountll(s & P.term_sign[k]) & 1) h += -c;
        else h += c;
    }
    t = s ^ F;
    return true;
}

struct SectorView {
    int N, n_up, G;
    const int* perms;
    const u64* flips;
    const double *chi_re, *chi_im;
    const int* key_bit;
    const u64* reps;
    const double* inv_norms;
    i64 dim;
    i64 index_proj(u64 t, cplx& proj) const {
        double pr = 0, pi = 0;
        const i64 k = orc_index_and_projection(N, n_up, perms, flips, G, chi_re, chi_im, key_bit, reps, inv_norms, dim,
                                               t, &pr, &pi);
        proj = cplx(pr, pi);
        return k;
    }
};

// Row j of O from the column sector to the row sector, merged: entries of the walk from the row's
// representative (include/ed/matvec/sector_rows.h:41-56 for_each_cross_entry: diagonal conj(h) at
// column j when the sectors are one and t == s, else inv_norm[j] * conj(h * proj(t)) at index(t)),
// then stable-sorted by column, equal columns summed in emission order, exact zeros dropped
// (sector_rows.h:147-158 merge_row).
static void merged_row(const Program& P, const SectorView& row, const SectorView& col, bool same, i64 j,
                       std::vector<std::pair<i64, cplx>>& out) {
    out.clear();
    const u64 s = row.reps[j];
    const double w = row.inv_norms[j];
    for (int g = 0; g < P.n_groups; ++g) {
        u64 t;
        cplx h;
        if (!connection(P, s, g, t, h)) continue;
        if (same && t == s) {
            out.emplace_back(j, std::conj(h));
            continue;
        }
        cplx proj;
        const i64 c = col.index_proj(t, proj);
        if (c < 0) continue;
        out.emplace_back(c, w * std::conj(h * proj));
    }
    std::stable_sort(out.begin(), out.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::size_t o = 0;
    for (std::size_t i = 0; i < out.size();) {
        cplx sum = out[i].second;
        std::size_t k = i + 1;
        for (; k < out.size() && out[k].first == out[i].first; ++k) sum += out[k].second;
        if (sum != cplx(0.0, 0.0)) out[o++] = {out[i].first, sum};
        i = k;
    }
    out.resize(o);
}

// The merged CSR of O between two sectors of one group (same: one sector). Fills row_ptr[rows+1]
// and, when col/val are non-null and cap suffices, the entries. Returns nnz.
i64 orc_sector_csr(int N, int row_nup, int col_nup, const int* perms, const u64* flips, int G, const double* rchi_re,
                   const double* rchi_im, const double* cchi_re, const double* cchi_im, const int* key_bit,
                   const u64* row_reps, const double* row_inv, i64 row_dim, const u64* col_reps, const double* col_inv,
                   i64 col_dim, int same, int n_groups, const u64* group_flip, const int* group_setbits,
                   const unsigned* group_vbegin, const u64* vsub_val, const unsigned* vsub_tbegin,
                   const u64* term_sign, const double* coeff_re, const double* coeff_im, i64* row_ptr, i64* col,
                   double* val_re, double* val_im, i64 cap) {
    Program P{n_groups, group_flip, group_setbits, group_vbegin, vsub_val, vsub_tbegin, term_sign, coeff_re, coeff_im};
    SectorView R{N, row_nup, G, perms, flips, rchi_re, rchi_im, key_bit, row_reps, row_inv, row_dim};
    SectorView C{N, col_nup, G, perms, flips, cchi_re, cchi_im, key_bit, col_reps, col_inv, col_dim};
    std::vector<std::pair<i64, cplx>> row;
    i64 nnz = 0;
    row_ptr[0] = 0;
    for (i64 j = 0; j < row_dim; ++j) {
        merged_row(P, R, C, same != 0, j, row);
        for (const auto& [c, v] : row) {
            if (col && nnz < cap) {
                col[nnz] = c;
                val_re[nnz] = v.real();
                val_im[nnz] = v.imag();
            }
            ++nnz;
        }
        row_ptr[j + 1] = nnz;
    }
    return nnz;
}

// y = A x with a CSR from orc_sector_csr (complex, interleaved re/im), row by row in entry order
// (include/ed/matvec/reduced_csr.h:77-103 spmv_with).
void orc_csr_matvec(i64 rows, const i64* row_ptr, const i64* col, const double* val_re, const double* val_im,
                    const double* x, double* y) {
    for (i64 r = 0; r < rows; ++r) {
        cplx acc(0.0, 0.0);
        for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
            acc += cplx(val_re[e], val_im[e]) * cplx(x[2 * col[e]], x[2 * col[e] + 1]);
        y[2 * r] = acc.real();
        y[2 * r + 1] = acc.imag();
    }
}

// ---- Lanczos (include/ed/krylov/lanczos.h:266-576) ----------------------------------------------
// The three-term recurrence with one of the reorthogonalisation policies: 0 None, 1 FullCGS2 with
// the DGKS-gated second pass (kappa = 1/sqrt 2), 2 LocalDGKS3 with K = 1 (project w on V_j when
// |<V_j, w>| > sqrt(eps), folded into the norm). Breakdown when beta < breakdown_tol. Sums are
// serial (a 1-thread CpuBackend). The operator is a CSR from orc_sector_csr. Returns steps done;
// alpha[steps], beta[steps + 1] (beta[0] = 0).
int orc_lanczos(i64 n, const i64* row_ptr, const i64* col, const double* val_re, const double* val_im,
                const double* v0, int max_iter, int reorth, double breakdown_tol, double* alpha, double* beta) {
    auto dot = [n](const std::vector<cplx>& a, const std::vector<cplx>& b) {
        cplx s(0.0, 0.0);
        for (i64 i = 0; i < n; ++i) s += std::conj(a[i]) * b[i];
        return s;
    };
    auto nrm = [n](const std::vector<cplx>& a) {
        double s = 0.0;
        for (i64 i = 0; i < n; ++i) s += std::norm(a[i]);
        return std::sqrt(s);
    };
    std::vector<cplx> vp(n), vc(n), w(n);
    for (i64 i = 0; i < n; ++i) vc[i] = cplx(v0[2 * i], v0[2 * i + 1]);
    const double n0 = nrm(vc);
    if (!(n0 > 0.0)) return -1;
    for (auto& x : vc) x *= cplx(1.0 / n0);
    std::vector<std::vector<cplx>> basis;
    if (reorth == 1) basis.push_back(vc);
    beta[0] = 0.0;
    const int cap = static_cast<int>(std::min<i64>(max_iter, n));
    int steps = 0;
    const double kappa = 0.7071067811865476, thr = 1.49011611938476562e-08;
    for (int j = 0; j < cap; ++j) {
        for (i64 r = 0; r < n; ++r) {
            cplx acc(0.0, 0.0);
            for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e) acc += cplx(val_re[e], val_im[e]) * vc[col[e]];
            w[r] = acc;
        }
        if (j > 0)
            for (i64 i = 0; i < n; ++i) w[i] += -beta[j] * vp[i];
        const double a = std::real(dot(vc, w));
        alpha[j] = a;
        for (i64 i = 0; i < n; ++i) w[i] += -a * vc[i];
        if (reorth == 1) {
            const double nb = nrm(w);
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<cplx> c(basis.size());
                for (std::size_t k = 0; k < basis.size(); ++k) c[k] = dot(basis[k], w);
                for (std::size_t k = 0; k < basis.size(); ++k)
                    for (i64 i = 0; i < n; ++i) w[i] -= c[k] * basis[k][i];
                if (pass == 0 && !(nrm(w) < kappa * nb)) break;
            }
        } else if (reorth == 2 && j > 0) {
            const cplx ov = dot(vc, w);
            if (std::abs(ov) > thr)
                for (i64 i = 0; i < n; ++i) w[i] += -ov * vc[i];
        }
        const double b = nrm(w);
        beta[j + 1] = b;
        steps = j + 1;
        if (b < breakdown_tol) break;
        for (i64 i = 0; i < n; ++i) w[i] *= cplx(1.0 / b);
        std::swap(vp, vc);
        std::swap(vc, w);
        if (j + 1 >= cap) break;
        if (reorth == 1) basis.push_back(vc);
    }
    return steps;
}

// Eigenvalues (ascending) and eigenvectors (column-major, column k of values[k]) of the leading
// m x m block of a Lanczos tridiagonal (include/ed/krylov/tridiag.h:38-57 convention: diagonal
// alpha, off-diagonal beta[1..m)). Cyclic Jacobi rotations on the dense matrix: slow and simple,
// an independent check of the product's implicit-QL solver. Returns 0 on success.
int orc_tridiag_eig(int m, const double* alpha, const double* beta, double* values, double* vectors) {
    std::vector<double> A(static_cast<std::size_t>(m) * m, 0.0), V(static_cast<std::size_t>(m) * m, 0.0);
    for (int i = 0; i < m; ++i) {
        A[i * m + i] = alpha[i];
        V[i * m + i] = 1.0;
        if (i + 1 < m) A[i * m + i + 1] = A[(i + 1) * m + i] = beta[i + 1];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) off += A[p * m + q] * A[p * m + q];
        if (off < 1e-300) break;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[p * m + q];
                if (std::abs(apq) < 1e-300) continue;
                const double theta = (A[q * m + q] - A[p * m + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = A[k * m + p], akq = A[k * m + q];
                    A[k * m + p] = c * akp - s * akq;
                    A[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = A[p * m + k], aqk = A[q * m + k];
                    A[p * m + k] = c * apk - s * aqk;
                    A[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = V[k * m + p], vkq = V[k * m + q];
                    V[k * m + p] = c * vkp - s * vkq;
                    V[k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return A[a * m + a] < A[b * m + b]; });
    for (int k = 0; k < m; ++k) {
        values[k] = A[order[k] * m + order[k]];
        if (vectors)
            for (int r = 0; r < m; ++r) vectors[k * m + r] = V[r * m + order[k]];
    }
    return 0;
}

// ---- FTLM (include/ed/thermal/ftlm.h:165-235) ---------------------------------------------------
// One sample's moments about its lowest weighted Ritz value (sample_moments), written as
// [e_min, Z[nT], E1[nT], E2[nT]].
void orc_ftlm_sample_moments(int m, const double* ritz, const double* weights, int nT, const double* betas,
                             double* out) {
    double e_min = std::numeric_limits<double>::infinity();
    for (int i = 0; i < m; ++i)
        if (weights[i] > 0.0) e_min = std::min(e_min, ritz[i]);
    if (!std::isfinite(e_min)) e_min = *std::min_element(ritz, ritz + m);
    out[0] = e_min;
    for (int t = 0; t < nT; ++t) {
        double z = 0.0, e1 = 0.0, e2 = 0.0;
        for (int i = 0; i < m; ++i) {
            const double x = ritz[i] - e_min;
            const double b = weights[i] * std::exp(-betas[t] * x);
            z += b;
            e1 += x * b;
            e2 += x * x * b;
        }
        out[1 + t] = z;
        out[1 + nT + t] = e1;
        out[1 + 2 * nT + t] = e2;
    }
}

// Jensen-correct combination of R samples' moments (combine_samples): lnZ, E, V per beta and e_ref.
void orc_ftlm_combine(int R, int nT, const double* moments, u64 n, const double* betas, double* lnZ, double* E,
                      double* V, double* e_ref_out) {
    const int stride = 1 + 3 * nT;
    double e_ref = moments[0];
    for (int s = 0; s < R; ++s) e_ref = std::min(e_ref, moments[s * stride]);
    *e_ref_out = e_ref;
    const double ln_n = std::log(static_cast<double>(n));
    for (int t = 0; t < nT; ++t) {
        const double beta = betas[t];
        double z = 0.0, e1 = 0.0, e2 = 0.0;
        for (int s = 0; s < R; ++s) {
            const double* ms = moments + static_cast<std::size_t>(s) * stride;
            const double Z = ms[1 + t], E1 = ms[1 + nT + t], E2 = ms[1 + 2 * nT + t];
            const double d = ms[0] - e_ref, r = std::exp(-beta * d);
            z += Z * r;
            e1 += (E1 + d * Z) * r;
            e2 += (E2 + 2.0 * d * E1 + d * d * Z) * r;
        }
        z /= static_cast<double>(R);
        e1 /= static_cast<double>(R);
        e2 /= static_cast<double>(R);
        if (z > 1e-300) {
            const double m1 = e1 / z, m2 = e2 / z;
            lnZ[t] = ln_n + std::log(z) - beta * e_ref;
            E[t] = e_ref + m1;
            V[t] = std::max(m2 - m1 * m1, 0.0);
        } else {
            lnZ[t] = -beta * e_ref;
            E[t] = e_ref;
            V[t] = 0.0;
        }
    }
}

// ---- seeds (include/ed/thermal/sample_seed.h:37-89) ---------------------------------------------
static u64 splitmix(u64 base, u64 idx) {
    u64 z = base + 0x9E3779B97F4A7C15ULL * (idx + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
u64 orc_block_seed(u64 base, u64 key) { return splitmix(base, key); }

// The unit complex Gaussian start of sample `sample`: std::mt19937 seeded with the low 32 bits of
// splitmix64(base + phi (sample + 1)); entries Complex(nd(gen), nd(gen)) of one
// std::normal_distribution<double>(0, 1); scaled by 1 / ||v|| (a plain sum of squares here; the
// reference's dznrm2 differs in the last bits). Interleaved re/im.
void orc_gaussian_vector(u64 base, u64 sample, i64 n, double* out) {
    std::mt19937 gen(static_cast<std::mt19937::result_type>(splitmix(base, sample)));
    std::normal_distribution<double> nd(0.0, 1.0);
    double s2 = 0.0;
    for (i64 i = 0; i < n; ++i) {
        const double im = nd(gen);   // GCC evaluates Complex(nd(gen), nd(gen))'s second argument first
        const double re = nd(gen);
        out[2 * i] = re;
        out[2 * i + 1] = im;
        s2 += re * re + im * im;
    }
    const double inv = 1.0 / std::sqrt(s2);
    for (i64 i = 0; i < 2 * n; ++i) out[i] *= inv;
}

// The same draw without normalisation (entries of std::normal_distribution as drawn).
void orc_gaussian_raw(u64 base, u64 sample, i64 n, double* out) {
    std::mt19937 gen(static_cast<std::mt19937::result_type>(splitmix(base, sample)));
    std::normal_distribution<double> nd(0.0, 1.0);
    for (i64 i = 0; i < n; ++i) {
        out[2 * i + 1] = nd(gen);    // the imaginary part is drawn first (see orc_gaussian_vector)
        out[2 * i] = nd(gen);
    }
}

// The draw restated from the published algorithms it is built of (no <random>), so the device
// generator's steps are checked on the CPU against the library:
//  - MT19937 (Matsumoto and Nishimura 1998): seeding x_0 = seed, x_i = 1812433253 (x_{i-1} xor
//    (x_{i-1} >> 30)) + i; the recurrence x_{k+624} = x_{k+397} xor twist(upper bit of x_k, lower 31
//    bits of x_{k+1}) with the matrix 0x9908b0df; tempering by 11 / 7 & 0x9d2c5680 / 15 & 0xefc60000 / 18;
//  - generate_canonical<double, 53> of a 32-bit engine (libstdc++ 11, bits/random.tcc:3348-3380):
//    u = (w0 + 2^32 w1) / 2^64, rounded once, 1 - 2^-53 when it reaches 1;
//  - Marsaglia's polar method (bits/random.tcc:1795-1835): x = 2 u - 1, y = 2 u' - 1 until
//    0 < r2 = x^2 + y^2 <= 1, then y m and x m (in that order) with m = sqrt(-2 ln r2 / r2).
struct OrcMt {
    std::uint32_t x[624];
    int i = 624;
    explicit OrcMt(std::uint32_t seed) {
        x[0] = seed;
        for (std::uint32_t k = 1; k < 624; ++k) x[k] = 1812433253u * (x[k - 1] ^ (x[k - 1] >> 30)) + k;
    }
    std::uint32_t next() {
        if (i == 624) {
            for (int k = 0; k < 624; ++k) {
                const std::uint32_t y = (x[k] & 0x80000000u) | (x[(k + 1) % 624] & 0x7fffffffu);
                x[k] = x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            i = 0;
        }
        std::uint32_t y = x[i++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    bool swap_words = false;   // a planted slip for the tests: the high word drawn first
    double canonical() {
        std::uint32_t lo = next();
        std::uint32_t hi = next();
        if (swap_words) std::swap(lo, hi);
        double sum = static_cast<double>(lo);
        sum = sum + static_cast<double>(hi) * 4294967296.0;
        const double r = sum / 18446744073709551616.0;
        return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
    }
};

// The k-th output (k >= 1) of MT19937 seeded with `seed`.
std::uint32_t orc_mt19937_nth(std::uint32_t seed, i64 k) {
    OrcMt g(seed);
    std::uint32_t y = 0;
    for (i64 j = 0; j < k; ++j) y = g.next();
    return y;
}

// n complex entries of the polar method on MT19937(seed), unnormalised; returns the candidates drawn.
// slip != 0 plants a mistake for the tests (1: words of a uniform swapped, 2: (x m, y m) per entry).
i64 orc_gaussian_polar(std::uint32_t seed, i64 n, double* out, int slip) {
    OrcMt g(seed);
    g.swap_words = slip == 1;
    i64 tried = 0;
    for (i64 i = 0; i < n; ++i) {
        double x, y, r2;
        do {
            x = 2.0 * g.canonical() - 1.0;
            y = 2.0 * g.canonical() - 1.0;
            r2 = x * x + y * y;
            ++tried;
        } while (r2 > 1.0 || r2 == 0.0);
        const double m = std::sqrt(-2.0 * std::log(r2) / r2);
        out[2 * i + 1] = (slip == 2 ? x : y) * m;   // the first normal of the pair: the imaginary part
        out[2 * i] = (slip == 2 ? y : x) * m;
    }
    return tried;
}

// ---- continued fraction (include/ed/dynamics/cf.h:39-59) --------------------------------------
void orc_continued_fraction(int M, const double* alpha, const double* beta, int nw, const double* omega, double eta,
                            double norm_sq, double* S) {
    const double kPi = 3.14159265358979323846;
    for (int iw = 0; iw < nw; ++iw) {
        const cplx z(omega[iw], eta);
        cplx G(0.0, 0.0);
        for (int n = M - 1; n >= 1; --n) {
            const double b2 = beta[n] * beta[n];
            const cplx d = z - cplx(alpha[n], 0.0) - G;
            G = std::abs(d) > 1e-300 ? cplx(b2, 0.0) / d : cplx(0.0, 0.0);
        }
        const cplx d = z - cplx(alpha[0], 0.0) - G;
        const cplx G0 = std::abs(d) > 1e-300 ? cplx(norm_sq, 0.0) / d : cplx(0.0, 0.0);
        S[iw] = -G0.imag() / kPi;
    }
}

}  // extern "C"

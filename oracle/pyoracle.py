"""ctypes wrapper of oracle/liboracle.so, plus the dense numpy oracle of term lists.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker. Nothing in qed_spin_amd/ imports it.

Parity unpinned: seed32 / gaussian_vector are pinned on the first sampled block of the reference's
recorded walks; the order in which a multi-block walk hands seeds to its blocks (and the source keys
of a multi-source T > 0 walk) is restated from the reference's source text, with no recorded value
beyond the first sampled block to pin it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

P_I = C.POINTER(C.c_int)
P_U64 = C.POINTER(C.c_uint64)
P_I64 = C.POINTER(C.c_int64)
P_D = C.POINTER(C.c_double)
P_U32 = C.POINTER(C.c_uint)


def build():
    subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "liboracle.so")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_binom.restype = C.c_uint64
        L.orc_binom.argtypes = [C.c_int, C.c_int]
        L.orc_rank_state.restype = C.c_int64
        L.orc_rank_state.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.orc_unrank_state.restype = C.c_uint64
        L.orc_unrank_state.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.orc_next_bit_permutation.restype = C.c_uint64
        L.orc_next_bit_permutation.argtypes = [C.c_uint64]
        L.orc_apply_element.restype = C.c_uint64
        L.orc_apply_element.argtypes = [P_I, C.c_uint64, C.c_int, C.c_uint64]
        L.orc_sublattice_key_bits.restype = C.c_int
        L.orc_sublattice_key_bits.argtypes = [P_I, P_U64, C.c_int, C.c_int, C.c_int, C.c_int, P_I]
        L.orc_sector_build.restype = C.c_int64
        L.orc_sector_build.argtypes = [C.c_int, C.c_int, P_I, P_U64, C.c_int, P_D, P_D, P_I, P_U64, P_D, C.c_int64,
                                       P_I64]
        L.orc_index_and_projection.restype = C.c_int64
        L.orc_index_and_projection.argtypes = [C.c_int, C.c_int, P_I, P_U64, C.c_int, P_D, P_D, P_I, P_U64, P_D,
                                               C.c_int64, C.c_uint64, P_D, P_D]
        L.orc_sector_csr.restype = C.c_int64
        L.orc_sector_csr.argtypes = [C.c_int, C.c_int, C.c_int, P_I, P_U64, C.c_int, P_D, P_D, P_D, P_D, P_I,
                                     P_U64, P_D, C.c_int64, P_U64, P_D, C.c_int64, C.c_int, C.c_int, P_U64, P_I,
                                     P_U32, P_U64, P_U32, P_U64, P_D, P_D, P_I64, P_I64, P_D, P_D, C.c_int64]
        L.orc_csr_matvec.restype = None
        L.orc_csr_matvec.argtypes = [C.c_int64, P_I64, P_I64, P_D, P_D, P_D, P_D]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_int64, P_I64, P_I64, P_D, P_D, P_D, C.c_int, C.c_int, C.c_double, P_D, P_D]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_tridiag_eig.argtypes = [C.c_int, P_D, P_D, P_D, P_D]
        L.orc_ftlm_sample_moments.restype = None
        L.orc_ftlm_sample_moments.argtypes = [C.c_int, P_D, P_D, C.c_int, P_D, P_D]
        L.orc_ftlm_combine.restype = None
        L.orc_ftlm_combine.argtypes = [C.c_int, C.c_int, P_D, C.c_uint64, P_D, P_D, P_D, P_D, P_D]
        L.orc_block_seed.restype = C.c_uint64
        L.orc_block_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_gaussian_vector.restype = None
        L.orc_gaussian_vector.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, P_D]
        L.orc_gaussian_raw.restype = None
        L.orc_gaussian_raw.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, P_D]
        L.orc_mt19937_nth.restype = C.c_uint32
        L.orc_mt19937_nth.argtypes = [C.c_uint32, C.c_int64]
        L.orc_gaussian_polar.restype = C.c_int64
        L.orc_gaussian_polar.argtypes = [C.c_uint32, C.c_int64, P_D, C.c_int]
        L.orc_continued_fraction.restype = None
        L.orc_continued_fraction.argtypes = [C.c_int, P_D, P_D, C.c_int, P_D, C.c_double, C.c_double, P_D]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class OracleSector:
    """A sector by brute force: the orbit scan of every state in the subspace."""

    def __init__(self, N, n_up, perms, chars, flips=None, key_bit=None):
        self.N, self.n_up = int(N), int(n_up)
        self.perms = np.ascontiguousarray(perms, np.int32)
        self.G = int(self.perms.shape[0])
        self.flips = np.ascontiguousarray(np.zeros(self.G, np.uint64) if flips is None else flips, np.uint64)
        ch = np.asarray(chars, np.complex128)
        self.chi_re = np.ascontiguousarray(ch.real)
        self.chi_im = np.ascontiguousarray(ch.imag)
        self.key_bit = None if key_bit is None else np.ascontiguousarray(key_bit, np.int32)
        L = lib()
        total = int(L.orc_binom(self.N, self.n_up)) if self.n_up >= 0 else (1 << self.N)
        cap = total
        self.reps = np.zeros(cap, np.uint64)
        self.inv_norms = np.zeros(cap, np.float64)
        orbits = C.c_int64()
        dim = L.orc_sector_build(self.N, self.n_up, _p(self.perms, C.c_int), _p(self.flips, C.c_uint64), self.G,
                                 _p(self.chi_re, C.c_double), _p(self.chi_im, C.c_double), _p(self.key_bit, C.c_int),
                                 _p(self.reps, C.c_uint64), _p(self.inv_norms, C.c_double), cap, C.byref(orbits))
        self.dim = int(dim)
        self.n_orbits = int(orbits.value)
        self.reps = self.reps[: self.dim].copy()
        self.inv_norms = self.inv_norms[: self.dim].copy()

    def index_and_projection(self, t):
        pr, pi = C.c_double(), C.c_double()
        k = lib().orc_index_and_projection(self.N, self.n_up, _p(self.perms, C.c_int), _p(self.flips, C.c_uint64),
                                           self.G, _p(self.chi_re, C.c_double), _p(self.chi_im, C.c_double),
                                           _p(self.key_bit, C.c_int), _p(self.reps, C.c_uint64),
                                           _p(self.inv_norms, C.c_double), self.dim, int(t), C.byref(pr),
                                           C.byref(pi))
        return int(k), complex(pr.value, pi.value)


def sublattice_key_bits(perms, flips, N, mode, relaxed=True):
    perms = np.ascontiguousarray(perms, np.int32)
    G = perms.shape[0]
    flips = np.ascontiguousarray(np.zeros(G, np.uint64) if flips is None else flips, np.uint64)
    kb = np.zeros(N, np.int32)
    L = lib().orc_sublattice_key_bits(_p(perms, C.c_int), _p(flips, C.c_uint64), G, N, mode, int(relaxed),
                                      _p(kb, C.c_int))
    return (kb if L else None), L


def sector_csr(prog, rows: OracleSector, cols: OracleSector = None):
    """The merged CSR of the program between two oracle sectors of one group (cols None: one)."""
    same = cols is None
    cols = rows if cols is None else cols
    L = lib()
    row_ptr = np.zeros(rows.dim + 1, np.int64)
    args_head = (rows.N, rows.n_up, cols.n_up, _p(rows.perms, C.c_int), _p(rows.flips, C.c_uint64), rows.G,
                 _p(rows.chi_re, C.c_double), _p(rows.chi_im, C.c_double), _p(cols.chi_re, C.c_double),
                 _p(cols.chi_im, C.c_double), _p(rows.key_bit, C.c_int), _p(rows.reps, C.c_uint64),
                 _p(rows.inv_norms, C.c_double), rows.dim, _p(cols.reps, C.c_uint64), _p(cols.inv_norms, C.c_double),
                 cols.dim, int(same))
    pa = {k: np.ascontiguousarray(v) for k, v in prog.items()}
    coeff = pa["term_coeff"].reshape(-1, 2)
    cre = np.ascontiguousarray(coeff[:, 0])
    cim = np.ascontiguousarray(coeff[:, 1])
    args_prog = (len(pa["group_flip"]), _p(pa["group_flip"], C.c_uint64), _p(pa["group_setbits"], C.c_int),
                 _p(pa["group_vbegin"], C.c_uint), _p(pa["vsub_val"], C.c_uint64), _p(pa["vsub_tbegin"], C.c_uint),
                 _p(pa["term_sign"], C.c_uint64), _p(cre, C.c_double), _p(cim, C.c_double))
    nnz = L.orc_sector_csr(*args_head, *args_prog, _p(row_ptr, C.c_int64), None, None, None, 0)
    col = np.zeros(nnz, np.int64)
    vre = np.zeros(nnz)
    vim = np.zeros(nnz)
    L.orc_sector_csr(*args_head, *args_prog, _p(row_ptr, C.c_int64), _p(col, C.c_int64), _p(vre, C.c_double),
                     _p(vim, C.c_double), nnz)
    val = np.empty(nnz, np.complex128)   # component-wise: keeps the sign of zero parts
    val.real = vre
    val.imag = vim
    return row_ptr, col, val


def csr_matvec(row_ptr, col, val, x):
    x = np.ascontiguousarray(x, np.complex128)
    y = np.zeros(len(row_ptr) - 1, np.complex128)
    vre = np.ascontiguousarray(val.real)
    vim = np.ascontiguousarray(val.imag)
    lib().orc_csr_matvec(len(row_ptr) - 1, _p(np.ascontiguousarray(row_ptr, np.int64), C.c_int64),
                         _p(np.ascontiguousarray(col, np.int64), C.c_int64), _p(vre, C.c_double),
                         _p(vim, C.c_double), _p(x.view(np.float64), C.c_double), _p(y.view(np.float64), C.c_double))
    return y


def lanczos(row_ptr, col, val, v0, max_iter, reorth=1, breakdown_tol=1e-300):
    n = len(row_ptr) - 1
    v0 = np.ascontiguousarray(v0, np.complex128)
    alpha = np.zeros(max_iter)
    beta = np.zeros(max_iter + 1)
    vre = np.ascontiguousarray(val.real)
    vim = np.ascontiguousarray(val.imag)
    m = lib().orc_lanczos(n, _p(np.ascontiguousarray(row_ptr, np.int64), C.c_int64),
                          _p(np.ascontiguousarray(col, np.int64), C.c_int64), _p(vre, C.c_double),
                          _p(vim, C.c_double), _p(v0.view(np.float64), C.c_double), int(max_iter), int(reorth),
                          float(breakdown_tol), _p(alpha, C.c_double), _p(beta, C.c_double))
    return alpha[:m], beta[: m + 1]


def tridiag_eig(alpha, beta):
    m = len(alpha)
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    w = np.zeros(m)
    v = np.zeros(m * m)
    lib().orc_tridiag_eig(m, _p(a, C.c_double), _p(b, C.c_double), _p(w, C.c_double), _p(v, C.c_double))
    return w, v.reshape(m, m).T   # columns are eigenvectors


def ftlm_moments(ritz, weights, betas):
    ritz = np.ascontiguousarray(ritz, np.float64)
    weights = np.ascontiguousarray(weights, np.float64)
    betas = np.ascontiguousarray(betas, np.float64)
    out = np.zeros(1 + 3 * len(betas))
    lib().orc_ftlm_sample_moments(len(ritz), _p(ritz, C.c_double), _p(weights, C.c_double), len(betas),
                                  _p(betas, C.c_double), _p(out, C.c_double))
    return out


def ftlm_combine(moments, n, betas):
    moments = np.ascontiguousarray(moments, np.float64)
    betas = np.ascontiguousarray(betas, np.float64)
    R = moments.shape[0]
    nT = len(betas)
    lnZ, E, V = np.zeros(nT), np.zeros(nT), np.zeros(nT)
    eref = C.c_double()
    lib().orc_ftlm_combine(R, nT, _p(moments, C.c_double), int(n), _p(betas, C.c_double), _p(lnZ, C.c_double),
                           _p(E, C.c_double), _p(V, C.c_double), C.byref(eref))
    return lnZ, E, V, eref.value


def gaussian_vector(base, sample, n):
    out = np.zeros(2 * n)
    lib().orc_gaussian_vector(int(base), int(sample), int(n), _p(out, C.c_double))
    return out.view(np.complex128)


def gaussian_raw(base, sample, n):
    """Sample `sample` of seed `base` as std::normal_distribution draws it (not normalised)."""
    out = np.zeros(2 * n)
    lib().orc_gaussian_raw(int(base), int(sample), int(n), _p(out, C.c_double))
    return out.view(np.complex128)


def mt19937_nth(seed, k):
    return int(lib().orc_mt19937_nth(int(seed) & 0xFFFFFFFF, int(k)))


def gaussian_polar(seed32, n, slip=0):
    """The restated MT19937 -> polar-method draw; returns (entries, candidates drawn). slip plants a
    mistake (1: the words of a uniform swapped, 2: (x m, y m) per entry) for the tests."""
    out = np.zeros(2 * n)
    tried = lib().orc_gaussian_polar(int(seed32) & 0xFFFFFFFF, int(n), _p(out, C.c_double), int(slip))
    return out.view(np.complex128), int(tried)


def seed32(base, sample):
    """The 32-bit Mersenne Twister seed of sample `sample` of seed `base`."""
    return int(lib().orc_block_seed(int(base), int(sample))) & 0xFFFFFFFF


def continued_fraction(alpha, beta, omega, eta, norm_sq):
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    w = np.ascontiguousarray(omega, np.float64)
    S = np.zeros(len(w))
    lib().orc_continued_fraction(len(a), _p(a, C.c_double), _p(b, C.c_double), len(w), _p(w, C.c_double), float(eta),
                                 float(norm_sq), _p(S, C.c_double))
    return S


# ---- the dense oracle of term lists (independent of every lane) ------------------------------------
def _apply_product(ops, sites, s):
    amp = 1.0 + 0j
    for op, site in zip(reversed(ops), reversed(sites)):
        bit = (s >> site) & 1
        if op == "I":
            continue
        if op == "z":
            amp *= 0.5 if bit else -0.5
        elif op == "u":
            if not bit:
                return 0.0, s
        elif op == "d":
            if bit:
                return 0.0, s
        elif op == "+":
            if bit:
                return 0.0, s
            s ^= 1 << site
        elif op == "-":
            if not bit:
                return 0.0, s
            s ^= 1 << site
        else:
            raise ValueError(op)
    return amp, s


def _expand_xy(ops, sites, coeff):
    out = [(complex(coeff), "", ())]
    for o, s in zip(ops, sites):
        parts = {"x": [(0.5, "+"), (0.5, "-")], "y": [(-0.5j, "+"), (0.5j, "-")]}.get(o, [(1.0, o)])
        out = [(c * pc, prev_o + po, prev_s + (s,)) for c, prev_o, prev_s in out for pc, po in parts]
    return out


def dense_matrix(products, N, states=None):
    """M[t, s] = <t|O|s> over `states` (default all 2^N) from (ops, sites, coeff) products."""
    if states is None:
        states = np.arange(1 << N, dtype=np.int64)
    index = {int(s): i for i, s in enumerate(states)}
    M = np.zeros((len(states), len(states)), np.complex128)
    for ops, sites, coeff in products:
        for c, o2, s2 in _expand_xy(ops, sites, coeff):
            for j, s in enumerate(states):
                a, t = _apply_product(o2, s2, int(s))
                if a != 0 and t in index:
                    M[index[t], j] += c * a
    return M


def sector_basis_matrix(sector: OracleSector, states):
    """E (len(states) x dim): column r = |G|^{-1/2} inv_norm_r sum_g conj(chi(g)) |g s_r>."""
    index = {int(s): i for i, s in enumerate(states)}
    E = np.zeros((len(states), sector.dim), np.complex128)
    L = lib()
    chi = sector.chi_re + 1j * sector.chi_im
    for r, s in enumerate(sector.reps):
        for g in range(sector.G):
            t = L.orc_apply_element(_p(np.ascontiguousarray(sector.perms[g]), C.c_int), int(sector.flips[g]),
                                    sector.N, int(s))
            E[index[int(t)], r] += np.conj(chi[g]) * sector.inv_norms[r]
    return E / np.sqrt(sector.G)


def irrep_fields(N, n_up, perms, D, flips=None):
    """A sector of an irrep of dimension d > 1 by brute force, from the definitions of the
    reference's d > 1 sectors (include/ed/basis/rep_sector.h:198-215, src/engine/group_sector.cpp:
    78-139): every orbit's least member r, its stabiliser S = {g : g r = r}, the projector
    P = (1/|S|) sum_{s in S} conj(D(s)), C_r = an orthonormal basis of P's range (numpy's eigh,
    eigenvalue 1) over sqrt(|S|); members of rank 0 dropped. Returns (reps, [C_r as d x rank])."""
    perms = np.ascontiguousarray(perms, np.int32)
    G = perms.shape[0]
    D = np.asarray(D, np.complex128)
    fl = np.zeros(G, np.uint64) if flips is None else np.asarray(flips, np.uint64)
    triv = OracleSector(N, n_up, perms, np.ones(G), fl)
    L = lib()
    reps, Cs = [], []
    for r in triv.reps:
        S = [g for g in range(G)
             if L.orc_apply_element(_p(np.ascontiguousarray(perms[g]), C.c_int), int(fl[g]), N, int(r)) == int(r)]
        P = np.conj(D[S]).sum(axis=0) / len(S)
        w, V = np.linalg.eigh(0.5 * (P + P.conj().T))
        cols = V[:, w > 0.5]
        if cols.shape[1] == 0:
            continue
        reps.append(int(r))
        Cs.append(cols / np.sqrt(len(S)))
    return np.array(reps, np.uint64), Cs


def irrep_basis_matrix(N, perms, D, reps, Cs, states, flips=None):
    """E (len(states) x states of the sector): the columns |r; a> = sqrt(d/|G|) sum_j C_r[j][a]
    sum_g conj(D(g)[0][j]) |g r>, representative by representative, a < rank(r)."""
    perms = np.ascontiguousarray(perms, np.int32)
    G = perms.shape[0]
    D = np.asarray(D, np.complex128)
    d = D.shape[1]
    fl = np.zeros(G, np.uint64) if flips is None else np.asarray(flips, np.uint64)
    index = {int(s): i for i, s in enumerate(states)}
    n = sum(Cr.shape[1] for Cr in Cs)
    E = np.zeros((len(states), n), np.complex128)
    L = lib()
    col = 0
    for r, Cr in zip(reps, Cs):
        imgs = [index[int(L.orc_apply_element(_p(np.ascontiguousarray(perms[g]), C.c_int), int(fl[g]), N, int(r)))]
                for g in range(G)]
        for a in range(Cr.shape[1]):
            coef = np.conj(D[:, 0, :]) @ Cr[:, a]   # per element: sum_j conj(D(g)[0][j]) C_r[j][a]
            for g in range(G):
                E[imgs[g], col] += coef[g]
            col += 1
    return E * np.sqrt(d / G)

"""Pins of the oracle (CPU): against the reference's own hot path (fixtures written by oracle/_ref),
brute force, the Burnside count and closed forms. A plausible slip -- a dropped character, a wrong
sign in the projection, a transposed permutation, a column merge in the wrong order -- fails here."""

import math
import os

import numpy as np
import pytest

from conftest import FIXTURES, fixture_names, fixture_program, fixture_products, load_fixture
from oracle import pyoracle as po

NAMES = fixture_names()


def oracle_sector(fx, key_bit=None):
    kb = fx["key_bit"] if "key_bit" in fx else key_bit
    return po.OracleSector(int(fx["N"]), int(fx["n_up"]), fx["perms"], fx["chars"], fx["flips"], kb)


def test_fixtures_present():
    assert len(NAMES) >= 10


@pytest.mark.parametrize("name", NAMES)
def test_sector_bit_exact(name):
    fx = load_fixture(name)
    s = oracle_sector(fx)
    assert s.n_orbits == int(fx["n_orbits"][0])
    assert s.dim == len(fx["reps"])
    assert np.array_equal(s.reps, fx["reps"])
    assert np.array_equal(s.inv_norms.view(np.uint64), fx["inv_norms"].view(np.uint64))


@pytest.mark.parametrize("name", NAMES)
def test_key_order_bit_exact(name):
    fx = load_fixture(name)
    mode = int(fx["sublattice_mode"])
    kb, L = po.sublattice_key_bits(fx["perms"], fx["flips"], int(fx["N"]), mode)
    if "key_bit" in fx:
        assert L > 0 and np.array_equal(kb, fx["key_bit"])
    else:
        assert L == 0


@pytest.mark.parametrize("name", NAMES)
def test_index_and_projection_bit_exact(name):
    fx = load_fixture(name)
    s = oracle_sector(fx)
    proj = fx["probe_proj"].view(np.complex128)
    for t, k_ref, p_ref in zip(fx["probes"], fx["probe_index"], proj):
        k, p = s.index_and_projection(int(t))
        assert k == k_ref
        if k >= 0:
            assert np.float64(p.real).view(np.uint64) == np.float64(p_ref.real).view(np.uint64)
            assert np.float64(p.imag).view(np.uint64) == np.float64(p_ref.imag).view(np.uint64)


@pytest.mark.parametrize("name", NAMES)
def test_merged_csr_bit_exact(name):
    fx = load_fixture(name)
    s = oracle_sector(fx)
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    assert np.array_equal(row_ptr.astype(np.uint64), fx["csr_row_ptr"])
    assert np.array_equal(col.astype(np.uint32), fx["csr_col"])
    ref = fx["csr_val"].view(np.complex128)
    assert np.array_equal(val.view(np.float64).view(np.uint64), ref.view(np.float64).view(np.uint64))
    assert int(fx["csr_meta"][0]) == len(col)


@pytest.mark.parametrize("name", NAMES)
def test_matvec_matches_reference_walk(name):
    fx = load_fixture(name)
    s = oracle_sector(fx)
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    x = fx["gather_in"].view(np.complex128)
    y = po.csr_matvec(row_ptr, col, val, x)
    ref = fx["gather_out"].view(np.complex128)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("name", ["j1j2_n12_k3", "j1j2_n12_d_k0_odd", "heis_n12_flip_k6", "chiral_n10_k2",
                                  "xxz_n10_nup4_k1", "heis_n8_full_k1"])
def test_sector_block_equals_projected_dense(name):
    """E^dag H E of the dense matrix (terms only, no lane) equals the sector CSR."""
    fx = load_fixture(name)
    N, n_up = int(fx["N"]), int(fx["n_up"])
    s = oracle_sector(fx)
    states = [x for x in range(1 << N) if n_up < 0 or bin(x).count("1") == n_up]
    H = po.dense_matrix(fixture_products(fx), N, states)
    E = po.sector_basis_matrix(s, states)
    assert np.allclose(E.conj().T @ E, np.eye(s.dim), atol=1e-12)
    Hs = E.conj().T @ H @ E
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    A = np.zeros((s.dim, s.dim), np.complex128)
    for r in range(s.dim):
        for e in range(row_ptr[r], row_ptr[r + 1]):
            A[r, col[e]] += val[e]
    scale = np.abs(H).max()
    assert np.abs(A - Hs).max() <= 1e-12 * scale


def test_burnside_tiles_the_sz_sector():
    """Momentum sectors of the translations partition C(N, n_up) (the sum rule)."""
    N, n_up = 12, 6
    perms = np.array([[(i + a) % N for i in range(N)] for a in range(N)], np.int32)
    total = 0
    for m in range(N):
        chars = np.exp(2j * np.pi * m * np.arange(N) / N)
        total += po.OracleSector(N, n_up, perms, chars).dim
    assert total == math.comb(N, n_up)


def test_rank_unrank_roundtrip():
    L = po.lib()
    N, k = 16, 7
    s = L.orc_unrank_state(0, N, k)
    for r in range(math.comb(N, k)):
        assert L.orc_rank_state(s, N, k) == r
        assert L.orc_unrank_state(r, N, k) == s
        if r + 1 < math.comb(N, k):
            s = L.orc_next_bit_permutation(s)


def test_key_order_planted_error_is_caught():
    """A transposed key order (blocks swapped) changes the representatives of a fixture: the fixture
    comparison is sensitive to the order it pins."""
    fx = load_fixture("j1j2_n16_k3_slc")
    kb = fx["key_bit"].copy()
    N = int(fx["N"])
    swapped = (kb + N // 2) % N
    s = po.OracleSector(N, int(fx["n_up"]), fx["perms"], fx["chars"], fx["flips"], swapped)
    assert s.dim == len(fx["reps"]) and not np.array_equal(s.reps, fx["reps"])


def dense_csr(row_ptr, col, val, n):
    A = np.zeros((n, n), np.complex128)
    for r in range(n):
        for e in range(row_ptr[r], row_ptr[r + 1]):
            A[r, col[e]] += val[e]
    return A


@pytest.mark.parametrize("reorth", [0, 1, 2])
def test_lanczos_against_dense(reorth):
    fx = load_fixture("j1j2_n16_k3_plain")
    s = oracle_sector(fx)
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    A = dense_csr(row_ptr, col, val, s.dim)
    w = np.linalg.eigvalsh(A)
    rng = np.random.default_rng(1)
    v0 = rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim)
    alpha, beta = po.lanczos(row_ptr, col, val, v0, 120, reorth)
    ritz, _ = po.tridiag_eig(alpha, beta)
    assert abs(ritz[0] - w[0]) <= 1e-10 * abs(w[0])
    if reorth == 1:   # a full basis: T = V^H A V, extremes converge on both ends
        assert abs(ritz[-1] - w[-1]) <= 1e-10 * abs(w[-1])


def test_tridiag_eig_against_numpy():
    rng = np.random.default_rng(3)
    a = rng.normal(size=40)
    b = np.concatenate([[0.0], rng.normal(size=39) ** 2 + 0.1])
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    w, V = po.tridiag_eig(a, b)
    assert np.allclose(w, np.linalg.eigvalsh(T), atol=1e-12)
    assert np.allclose(T @ V, V * w, atol=1e-10)


def test_ftlm_full_trace_is_exact():
    """With start vectors spanning the space (the unit vectors, one Lanczos each run to the full
    dimension) FTLM's moments sum to the exact partition function of a small block."""
    fx = load_fixture("xxz_n10_nup4_k1")
    s = oracle_sector(fx)
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    A = dense_csr(row_ptr, col, val, s.dim)
    ev = np.linalg.eigvalsh(A)
    betas = np.array([0.3, 1.0, 2.5])
    moments = []
    for i in range(s.dim):
        v0 = np.zeros(s.dim, np.complex128)
        v0[i] = 1.0
        alpha, beta = po.lanczos(row_ptr, col, val, v0, s.dim, 1, 1e-13)
        ritz, V = po.tridiag_eig(alpha, beta)
        moments.append(po.ftlm_moments(ritz, V[0] ** 2, betas))
    # R samples average to Z / R; scaled by n = dim inside combine, so use n = R = dim
    lnZ, E, Vv, eref = po.ftlm_combine(np.array(moments), s.dim, betas)
    Z = np.array([np.exp(-b * ev).sum() for b in betas])
    Em = np.array([(ev * np.exp(-b * ev)).sum() for b in betas]) / Z
    assert np.allclose(lnZ, np.log(Z), rtol=1e-10, atol=1e-10)
    assert np.allclose(E, Em, rtol=1e-10, atol=1e-10)


def test_continued_fraction_two_level_closed_form():
    """A 2x2 tridiagonal: G(z) = (z - a1) / ((z - a0)(z - a1) - b1^2)."""
    a = np.array([0.3, 1.7])
    b = np.array([0.0, 0.8])
    w = np.linspace(-2, 4, 31)
    eta = 0.07
    S = po.continued_fraction(a, b, w, eta, 2.5)
    z = w + 1j * eta
    G = 2.5 * (z - a[1]) / ((z - a[0]) * (z - a[1]) - b[1] ** 2)
    assert np.allclose(S, -G.imag / np.pi, rtol=1e-13, atol=1e-15)


def test_seeded_gaussian_is_unit_and_reproducible():
    a = po.gaussian_vector(11, 0, 1000)
    b = po.gaussian_vector(11, 0, 1000)
    c = po.gaussian_vector(11, 1, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(np.linalg.norm(a) - 1.0) < 1e-14


def test_mt19937_known_answer():
    # The C++ standard's check of std::mt19937 ([rand.predef]): the 10000th output of the
    # default-seeded (5489) engine is 4123659995.
    assert po.mt19937_nth(5489, 10000) == 4123659995
    assert po.mt19937_nth(5489, 1) == 3499211612   # the first output of MT19937 at its default seed


@pytest.mark.parametrize("base,sample,n", [(11, 0, 1), (11, 0, 155), (11, 3, 157), (7, 1, 5000), (2**63 + 5, 39, 20011)])
def test_polar_restatement_equals_library_draw(base, sample, n):
    """The device generator's algorithm restated (MT19937 words, 53-bit canonical uniforms, the polar
    method, (y m, x m) per kept candidate) equals std::normal_distribution on std::mt19937 bit for
    bit on the CPU (same libm log on both sides)."""
    want = po.gaussian_raw(base, sample, n)
    got, tried = po.gaussian_polar(po.seed32(base, sample), n)
    assert np.array_equal(got, want)
    # about pi / 4 of the candidates are kept
    if n >= 5000:
        assert abs(n / tried - np.pi / 4) < 0.02
    # planted slips in the restatement -- the words of a uniform swapped, the kept pair's normals in
    # the other order -- are told apart
    for slip in (1, 2):
        assert not np.array_equal(po.gaussian_polar(po.seed32(base, sample), n, slip)[0], want)


def _reference_records():
    import json

    with open(os.path.join(FIXTURES, "reference_records.json")) as f:
        return json.load(f)


def _full_space_j1j2_chain12():
    """The J1-J2 ring N = 12, J2 = 0.5 on the whole 2^12 space (trivial group): the reference's
    thermal 'none' block, states in ascending order."""
    fx = load_fixture("j1j2_n12_k0")   # the same Hamiltonian's row program
    s = po.OracleSector(12, -1, np.arange(12, dtype=np.int32)[None, :], [1.0])
    row_ptr, col, val = po.sector_csr(fixture_program(fx), s)
    return s, row_ptr, col, val


def test_ftlm_reproduces_reference_golden_record():
    """The oracle's FTLM (sample seeds, LocalDGKS3 Lanczos, Ritz weights, moments, Jensen combine)
    against the reference's own recorded FTLM of j1j2_chain12 with no symmetry (one 4096-state
    block, 8 samples, krylov 40, seed 11: the block's seed is the walk's first LCG step)."""
    rec = _reference_records()
    r = rec["records"]["j1j2_chain12/api/thermal/ftlm/none"]["values"]
    T = np.array(r["T"])
    betas = 1.0 / T
    s, row_ptr, col, val = _full_space_j1j2_chain12()
    assert s.dim == 4096
    seed = (11 * 6364136223846793005 + 1442695040888963407) % (1 << 64)
    moments = []
    for k in range(8):
        v0 = po.gaussian_vector(seed, k, s.dim)
        alpha, beta = po.lanczos(row_ptr, col, val, v0, 40, 2, 64 * np.finfo(float).eps * 36.0)
        ritz, V = po.tridiag_eig(alpha, beta)
        moments.append(po.ftlm_moments(ritz, V[0] ** 2, betas))
    lnZ, E, Vv, _ = po.ftlm_combine(np.array(moments), s.dim, betas)
    C = betas ** 2 * Vv
    S = lnZ + betas * E
    assert np.allclose(E, r["E"], rtol=1e-10, atol=1e-10)
    assert np.allclose(C, r["C"], rtol=1e-10, atol=1e-10)
    assert np.allclose(S, r["S"], rtol=1e-10, atol=1e-10)


def test_exact_thermal_reproduces_reference_golden_record():
    """Exact thermodynamics of the whole spectrum (dense eigvalsh of the oracle's matrix) against
    the reference's recorded exact curves of j1j2_chain12."""
    rec = _reference_records()
    r = rec["records"]["j1j2_chain12/api/thermal/exact/none"]["values"]
    betas = 1.0 / np.array(r["T"])
    s, row_ptr, col, val = _full_space_j1j2_chain12()
    ev = np.linalg.eigvalsh(dense_csr(row_ptr, col, val, s.dim))
    e0 = ev.min()
    for i, b in enumerate(betas):
        w = np.exp(-b * (ev - e0))
        Z = w.sum()
        E = (w * ev).sum() / Z
        V = (w * (ev - E) ** 2).sum() / Z
        assert abs(E - r["E"][i]) < 1e-10 * max(1.0, abs(E))
        assert abs(b * b * V - r["C"][i]) < 1e-10 * max(1.0, abs(r["C"][i]))
        assert abs(np.log(Z) - b * e0 + b * E - r["S"][i]) < 1e-10 * max(1.0, abs(r["S"][i]))


def test_reference_eigs_records_are_the_dense_spectrum():
    """The reference's recorded eigs levels of j1j2_chain12 (k = 4) are the lowest levels of the
    oracle's dense matrix: of the whole space (no symmetry, every symmetry, Sz blocks without spatial
    symmetry) and of the Sz = 0 block alone (popcount-6 states of the same matrix)."""
    rec = _reference_records()["records"]
    s, row_ptr, col, val = _full_space_j1j2_chain12()
    Hd = dense_csr(row_ptr, col, val, s.dim)
    ev = np.linalg.eigvalsh(Hd)
    for key in ("none", "auto", "spatial=none"):
        got = rec[f"j1j2_chain12/api/eigs/{key}/k=4"]["values"]["energies"]
        assert np.allclose(got, ev[:4], rtol=1e-10, atol=1e-10), key
    sz0 = np.array([i for i in range(1 << 12) if bin(i).count("1") == 6])
    ev0 = np.linalg.eigvalsh(Hd[np.ix_(sz0, sz0)])
    for key in ("sz=6", "auto/sz=6"):
        got = rec[f"j1j2_chain12/api/eigs/{key}/k=4"]["values"]["energies"]
        assert np.allclose(got, ev0[:4], rtol=1e-10, atol=1e-10), key


def test_reference_spectrum_records_are_the_dense_spectrum():
    """The reference's recorded whole spectra of j1j2_chain12 are the oracle's dense spectrum: the
    whole space under any symmetry selection, and the Sz = 0 block."""
    rec = _reference_records()["records"]
    s, row_ptr, col, val = _full_space_j1j2_chain12()
    Hd = dense_csr(row_ptr, col, val, s.dim)
    ev = np.linalg.eigvalsh(Hd)
    for key in ("none", "auto", "spatial=none"):
        got = np.array(rec[f"j1j2_chain12/api/spectrum/{key}"]["values"]["eigenvalues"])
        assert got.size == 4096 and np.allclose(np.sort(got), ev, rtol=1e-10, atol=1e-10), key
    sz0 = np.array([i for i in range(1 << 12) if bin(i).count("1") == 6])
    ev0 = np.linalg.eigvalsh(Hd[np.ix_(sz0, sz0)])
    for key in ("sz=6", "auto/sz=6"):
        got = np.array(rec[f"j1j2_chain12/api/spectrum/{key}"]["values"]["eigenvalues"])
        assert np.allclose(np.sort(got), ev0, rtol=1e-10, atol=1e-10), key


def _sz0_momentum_levels(Hd, n=12, n_up=6, per_k=2):
    """(energy, <S_0 . S_1>) of the lowest levels of every momentum block of the Sz = 0 states of a
    dense ring Hamiltonian, by projecting onto translation eigenspaces (a check independent of the
    sector code)."""
    st = np.array([i for i in range(1 << n) if bin(i).count("1") == n_up])
    pos = {int(s): i for i, s in enumerate(st)}
    H0 = Hd[np.ix_(st, st)]
    d = st.size
    T = np.zeros((d, d))
    for i, s in enumerate(st):
        t = ((int(s) << 1) | (int(s) >> (n - 1))) & ((1 << n) - 1)   # site j -> j + 1
        T[pos[t], i] = 1.0
    B = np.zeros((d, d))
    for i, s in enumerate(st):
        b0, b1 = int(s) & 1, (int(s) >> 1) & 1
        B[i, i] = 0.25 if b0 == b1 else -0.25
        if b0 != b1:
            B[pos[int(s) ^ 3], i] = 0.5
    out = []
    Tp = [np.linalg.matrix_power(T, a) for a in range(n)]
    for k in range(n):
        P = sum(np.exp(-2j * np.pi * k * a / n) * Tp[a] for a in range(n)) / n
        w, U = np.linalg.eigh((P + P.conj().T) / 2)
        Q = U[:, w > 0.5]
        e, V = np.linalg.eigh(Q.conj().T @ H0 @ Q)
        for j in range(min(per_k, e.size)):
            v = Q @ V[:, j]
            out.append((e[j], float(np.real(np.vdot(v, B @ v)))))
    return out


def test_reference_expect_record_matches_momentum_resolved_dense_states():
    """The reference's recorded <S_0 . S_1> in the 4 lowest levels of j1j2_chain12: every recorded
    (energy, bond) pair is a momentum eigenstate's of the oracle's dense Sz = 0 block (translation
    invariant states carry one value of the bond; the triplet's members share theirs)."""
    r = _reference_records()["records"]["j1j2_chain12/api/expect/auto/S0.S1/k=4"]["values"]
    s, row_ptr, col, val = _full_space_j1j2_chain12()
    lv = _sz0_momentum_levels(dense_csr(row_ptr, col, val, s.dim))
    for e, b in zip(r["energy"], r["bond"]):
        assert any(abs(e - e2) < 1e-9 and abs(b - b2) < 1e-9 for e2, b2 in lv), (e, b)
    # the two ground states (momentum 0 and pi) carry different bonds, so the matching tells them apart
    assert abs(r["bond"][0] - r["bond"][1]) > 1e-3

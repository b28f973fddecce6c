"""Level vectors on the device: the expansion of sector vectors into the Sz sector and the full space
(against the oracle's expansion matrix E), eigs(vectors=True) against the dense Hamiltonian of the
full space, expect / correlations against dense expectation values, and the .npz round trip.

Tolerances: expanded vectors within 1e-13 of E u / |E u|; eigenpairs within 1e-10 (energies) and
residual 1e-8 (the Krylov-Schur lock tolerance is 1e-10 s_H); expectation values within 1e-10."""

import sys

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import eigs as ev
from qed_spin_amd import errors
from qed_spin_amd import levels as lv
from qed_spin_amd import symmetry as sym

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _ring(n, j2=0.0):
    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    if j2:
        b.heisenberg([(i, (i + 2) % n) for i in range(n)], j2)
    return b.to_operator()


def _flip_group(N, m, z):
    """Translations times the global spin flip, characters chi(T^j) z^f."""
    T = sym.translations(N)
    ck = sym.momentum_chars(N, m)
    full = np.uint64((1 << N) - 1)
    perms = np.concatenate([T, T])
    flips = np.concatenate([np.zeros(N, np.uint64), np.full(N, full, np.uint64)])
    return qs.GroupSpec(perms=perms, chars=np.concatenate([ck, z * ck]), flips=flips)


GROUPS = [
    ("chain k=0", lambda N: sym.chain_group(N, 0)),
    ("chain k=3", lambda N: sym.chain_group(N, 3)),
    ("dihedral k=0 R=-1", lambda N: sym.dihedral_group(N, 0, -1)),
    ("chain k=5 R=+1", lambda N: sym.dihedral_group(N, 5, 1)),
    ("chain k=2 x flip -1", lambda N: _flip_group(N, 2, -1)),
]


@pytest.mark.parametrize("name,make", GROUPS, ids=[g[0] for g in GROUPS])
@pytest.mark.parametrize("full", [False, True])
def test_expand_matches_the_oracle_expansion(name, make, full):
    say("expand", name, "full" if full else "sz")
    N, n_up = 10, 5
    g = make(N)
    s = qs.Sector(g, n_up)
    if s.dim == 0:
        pytest.skip("empty sector")
    kb, L = s.key_bits()
    os_ = po.OracleSector(N, n_up, g.perms, g.chars, g.flips, kb if L else None)
    reps, inv = s.export()
    assert np.array_equal(reps, os_.reps)
    states = list(range(1 << N)) if full else [x for x in range(1 << N) if bin(x).count("1") == n_up]
    E = po.sector_basis_matrix(os_, states)
    rng = np.random.default_rng(7)
    U = rng.normal(size=(3, s.dim)) + 1j * rng.normal(size=(3, s.dim))
    U[2, ::2] = 0.0     # zero amplitudes are skipped
    B = lv.LevelBasis.of(g, s)
    got = lv.expand(B, list(U), full=full)
    for u, psi in zip(U, got):
        want = E @ u
        want /= np.linalg.norm(want)
        assert psi.shape == want.shape
        assert np.abs(psi - want).max() < 1e-13


def test_expand_refuses_what_it_cannot_write():
    say("expand refusals")
    N = 10
    s = qs.Sector(sym.trivial_group(N), -1)
    B = lv.LevelBasis.of(sym.trivial_group(N), s)
    with pytest.raises(errors.InvalidRequest):
        lv.expand(B, [np.ones(s.dim)], full=False)     # no fixed n_up: only the full space
    s2 = qs.Sector(sym.chain_group(N, 0), 5)
    B2 = lv.LevelBasis.of(sym.chain_group(N, 0), s2)
    with pytest.raises(errors.InvalidRequest):
        lv.expand(B2, [np.zeros(s2.dim)], full=False)  # the zero vector


def _dense(H, N):
    return po.dense_matrix(H.products, N)


@pytest.fixture(scope="module")
def ring10():
    N = 10
    H = _ring(N, 0.35)
    r = ev.eigs(H, 6, vectors=True)
    return N, H, r, _dense(H, N)


def test_eigs_vectors_full_space(ring10):
    """Every (Sz, k) block of the N = 10 J1-J2 ring: the 6 lowest levels' vectors are orthonormal
    eigenvectors of the dense H of the full space with the returned energies."""
    say("eigs vectors full")
    N, H, r, Hd = ring10
    w = np.linalg.eigvalsh(Hd)
    assert np.allclose(r.energies, w[:6], atol=1e-10)
    V = np.array(r.vectors("full")).T
    assert V.shape == (1 << N, 6)
    assert np.abs(V.conj().T @ V - np.eye(6)).max() < 1e-10
    HV = Hd @ V
    for i, e in enumerate(r.energies):
        assert np.linalg.norm(HV[:, i] - e * V[:, i]) < 1e-8
        assert abs(np.vdot(V[:, i], HV[:, i]).real - e) < 1e-10
    assert [L["vector"] for L in r.levels] == list(range(6))


def test_eigs_vectors_sz_basis(ring10):
    """basis='sz': the levels of that n_up only, each the full-space vector restricted to the
    popcount-n_up states (ascending), exactly."""
    say("eigs vectors sz")
    N, H, r, Hd = ring10
    full = r.vectors("full")
    for n_up in sorted({L["n_up"] for L in r.levels}):
        idx = np.array([x for x in range(1 << N) if bin(x).count("1") == n_up])
        sz = r.vectors("sz", n_up=n_up)
        mine = [f for f, L in zip(full, r.levels) if L["n_up"] == n_up]
        assert len(sz) == len(mine) > 0
        for a, b in zip(sz, mine):
            assert np.abs(a - b[idx]).max() < 1e-14
    assert r.vectors("sz", n_up=0) == []
    with pytest.raises(errors.InvalidRequest):
        r.vectors("sz")
    with pytest.raises(errors.InvalidRequest):
        ev.eigs(H, 2).vectors()


def _site_ops(N):
    ops = []
    for i in range(3):
        ops.append(qs.Operator(N).add_product("z", (i,), 1.0))
    ops.append(qs.Operator(N).add_product("x", (0,), 1.0))
    ops.append(qs.Operator(N).add_product("-", (1,), 1.0))
    ops.append(qs.Operator(N).add_product("+-", (0, 2), 1.0).add_product("zz", (1, 3), 0.5))
    ops.append(qs.Operator(N).add_product("y", (4,), 1.0).add_product("zy", (0, 5), 2.0))
    return ops


def test_expect_and_correlations_against_dense(ring10):
    """<v|O|v> (observables that break the symmetry included) and <A_a^dag B_b> for operators that
    change n_up by 0, -1 or both (S^x, S^y), against the dense matrices of the full space."""
    say("expect and correlations")
    N, H, r, Hd = ring10
    V = r.vectors("full")
    ops = _site_ops(N)
    M = [_dense(O, N) for O in ops]
    ex = r.expect(ops)
    for i, v in enumerate(V):
        for o in range(len(ops)):
            assert abs(ex[i, o] - np.vdot(v, M[o] @ v)) < 1e-10, (i, o)
    Cm = r.correlations(ops)
    assert Cm.shape == (len(V), len(ops), len(ops))
    for i, v in enumerate(V):
        for a in range(len(ops)):
            for b in range(len(ops)):
                want = np.vdot(M[a] @ v, M[b] @ v)
                assert abs(Cm[i, a, b] - want) < 1e-10, (i, a, b)
    Cab = r.correlations(ops[:2], ops[3:5])
    assert np.allclose(Cab, Cm[:, :2, 3:5], atol=1e-12)


def test_save_and_load_round_trip(ring10, tmp_path):
    """save() then load_eigs(): energies, levels and vectors as they were; vectors(), expect() and
    correlations() work on the loaded result (its bases imported, no H)."""
    say("save/load")
    N, H, r, Hd = ring10
    p = tmp_path / "levels.npz"
    r.save(p)
    s = ev.load_eigs(p)
    assert np.array_equal(s.energies, r.energies)
    assert s.k == r.k and s.n_sites == N
    assert [L["n_up"] for L in s.levels] == [L["n_up"] for L in r.levels]
    assert [L["momentum"] for L in s.levels] == [L["momentum"] for L in r.levels]
    a, b = r.vectors("full"), s.vectors("full")
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() < 1e-14
    ops = _site_ops(N)[:3]
    assert np.allclose(s.expect(ops), r.expect(ops), atol=1e-13)
    with np.load(p) as f:
        assert int(f["format_version"]) == 2
        assert f["level_tag"].size == 11 * len(r.levels)
        assert int(f["n_bases"][0]) == len({L["label"] for L in r.levels})


def test_every_level_of_a_small_ring(tmp_path):
    """N = 6, k = 64: every level of every (Sz, k) block, blocks of one and two states included --
    the vectors form a unitary that diagonalises the dense H."""
    say("all levels N=6")
    N = 6
    H = _ring(N, 0.2)
    r = ev.eigs(H, 1 << N, vectors=True)
    Hd = _dense(H, N)
    V = np.array(r.vectors("full")).T
    assert V.shape == (1 << N, 1 << N)
    assert np.abs(V.conj().T @ V - np.eye(1 << N)).max() < 1e-10
    assert np.abs(V.conj().T @ Hd @ V - np.diag(r.energies)).max() < 1e-9
    assert min(L["dim"] for L in r.levels) == 1


def test_sublattice_bases_survive_the_file(tmp_path):
    """A sector whose representatives are least in the sublattice key order (the dihedral group of
    the N = 16 ring, sublattice code forced on): its fingerprint is written and checked again on
    load; the loaded basis expands a vector exactly as the original; a wrong fingerprint is refused."""
    say("sublattice file")
    N = 16
    g = sym.dihedral_group(N, 0, 1)
    sec = qs.Sector(g, N // 2, sublattice=1)
    B = lv.LevelBasis.of(g, sec)
    assert B.sublattice != 0
    rng = np.random.default_rng(5)
    u = rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim)
    r = ev.EigsResult(energies=np.array([0.0]), k=1, n_sites=N, _bases={0: B}, _vectors=[(0, u / np.linalg.norm(u))],
                      levels=[{"energy": 0.0, "n_up": N // 2, "momentum": 0, "parity": 1, "dim": sec.dim,
                               "label": "n_up=8,k=0,R=+1", "vector": 0}])
    p = tmp_path / "slc.npz"
    r.save(p)
    s = ev.load_eigs(p)
    a, b = r.vectors("sz", n_up=N // 2), s.vectors("sz", n_up=N // 2)
    assert len(a) == len(b) == 1
    assert np.abs(a[0] - b[0]).max() < 1e-14
    with np.load(p) as f:
        d = {k: f[k] for k in f.files}
    d["basis0_sublattice"] = np.array([B.sublattice ^ 1], np.uint64)
    bad = tmp_path / "bad.npz"
    np.savez(bad, **d)
    with pytest.raises(errors.InvalidRequest):
        ev.load_eigs(bad).vectors("full")


@pytest.mark.parametrize("total_spin", [0.5, 1.5])
def test_total_spin_against_the_dense_spin_subspace(total_spin):
    """N = 9 J1-J2 ring: the 4 lowest energies of spin S (each multiplet counted 2S + 1 times, as
    the reference counts them) against eigvalsh of H restricted to the S(S+1) eigenspace of the
    dense S^2 in the Sz = S sector (1e-10); the levels' vectors carry <S^2> = S(S+1) (1e-9)."""
    say("total_spin dense", total_spin)
    N = 9
    H = _ring(N, 0.3)
    two_S = int(2 * total_spin)
    n_up = (N + two_S) // 2
    states = [x for x in range(1 << N) if bin(x).count("1") == n_up]
    S2 = ev.spin_squared(N)
    Hd = po.dense_matrix(H.products, N, states)
    S2d = po.dense_matrix(S2.products, N, states)
    w, Q = np.linalg.eigh(S2d)
    lam = total_spin * (total_spin + 1)
    Q = Q[:, np.abs(w - lam) < 1e-8]
    want = np.repeat(np.linalg.eigvalsh(Q.conj().T @ Hd @ Q), two_S + 1)[:4]   # a multiplet counts 2S + 1
    for dmd in (None, 0):
        r = ev.eigs(H, 4, total_spin=total_spin, vectors=True, dense_max_dim=dmd)
        assert np.allclose(r.energies, want, atol=1e-10), (dmd, r.energies, want)
        assert len(r.levels) == -(-4 // (two_S + 1))
        assert all(L["multiplicity"] == two_S + 1 for L in r.levels)
        s2 = r.expect([S2])[:, 0].real
        assert np.allclose(s2, lam, atol=1e-9), s2
    with pytest.raises(errors.InvalidRequest):
        ev.eigs(H, 2, total_spin=1.0)            # not a spin of 9 spins 1/2
    with pytest.raises(errors.InvalidRequest):
        ev.eigs(H, 2, total_spin=0.5, sym=qs.Symmetry(sz=[4]))


@pytest.mark.parametrize("total_spin", [0.5, 1.5])
def test_tower_vectors_complete_the_multiplets(total_spin):
    """vectors() of a spin tower: the k = 4 vectors are the multiplets' members from Sz = S down,
    lowered by S^- on the device -- orthonormal eigenvectors of the dense H with the returned
    energies, <S^2> = S(S+1), <S^z> stepping down by one; basis='sz' gives the member of that n_up,
    the full vector restricted to it."""
    say("tower vectors", total_spin)
    N = 9
    H = _ring(N, 0.3)
    r = ev.eigs(H, 4, total_spin=total_spin, vectors=True)
    V = np.array(r.vectors("full")).T
    assert V.shape == (1 << N, 4)
    assert np.abs(V.conj().T @ V - np.eye(4)).max() < 1e-10
    Hd = _dense(H, N)
    S2 = _dense(ev.spin_squared(N), N)
    Sz = _dense(qs.Operator(N).add_terms([(1.0, (("z", i),)) for i in range(N)]), N)
    lam = total_spin * (total_spin + 1)
    two_S = int(2 * total_spin)
    for i in range(4):
        v = V[:, i]
        e = r.energies[i]
        assert np.linalg.norm(Hd @ v - e * v) < 1e-8, i
        assert abs(np.vdot(v, S2 @ v).real - lam) < 1e-9, i
        assert abs(np.vdot(v, Sz @ v).real - (total_spin - (i % (two_S + 1)))) < 1e-9, i
    n_low = (N + two_S) // 2 - 1            # one step below the top member
    idx = np.array([x for x in range(1 << N) if bin(x).count("1") == n_low])
    sz = r.vectors("sz", n_up=n_low)
    assert len(sz) == len(r.levels)
    assert np.abs(sz[0] - V[idx, 1]).max() < 1e-12


def test_tower_result_survives_the_file(tmp_path):
    """A total_spin result saved and loaded: 2S travels in spec_scalars[5], the levels keep their
    multiplicities, and the loaded result's vectors() completes the multiplets as the original's."""
    say("tower file")
    N = 9
    r = ev.eigs(_ring(N, 0.3), 4, total_spin=0.5, vectors=True)
    p = tmp_path / "tower.npz"
    r.save(p)
    with np.load(p) as f:
        assert int(f["spec_scalars"][5]) == 1
        assert list(f["level_multiplicity"]) == [2] * len(r.levels)
    s = ev.load_eigs(p)
    assert all(L["total_spin"] == 0.5 for L in s.levels)
    a, b = r.vectors("full"), s.vectors("full")
    assert len(a) == len(b) == 4
    for x, y in zip(a, b):
        assert np.abs(x - y).max() < 1e-13

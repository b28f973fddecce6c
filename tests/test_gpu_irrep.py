"""Sectors of irreps of dimension d > 1 on the device (qedh_sector_create_irrep and the CSR built
representative by representative), against the oracle and what the mathematics fixes:
- representatives and ranks equal the oracle's brute force (irrep_fields);
- the basis defined by the device's own C matrices is orthonormal, and the device CSR equals
  E^dag H E over the Sz sector entry by entry (1e-12), its applies the same within 1e-13;
- the spectrum equals the momentum-m block's (the E_m copies and the momentum-m states correspond
  one to one), at small sizes by dense eigensolves and at N = 24 through the device Krylov lanes;
- d = 4..8 from 2-transitive groups, against the same checks;
- the operator between two sectors of one irrep (total S-) against E^dag S- E;
- refusals: the on-the-fly lane, d above the reference's limit of 8, matrices that are not a
  representation, the per-state lookup, two sectors of different irreps.
Every case prints its name before it touches the device."""

import sys

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import symmetry as sym
from qed_spin_amd.errors import DeviceUnsupported, InvalidRequest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ring_h(N, J2=0.5, chi=0.0):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    b.heisenberg([(i, (i + 2) % N) for i in range(N)], J2)
    return b.to_operator()


def sz_hamiltonian(H, N, n_up):
    triv = po.OracleSector(N, n_up, np.arange(N, dtype=np.int32)[None, :], np.ones(1))
    rp, col, val = po.sector_csr(qs.Program(H).export(), triv)
    A = np.zeros((triv.dim, triv.dim), complex)
    for r in range(triv.dim):
        A[r, col[rp[r]:rp[r + 1]]] += val[rp[r]:rp[r + 1]]
    return A, triv.reps


def dense(m):
    """The CSR as a dense matrix; repeated columns of a row are summed."""
    row_ptr, col, val = m.export_csr()
    A = np.zeros((m.dim, m.dim), complex)
    for r in range(m.dim):
        np.add.at(A[r], col[row_ptr[r]:row_ptr[r + 1]].astype(np.int64), val[row_ptr[r]:row_ptr[r + 1]])
    return A


def merged(m):
    """Whether every row of the CSR lists each column once, ascending."""
    row_ptr, col, _ = m.export_csr()
    for r in range(m.dim):
        c = col[row_ptr[r]:row_ptr[r + 1]].astype(np.int64)
        if np.any(np.diff(c) <= 0):
            return False
    return True


@pytest.mark.parametrize("N,n_up,m", [(8, 4, 1), (8, 4, 2), (8, 3, 3), (10, 5, 2), (10, 4, 1), (12, 6, 5), (12, 5, 3)])
def test_irrep_sector_against_the_oracle(N, n_up, m):
    say("irrep sector", N, n_up, m)
    H = ring_h(N)
    gD, D = sym.dihedral_irrep(N, m)
    s = qs.Sector(gD, n_up, irrep=D)
    assert s.irrep_dim == 2
    reps_o, Cs_o = po.irrep_fields(N, n_up, gD.perms, D)
    reps, inv = s.export()
    assert np.array_equal(reps, reps_o)
    rep_class, offset, rank, Cm = s.export_irrep()
    ranks = np.diff(offset.astype(np.int64))
    assert list(ranks) == [C.shape[1] for C in Cs_o]
    assert s.dim == int(ranks.sum()) == int(offset[-1])
    assert all(ranks[i] == rank[rep_class[i]] for i in range(len(reps)))
    Cs = [Cm[rep_class[i]][:, : ranks[i]] for i in range(len(reps))]
    Hs, states = sz_hamiltonian(H, N, n_up)
    E = po.irrep_basis_matrix(N, gD.perms, D, reps, Cs, states)
    assert np.abs(E.conj().T @ E - np.eye(s.dim)).max() <= 1e-12
    mv = qs.SectorMatvec(H, s, lane=qs.LANE_CSR, tune=False)
    assert mv.dim == s.dim and not mv.is_real
    assert merged(mv)
    A = dense(mv)
    ref = E.conj().T @ Hs @ E
    assert np.abs(A - ref).max() <= 1e-12 * max(1.0, np.abs(Hs).max())
    rng = np.random.default_rng(N * 100 + m)
    x = rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim)
    y = mv.apply_host(x, complex_vectors=True)
    assert np.linalg.norm(y - ref @ x) <= 1e-13 * np.linalg.norm(ref @ x)
    g = sym.chain_group(N, m)
    o = po.OracleSector(N, n_up, g.perms, g.chars)
    rk, ck, vk = po.sector_csr(qs.Program(H).export(), o)
    Ak = np.zeros((o.dim, o.dim), complex)
    for r in range(o.dim):
        Ak[r, ck[rk[r]:rk[r + 1]]] += vk[rk[r]:rk[r + 1]]
    assert np.abs(np.linalg.eigvalsh(A) - np.linalg.eigvalsh(Ak)).max() <= 1e-12 * max(1.0, np.abs(Hs).max())


@pytest.mark.parametrize("n,layers,n_up", [(3, 2, 3), (4, 2, 4), (4, 2, 3), (5, 2, 5), (5, 2, 4)])
def test_standard_irreps_of_symmetric_groups_d2_to_d4(n, layers, n_up):
    from irrep_groups import layered_heisenberg, layered_symmetric_group, standard_irrep
    from test_irrep_oracle import isotypic_spectrum

    say("symmetric-group irrep sector", n, layers, n_up)
    N = n * layers
    perms, els = layered_symmetric_group(n, layers)
    D = standard_irrep(els, n)
    d = D.shape[1]
    g = sym.GroupSpec(perms=perms, chars=np.trace(D, axis1=1, axis2=2))
    s = qs.Sector(g, n_up, irrep=D)
    assert s.irrep_dim == d
    reps_o, Cs_o = po.irrep_fields(N, n_up, perms, D)
    reps, _ = s.export()
    assert np.array_equal(reps, reps_o)
    rep_class, offset, rank, Cm = s.export_irrep()
    ranks = np.diff(offset.astype(np.int64))
    assert list(ranks) == [C.shape[1] for C in Cs_o]
    Cs = [Cm[rep_class[i]][:, : ranks[i]] for i in range(len(reps))]
    H = layered_heisenberg(qs, n, layers)
    Hs, states = sz_hamiltonian(H, N, n_up)
    E = po.irrep_basis_matrix(N, perms, D, reps, Cs, states)
    assert np.abs(E.conj().T @ E - np.eye(s.dim)).max() <= 1e-12
    mv = qs.SectorMatvec(H, s, lane=qs.LANE_CSR, tune=False)
    assert merged(mv)
    A = dense(mv)
    assert np.abs(A - E.conj().T @ Hs @ E).max() <= 1e-12 * max(1.0, np.abs(Hs).max())
    iso, _ = isotypic_spectrum(Hs, states, N, perms, D)
    assert np.abs(np.sort(np.repeat(np.linalg.eigvalsh(A), d)) - np.sort(iso)).max() <= 1e-11


def test_irrep_sector_n24_lowest_levels_equal_the_momentum_block():
    say("irrep sector N=24, E_5")
    N, m = 24, 5
    H = ring_h(N, 0.35)
    gD, D = sym.dihedral_irrep(N, m)
    s = qs.Sector(gD, N // 2, irrep=D)
    k = qs.Sector(sym.chain_group(N, m), N // 2)
    assert s.dim == k.dim
    a = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    b = qs.SectorMatvec(H, k, lane=qs.LANE_CSR)
    va, ca, _ = a.lowest(3)
    vb, cb, _ = b.lowest(3)
    assert ca and cb
    assert np.allclose(va, vb, rtol=1e-10, atol=1e-10)


def test_irrep_sector_refusals():
    say("irrep sector refusals")
    N = 8
    H = ring_h(N)
    gD, D = sym.dihedral_irrep(N, 1)
    s = qs.Sector(gD, 4, irrep=D)
    with pytest.raises(DeviceUnsupported):
        qs.SectorMatvec(H, s, lane=qs.LANE_GATHER)
    with pytest.raises(DeviceUnsupported):
        s.index_and_projection(np.array([0b00001111], np.uint64))
    # two handles of one irrep sector: the cross operator (no diagonal shortcut) is the one-sector
    # operator within rounding
    a = dense(qs.SectorMatvec(H, s, lane=qs.LANE_CSR, tune=False))
    b = dense(qs.SectorMatvec(H, s, qs.Sector(gD, 4, irrep=D), lane=qs.LANE_CSR, tune=False))
    assert np.abs(a - b).max() <= 1e-13
    big = np.zeros((gD.G, 9, 9), complex)
    big[:, :, :] = np.eye(9)
    with pytest.raises(DeviceUnsupported):
        qs.Sector(gD, 4, irrep=big)
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.normal(size=(gD.G, 2, 2)) + 1j * rng.normal(size=(gD.G, 2, 2)))[0]
    with pytest.raises(InvalidRequest):
        qs.Sector(gD, 4, irrep=Q)


@pytest.mark.parametrize("name,layers,n_up", [("agl5", 2, 4), ("psl2_5", 2, 5), ("agl7", 2, 4), ("agl8", 2, 4),
                                              ("agl9", 2, 4)])
def test_standard_irreps_of_two_transitive_groups_d4_to_d8(name, layers, n_up):
    """d = 4..8 (up to the reference's kMaxIrrepDim): representatives and ranks equal the oracle's,
    the device's C matrices give an orthonormal basis, the device CSR equals E^dag H E entry by
    entry, and its spectrum is the isotypic spectrum of the Sz sector."""
    from irrep_groups import layered_heisenberg, standard_irrep
    from test_irrep_oracle import isotypic_spectrum, two_transitive

    say("two-transitive irrep sector", name, layers, n_up)
    perms, els, n = two_transitive(name, layers)
    N = n * layers
    D = standard_irrep(els, n)
    d = D.shape[1]
    g = sym.GroupSpec(perms=perms, chars=np.trace(D, axis1=1, axis2=2))
    s = qs.Sector(g, n_up, irrep=D)
    assert s.irrep_dim == d
    reps_o, Cs_o = po.irrep_fields(N, n_up, perms, D)
    reps, _ = s.export()
    assert np.array_equal(reps, reps_o)
    rep_class, offset, rank, Cm = s.export_irrep()
    ranks = np.diff(offset.astype(np.int64))
    assert list(ranks) == [C.shape[1] for C in Cs_o]
    Cs = [Cm[rep_class[i]][:, : ranks[i]] for i in range(len(reps))]
    H = layered_heisenberg(qs, n, layers)
    Hs, states = sz_hamiltonian(H, N, n_up)
    E = po.irrep_basis_matrix(N, perms, D, reps, Cs, states)
    assert np.abs(E.conj().T @ E - np.eye(s.dim)).max() <= 1e-12
    mv = qs.SectorMatvec(H, s, lane=qs.LANE_CSR, tune=False)
    assert merged(mv)
    A = dense(mv)
    assert np.abs(A - E.conj().T @ Hs @ E).max() <= 1e-12 * max(1.0, np.abs(Hs).max())
    iso, _ = isotypic_spectrum(Hs, states, N, perms, D)
    assert np.abs(np.sort(np.repeat(np.linalg.eigvalsh(A), d)) - np.sort(iso)).max() <= 1e-11


def sz_states(N, n_up):
    return np.array([s for s in range(1 << N) if bin(s).count("1") == n_up], np.int64)


@pytest.mark.parametrize("name,n_up", [("s4", 4), ("agl7", 4), ("agl9", 4)])
def test_cross_operator_between_sectors_of_one_irrep(name, n_up):
    """Total S- from the irrep sector at n_up to the one at n_up - 1 (the reference's
    CrossSectorMatVec between two sectors of one irrep): the device CSR equals E_rows^dag S- E_cols
    entry by entry, its applies agree with it, and S- carries an eigenvector of the SU(2)-invariant
    H at n_up to one at n_up - 1 with the same energy. Sectors of different irreps refuse."""
    from irrep_groups import layered_heisenberg, layered_symmetric_group, standard_irrep
    from test_irrep_oracle import two_transitive

    say("cross operator between irrep sectors", name, n_up)
    if name == "s4":
        n = 4
        perms, els = layered_symmetric_group(n, 2)
    else:
        perms, els, n = two_transitive(name, 2)
    N = 2 * n
    D = standard_irrep(els, n)
    g = sym.GroupSpec(perms=perms, chars=np.trace(D, axis1=1, axis2=2))
    src, tgt = qs.Sector(g, n_up, irrep=D), qs.Sector(g, n_up - 1, irrep=D)
    Sm = qs.Operator(N)
    for i in range(N):
        Sm.add_product("-", (i,), 1.0)
    mv = qs.SectorMatvec(Sm, tgt, src, lane=qs.LANE_CSR)
    assert (mv.dim, src.dim) == (tgt.dim, src.dim)
    row_ptr, col, val = mv.export_csr()
    A = np.zeros((tgt.dim, src.dim), complex)
    for r in range(tgt.dim):
        np.add.at(A[r], col[row_ptr[r]:row_ptr[r + 1]].astype(np.int64), val[row_ptr[r]:row_ptr[r + 1]])

    def basis(sec, nu):
        st = sz_states(N, nu)
        rc, off, _, Cm = sec.export_irrep()
        rk = np.diff(off.astype(np.int64))
        reps, _ = sec.export()
        Cs = [Cm[rc[i]][:, : rk[i]] for i in range(len(reps))]
        return po.irrep_basis_matrix(N, perms, D, reps, Cs, st), st

    Es, ss = basis(src, n_up)
    Et, st = basis(tgt, n_up - 1)
    allst = np.concatenate([st, ss])
    M = po.dense_matrix(Sm.products, N, allst)[: len(st), len(st):]
    ref = Et.conj().T @ M @ Es
    assert np.abs(A - ref).max() <= 1e-12
    rng = np.random.default_rng(7)
    x = rng.normal(size=src.dim) + 1j * rng.normal(size=src.dim)
    y = mv.apply_host(x, complex_vectors=True)
    assert np.abs(y - ref @ x).max() <= 1e-12 * max(1.0, np.abs(ref @ x).max())
    H = layered_heisenberg(qs, n, 2)
    Hs = dense(qs.SectorMatvec(H, src, lane=qs.LANE_CSR, tune=False))
    Ht = dense(qs.SectorMatvec(H, tgt, lane=qs.LANE_CSR, tune=False))
    w, V = np.linalg.eigh(Hs)
    moved = 0
    for k in range(len(w)):
        u = ref @ V[:, k]
        if np.linalg.norm(u) > 1e-6:
            assert np.abs(Ht @ u - w[k] * u).max() <= 1e-10 * max(1.0, abs(w[k])) * np.linalg.norm(u)
            moved += 1
    assert moved > 0
    # an equivalent irrep in another basis is another D: refused, as is a sector of characters
    Q = np.linalg.qr(rng.normal(size=(D.shape[1], D.shape[1])))[0]
    other = qs.Sector(g, n_up - 1, irrep=Q.T @ D @ Q)
    with pytest.raises(InvalidRequest):
        qs.SectorMatvec(Sm, other, src, lane=qs.LANE_CSR)
    trivial = sym.GroupSpec(perms=perms, chars=np.ones(len(els)))
    with pytest.raises(InvalidRequest):
        qs.SectorMatvec(Sm, qs.Sector(trivial, n_up - 1), src, lane=qs.LANE_CSR)


@pytest.mark.parametrize("name", ["agl8", "agl9"])
def test_irrep_sector_krylov_and_multi_vector_applies(name):
    """On d = 7 and 8 sectors: the device Krylov lane's lowest level equals the dense spectrum's of
    the device CSR (1e-10; the all-to-all layers' larger symmetry can make higher levels
    degenerate, which one Krylov run need not resolve), multi-vector and block-interleaved applies equal single applies bit for
    bit, on one sector and on the cross operator total S- between two sectors of the irrep."""
    from irrep_groups import layered_heisenberg, standard_irrep
    from test_irrep_oracle import two_transitive

    say("irrep sector krylov and multi-vector applies", name)
    perms, els, n = two_transitive(name, 2)
    N = 2 * n
    D = standard_irrep(els, n)
    g = sym.GroupSpec(perms=perms, chars=np.trace(D, axis1=1, axis2=2))
    n_up = 5
    s = qs.Sector(g, n_up, irrep=D)
    H = layered_heisenberg(qs, n, 2, rung=0.45)
    mv = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    ev = np.linalg.eigvalsh(dense(mv))
    vals, conv, _ = mv.lowest(1)
    assert conv
    assert abs(vals[0] - ev[0]) <= 1e-10 * max(1.0, abs(ev[0]))
    Sm = qs.Operator(N)
    for i in range(N):
        Sm.add_product("-", (i,), 1.0)
    down = qs.Sector(g, n_up - 1, irrep=D)
    cross = qs.SectorMatvec(Sm, down, s, lane=qs.LANE_CSR)
    rng = np.random.default_rng(11)
    for m, cols in ((mv, s.dim), (cross, s.dim)):
        xs = [rng.normal(size=cols) + 1j * rng.normal(size=cols) for _ in range(4)]
        singles = [m.apply_host(x, complex_vectors=True) for x in xs]
        X = [qs.DeviceArray.from_host(x) for x in xs]
        Y = [qs.DeviceArray(m.dim) for _ in xs]
        m.apply_multi(X, Y, complex_vectors=True)
        for y, ref in zip(Y, singles):
            assert np.array_equal(y.to_host().view(np.uint8), ref.view(np.uint8))
        XB = qs.DeviceArray.from_host(np.ascontiguousarray(np.stack(xs, axis=1)).reshape(-1))
        YB = qs.DeviceArray(m.dim * 4)
        m.apply_block(4, XB, YB, complex_vectors=True)
        got = YB.to_host().reshape(m.dim, 4)
        for i, ref in enumerate(singles):
            assert np.array_equal(np.ascontiguousarray(got[:, i]).view(np.uint8), ref.view(np.uint8))

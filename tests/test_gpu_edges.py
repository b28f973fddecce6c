"""GPU edge cases: the smallest and empty sectors, Krylov depths beyond the block, the batched engine
on blocks of one and two states, and refused requests (no kernel runs on an invalid input: the host
checks turn it into an error first). Every case prints its name before it touches the device."""

import sys

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import eigs as ev
from qed_spin_amd import errors
from qed_spin_amd import symmetry as sym
from qed_spin_amd import thermal as th

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


@pytest.mark.parametrize("n_up", [0, 8])
def test_one_state_sectors(n_up):
    """All spins down / all up (one state, momentum 0): the sector, its 1x1 CSR, the level
    (J N / 4 for the ring) by the scan and by the dense lane, the exact thermodynamics."""
    say("one-state sector", n_up)
    n = 8
    H = _ring(n)
    s = qs.Sector(sym.chain_group(n, 0), n_up)
    assert s.dim == 1
    m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    rp, col, val = m.export_csr()
    assert list(rp) == [0, 1] and list(col) == [0] and val[0] == pytest.approx(n / 4)
    v, conv, _ = m.lowest(1)
    assert conv and v[0] == pytest.approx(n / 4, abs=1e-14)
    lnZ, E, V, emin = th.exact_thermal([m], np.array([0.5, 2.0]))
    assert np.allclose(E[0], n / 4) and np.allclose(V[0], 0.0, atol=1e-12)


def test_empty_momentum_sector_is_skipped_and_reported():
    """A momentum block with no state (N = 4, n_up = 0 at k = 1) has dimension 0; a walk over it skips
    it, and asking for a walk of only empty blocks is refused."""
    say("empty sector")
    n = 4
    s = qs.Sector(sym.chain_group(n, 1), 0)
    assert s.dim == 0
    H = _ring(n)
    blocks = th.walk_blocks(H, qs.Symmetry(sz=[0], momenta=[1]))
    assert blocks == []
    with pytest.raises(errors.InvalidRequest):
        ev.eigs(H, 1, sym=qs.Symmetry(sz=[0], momenta=[1]))


def test_krylov_depth_beyond_the_block():
    """Lanczos asked for more steps than the block has states stops at the invariant subspace, and
    its Ritz values are then the block's spectrum (dense eigvalsh of the oracle's CSR)."""
    say("krylov beyond the block")
    n = 8
    H = _ring(n, 0.3)
    g = sym.chain_group(n, 0)
    s = qs.Sector(g, 4)
    m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    a, b, _ = m.lanczos(np.ones(s.dim) / np.sqrt(s.dim), 50, reorth=1)
    assert len(a) <= s.dim
    T = np.diag(a) + np.diag(b[1:len(a)], 1) + np.diag(b[1:len(a)], -1)
    os_ = po.OracleSector(n, 4, g.perms, g.chars)
    rp, col, val = po.sector_csr(qs.Program(H, adjoint=True).export(), os_)
    Hd = np.zeros((os_.dim, os_.dim), complex)
    for r in range(os_.dim):
        for e in range(rp[r], rp[r + 1]):
            Hd[r, col[e]] += val[e]
    ev_dense = np.linalg.eigvalsh(Hd)
    ritz = np.linalg.eigvalsh(T)
    # the start is momentum-0 and reflection-symmetric: its Krylov space holds a subset of the levels
    assert np.all(np.min(np.abs(ritz[:, None] - ev_dense[None, :]), axis=1) < 1e-10)


def test_batched_ftlm_on_blocks_of_one_and_two_states():
    """The batched engine with blocks whose Krylov space is full after one or two steps (the per-block
    cap ends those pairs on the device) against the per-block FTLM driver (1e-8)."""
    say("batched ftlm tiny blocks")
    n = 6
    H = _ring(n, 0.2)
    mats = []
    for n_up, k in ((0, 0), (1, 0), (1, 3), (2, 1), (3, 0)):
        s = qs.Sector(sym.chain_group(n, k), n_up)
        if s.dim:
            mats.append(qs.SectorMatvec(H, s, lane=qs.LANE_CSR))
    dims = [m.dim for m in mats]
    assert 1 in dims
    betas = np.array([0.5, 1.0, 4.0])
    seeds = [11 + 7 * i for i in range(len(mats))]
    lnZ, E, V, emin, _ = th.ftlm_batched(mats, 9, 30, seeds, betas)
    for b, m in enumerate(mats):
        l1, e1, v1, _ = m.ftlm(9, 30, seeds[b], betas)
        assert np.allclose(lnZ[b], l1, rtol=1e-8, atol=1e-8), dims[b]
        assert np.allclose(E[b], e1, rtol=1e-8, atol=1e-8), dims[b]


def test_refused_inputs():
    """Invalid requests are refused on the host before any kernel runs: a sector with n_up > N, a
    permutation that is not one, a serve request above the whole HBM budget, a spectrum block above
    the dense lanes' limit is not tried."""
    say("refused inputs")
    with pytest.raises(errors.InvalidRequest):
        qs.Sector(sym.chain_group(6, 0), 7)
    bad = qs.GroupSpec(perms=np.array([[0, 0, 1, 2]], np.int32), chars=np.array([1.0 + 0j]))
    with pytest.raises(errors.InvalidRequest):
        qs.Sector(bad, 2)
    from qed_spin_amd import serve as sv

    with pytest.raises(errors.QEDError):
        sv.serve([dict(kind="eigs", n_sites=20)], hbm_budget=1 << 20)


@pytest.mark.parametrize("N,group,n_up,sublattice,stride", [(18, "dihedral", 9, 0, 1), (16, "chain", -1, 0, 1),
                                                          (20, "dihedral", 10, 1, 3)])
def test_rank_directory_against_binary_search(N, group, n_up, sublattice, stride):
    """The rank directory (a bit per state of the subspace and a base per 64 states) and the binary
    search over the representatives (rank_budget = 0) resolve every lookup alike: the same merged CSR
    bit for bit, the same gather-lane apply, index_and_projection equal on every state of the subspace
    (every third one at N = 20, where the sublattice key order is engaged). An imported sector builds
    its directory from the representatives: the same again."""
    say("rank directory vs binary search", N, group, n_up, sublattice)
    g = sym.dihedral_group(N, 0, +1) if group == "dihedral" else sym.chain_group(N, 1)
    H = _ring(N, 0.4)
    a = qs.Sector(g, n_up, sublattice=sublattice)
    b = qs.Sector(g, n_up, sublattice=sublattice, rank_budget=0)
    ra, ia = a.export()
    rb, ib = b.export()
    assert np.array_equal(ra, rb) and np.array_equal(ia, ib)
    c = qs.Sector(g, n_up, sublattice=sublattice, reps=ra, inv_norms=ia)
    assert a.sublattice_block == b.sublattice_block == c.sublattice_block
    assert (a.sublattice_block > 0) == (sublattice == 1)
    csr = [qs.SectorMatvec(H, s, lane=qs.LANE_CSR).export_csr() for s in (a, b, c)]
    for other in csr[1:]:
        for x, y in zip(csr[0], other):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    x = np.random.default_rng(5).normal(size=a.dim) + 1j * np.random.default_rng(6).normal(size=a.dim)
    ys = [qs.SectorMatvec(H, s, lane=qs.LANE_GATHER).apply_host(x, complex_vectors=True) for s in (a, b, c)]
    assert np.array_equal(ys[0].view(np.uint8), ys[1].view(np.uint8))
    assert np.array_equal(ys[0].view(np.uint8), ys[2].view(np.uint8))
    if n_up >= 0:
        states = np.array([s for s in range(1 << N) if bin(s).count("1") == n_up], np.uint64)[::stride]
    else:
        states = np.arange(1 << N, dtype=np.uint64)[::stride]
    if n_up >= 0:   # states of another magnetisation find no index
        states = np.concatenate([states, np.array([(1 << (n_up + 1)) - 1, (1 << (n_up - 1)) - 1], np.uint64)])
    ja, pa = a.index_and_projection(states)
    for s in (b, c):
        jb, pb = s.index_and_projection(states)
        assert np.array_equal(ja, jb) and np.array_equal(np.asarray(pa).view(np.uint8), np.asarray(pb).view(np.uint8))
    assert (ja >= 0).sum() > 0
    if n_up >= 0:
        assert np.all(ja[-2:] == -1)

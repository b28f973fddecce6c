"""The merged CSR built on the device, on the shapes the reference fixtures leave out: rows with more
than 64 acting groups (two entries per lane in the row sort), rows whose equal columns cancel to an
exact zero (the entry is dropped), and the sublattice key order on such rows. Every case against the
oracle's merged CSR (the reference's sector_rows restated, itself pinned bit for bit to the
reference's own CSR on the fixtures), bit for bit."""

import sys

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import symmetry as sym

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def all_to_all(N, seed):
    """Heisenberg couplings on every pair with random weights from a small set (so that values
    repeat and the dictionary engages), plus a field-free XY part on every third pair."""
    rng = np.random.default_rng(seed)
    b = qs.HamiltonianBuilder(N)
    for i in range(N):
        for j in range(i + 1, N):
            b.heisenberg([(i, j)], float(rng.choice([0.25, 0.5, -0.75, 1.0])))
    return b.to_operator()


def cancelling(N, sign=-1.0):
    """S_0.S_1 - S_1.S_2: in a translation sector the two bonds' images of a state are often
    translates of each other, so their entries meet in one column and cancel exactly (sign = +1:
    the same entries add up instead)."""
    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(0, 1)], 1.0)
    b.heisenberg([(1, 2)], sign)
    return b.to_operator()


def check(op, g, n_up, sublattice=-1):
    s = qs.Sector(g, n_up, sublattice=sublattice)
    kb, L = s.key_bits()
    o = po.OracleSector(g.N, n_up, g.perms, g.chars, key_bit=kb)
    reps, inv = s.export()
    assert np.array_equal(reps, o.reps) and np.array_equal(bits(inv), bits(o.inv_norms))
    prog = qs.Program(op)
    m = qs.SectorMatvec(op, s, lane=qs.LANE_CSR)
    row_ptr, col, val = m.export_csr()
    rp, oc, ov = po.sector_csr(prog.export(), o)
    assert np.array_equal(row_ptr.astype(np.int64), rp)
    assert np.array_equal(col.astype(np.int64), oc)
    if m.is_real:
        assert np.array_equal(bits(val.real), bits(ov.real))
    else:
        assert np.array_equal(bits(val.view(np.float64)), bits(ov.view(np.float64)))
    return prog, L, int(rp[-1])


@pytest.mark.parametrize("N,k,flip_sym", [(12, 0, False), (12, 1, False), (12, 0, True), (13, 2, False)])
def test_more_than_64_groups(N, k, flip_sym):
    say("csr all-to-all", N, k, flip_sym)
    op = all_to_all(N, seed=N + k)
    g = sym.dihedral_group(N, 0, +1) if flip_sym else sym.chain_group(N, k)
    prog, _, nnz = check(op, g, N // 2)
    assert prog.counts()[0] > 64 and nnz > 0


def test_more_than_64_groups_key_order():
    say("csr all-to-all sublattice")
    N = 16
    op = all_to_all(N, seed=3)
    prog, L, _ = check(op, sym.dihedral_group(N, 0, +1), N // 2, sublattice=1)
    assert L > 0 and prog.counts()[0] > 64


@pytest.mark.parametrize("N,k", [(10, 0), (12, 3), (14, 0)])
def test_exact_cancellations_dropped(N, k):
    say("csr cancellations", N, k)
    g = sym.chain_group(N, k)
    _, _, nnz_minus = check(cancelling(N, -1.0), g, N // 2)
    _, _, nnz_plus = check(cancelling(N, +1.0), g, N // 2)
    assert nnz_minus < nnz_plus   # columns met and cancelled to exact zeros, which are dropped


def j123(N):
    """Heisenberg couplings at distances 1, 2 and 3 (3 N bonds: rows of up to 3 N + 1 entries)."""
    b = qs.HamiltonianBuilder(N)
    for d, J in ((1, 1.0), (2, 0.45), (3, -0.2)):
        b.heisenberg([(i, (i + d) % N) for i in range(N)], J)
    return b.to_operator()


@pytest.mark.parametrize("model,stage_bytes", [("cancelling", 100_000), ("j123", 600_000)])
def test_chunked_build_equals_one_walk(model, stage_bytes, monkeypatch):
    """A stage too small for the whole sector: the rows go in chunks (two passes: lengths and
    dictionary, then the walk again and the compaction into place). The CSR equals the one-chunk
    build bit for bit, and the oracle's."""
    say("csr chunked", model, stage_bytes)
    N = 20
    g = sym.chain_group(N, 0)
    s = qs.Sector(g, N // 2)
    H = cancelling(N) if model == "cancelling" else j123(N)
    groups = qs.Program(H).counts()[0]
    assert s.dim * groups * 20 > 4 * stage_bytes   # several chunks
    one = qs.SectorMatvec(H, s, lane=qs.LANE_CSR).export_csr()
    monkeypatch.setenv("QED_SPIN_AMD_CSR_STAGE_BYTES", str(stage_bytes))
    chunked = qs.SectorMatvec(H, s, lane=qs.LANE_CSR).export_csr()
    for x, y in zip(one, chunked):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    check(H, g, N // 2)

"""The fused observable sweep (qedh_matrix_elements, Python matrix_elements): every observable
and vector pair in one walk of the bra sector, against the dense oracle E_t^dag O E_s of each pair and
against the per-observable device path (apply, then dot), within 1e-12 relative; one sector (two
handles, observables not group invariant) and two sectors (n_up + 1, momentum k -> k'); the
expectation values of the levels verb through it. Every case prints its name first."""

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


def observables(N):
    ops = []
    for j in range(1, N // 2 + 1):   # S_0 . S_j
        O = qs.Operator(N)
        O.add_product("zz", (0, j), 1.0)
        O.add_product("+-", (0, j), 0.5)
        O.add_product("-+", (0, j), 0.5)
        ops.append(O)
    for j in range(N):               # S^z_j
        ops.append(qs.Operator(N).add_product("z", (j,), 1.0))
    chi = qs.Operator(N)             # a complex, not group-invariant three-spin term
    chi.add_product("+-z", (0, 1, 2), 0.5j)
    chi.add_product("-+z", (0, 1, 2), -0.5j)
    ops.append(chi)
    rnd = qs.Operator(N)
    rng = np.random.default_rng(N)
    for j in range(N):
        rnd.add_product("+-", (j, (j + 3) % N), complex(rng.normal(), rng.normal()))
        rnd.add_product("-+", (j, (j + 3) % N), complex(rng.normal(), rng.normal()))
    ops.append(rnd)
    return ops


@pytest.mark.parametrize("N,n_up,k", [(10, 5, 0), (10, 4, 3), (12, 6, 5)])
def test_sweep_one_sector_against_dense_and_per_observable(N, n_up, k):
    say("observable sweep, one sector", N, n_up, k)
    g = sym.chain_group(N, k)
    sec = qs.Sector(g, n_up)
    reps, inv = sec.export()
    twin = qs.Sector(g, n_up, reps=reps, inv_norms=inv)
    ops = observables(N)
    rng = np.random.default_rng(N + k)
    vecs = [rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim) for _ in range(3)]
    dv = [qs.DeviceArray.from_host(v) for v in vecs]
    pairs = [(0, 0), (1, 2), (2, 1), (1, 1)]
    got = qs.matrix_elements(ops, twin, sec, [dv[a] for a, _ in pairs], [dv[b] for _, b in pairs])
    states = np.arange(1 << N, dtype=np.uint64)
    E = po.sector_basis_matrix(po.OracleSector(N, n_up, g.perms, g.chars), states)
    for o, O in enumerate(ops):
        A = E.conj().T @ po.dense_matrix(O.products, N, states) @ E
        mv = qs.SectorMatvec(O, twin, sec, lane=qs.LANE_GATHER)
        for p, (a, b) in enumerate(pairs):
            want = np.vdot(vecs[a], A @ vecs[b])
            scale = np.linalg.norm(vecs[a]) * np.linalg.norm(vecs[b]) * max(1.0, np.linalg.norm(A, 2))
            assert abs(got[p, o] - want) <= 1e-12 * scale, (o, p, got[p, o], want)
            y = mv.apply_host(vecs[b], complex_vectors=True)
            assert abs(got[p, o] - np.vdot(vecs[a], y)) <= 1e-12 * scale, (o, p)


def test_sweep_between_two_sectors():
    say("observable sweep, two sectors")
    N = 10
    ss = qs.Sector(sym.chain_group(N, 2), 4)
    st = qs.Sector(sym.chain_group(N, 5), 5)
    ops = []
    for m in range(N):   # S^+(q) for every q: only q taking momentum 2 to 5 survives the projection
        O = qs.Operator(N)
        for j in range(N):
            O.add_product("+", (j,), np.exp(-2j * np.pi * m * j / N) / np.sqrt(N))
        ops.append(O)
    rng = np.random.default_rng(4)
    kets = [rng.normal(size=ss.dim) + 1j * rng.normal(size=ss.dim) for _ in range(2)]
    bras = [rng.normal(size=st.dim) + 1j * rng.normal(size=st.dim) for _ in range(2)]
    got = qs.matrix_elements(ops, st, ss, [qs.DeviceArray.from_host(b) for b in bras],
                             [qs.DeviceArray.from_host(k) for k in kets])
    states = np.arange(1 << N, dtype=np.uint64)
    Es = po.sector_basis_matrix(po.OracleSector(N, 4, sym.translations(N), sym.momentum_chars(N, 2)), states)
    Et = po.sector_basis_matrix(po.OracleSector(N, 5, sym.translations(N), sym.momentum_chars(N, 5)), states)
    nonzero = 0
    for o, O in enumerate(ops):
        A = Et.conj().T @ po.dense_matrix(O.products, N, states) @ Es
        for p in range(2):
            want = np.vdot(bras[p], A @ kets[p])
            scale = np.linalg.norm(bras[p]) * np.linalg.norm(kets[p]) * max(1.0, np.abs(A).max())
            assert abs(got[p, o] - want) <= 1e-12 * scale, (o, p)
            nonzero += abs(want) > 1e-8
    assert nonzero == 2   # only q = 2 pi 7/10 (k_t = k_s - m) connects them


def test_levels_expect_through_the_sweep():
    say("levels expect via the sweep")
    from qed_spin_amd import eigs as ev

    N = 10
    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    b.heisenberg([(i, (i + 2) % N) for i in range(N)], 0.3)
    H = b.to_operator()
    r = ev.eigs(H, 2, sym=qs.Symmetry(sz=[5]), vectors=True)
    ops = observables(N)
    vals = r.expect(ops)
    # the dense ground state of the Sz sector
    states = np.array([s for s in range(1 << N) if bin(s).count("1") == 5], np.uint64)
    Hd = po.dense_matrix(H.products, N, states)
    w, V = np.linalg.eigh(Hd)
    assert abs(w[1] - w[0]) > 1e-6   # a nondegenerate ground state
    g = V[:, 0]
    for o, O in enumerate(ops):
        want = np.vdot(g, po.dense_matrix(O.products, N, states) @ g)
        assert abs(vals[0][o] - want) <= 1e-10, (o, vals[0][o], want)


def test_sweep_is_reproducible_and_chunks_pairs():
    """The sweep's sums have a fixed order (a xor tree per wavefront, rows in grid-stride order,
    workgroups in order): two calls give the same bits. 1100 pairs take two chunks of pairs (at most
    1024 pair-observable cells per walk), each pair against the per-pair result of a one-pair call."""
    say("observable sweep: bits and pair chunks")
    N = 12
    g = sym.chain_group(N, 0)
    sec = qs.Sector(g, N // 2)
    ops = observables(N)[:3]
    rng = np.random.default_rng(9)
    vecs = [qs.DeviceArray.from_host(rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim)) for _ in range(4)]
    pairs = [(a % 4, (a * 7 + 1) % 4) for a in range(1100)]
    bras, kets = [vecs[a] for a, _ in pairs], [vecs[b] for _, b in pairs]
    one = qs.matrix_elements(ops, sec, sec, bras, kets)
    two = qs.matrix_elements(ops, sec, sec, bras, kets)
    assert np.array_equal(np.asarray(one).view(np.uint64), np.asarray(two).view(np.uint64))
    for p in (0, 1, 513, 1023, 1024, 1099):
        a, b = pairs[p]
        single = qs.matrix_elements(ops, sec, sec, [vecs[a]], [vecs[b]])
        assert np.array_equal(np.asarray(single[0]).view(np.uint64), np.asarray(one[p]).view(np.uint64)), p

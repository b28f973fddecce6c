"""GPU parity of the batched Krylov engine and the thermal verb.

- batch_lanczos (all blocks x R vectors in one packed multi-vector matvec per step) against the
  one-vector LocalDGKS3 Lanczos of each (block, vector) pair: equal steps, alpha and beta of the
  first 12 steps within 1e-12 (later steps of a recurrence without full reorthogonalisation
  amplify rounding differences of the reduction order exponentially), the extreme Ritz values of
  the whole run within 1e-9;
  graph replay equals eager launches bit for bit, and two graph runs agree bit for bit;
- ftlm_batched against the per-block FTLM driver at equal seeds (1e-8, the FTLM tolerance), also
  with more samples than one batch holds;
- exact_thermal against numpy's dense eigensolve of the exported block;
- thermal() against the reference's own recorded curves of j1j2_chain12 (FTLM of the 4096-state
  block at seed 11; exact curves over every (Sz, k) block), and a mixed walk (exact small blocks,
  sampled large ones) against the oracle block by block.

Every case prints its name before it touches the device."""

import json
import os
import sys

import numpy as np
import pytest

from conftest import FIXTURES
from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import symmetry as sym
from qed_spin_amd import thermal as th

pytestmark = pytest.mark.gpu

EPS64 = 64 * np.finfo(np.float64).eps


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ring(N, d=1):
    return [(i, (i + d) % N) for i in range(N)]


def j1j2(N, J2=0.5):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg(ring(N), 1.0)
    b.heisenberg(ring(N, 2), J2)
    return b.to_operator()


def xxz(N, delta=0.5):
    return qs.HamiltonianBuilder(N).xxz(ring(N), 1.0, delta).to_operator()


def chiral_j1j2(N, chi=0.2):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg(ring(N), 1.0)
    b.heisenberg(ring(N, 2), 0.3)
    b.chirality([(i, (i + 1) % N, (i + 2) % N) for i in range(N)], chi)
    return b.to_operator()


def blocks(H, N, pairs):
    out = []
    for n_up, k in pairs:
        s = qs.Sector(sym.chain_group(N, k), n_up)
        out.append(qs.SectorMatvec(H, s, lane=qs.LANE_CSR))
    return out


def records():
    with open(os.path.join(FIXTURES, "reference_records.json")) as f:
        return json.load(f)


def test_batch_lanczos_matches_one_vector_lanczos():
    say("batch_lanczos vs lanczos LocalDGKS3")
    N, M, R = 12, 30, 4
    H = chiral_j1j2(N)
    mats = blocks(H, N, [(6, 0), (6, 1), (5, 2), (2, 1), (6, 5)])
    dims = [m.dim for m in mats]
    assert min(dims) < M < max(dims)
    rows = sum(dims)
    rng = np.random.default_rng(3)
    start = rng.normal(size=(rows, R)) + 1j * rng.normal(size=(rows, R))
    tol = [EPS64 * H.norm_bound()] * len(mats)
    alpha, beta, steps, _ = th.batch_lanczos(mats, R, M, start, tol)
    off = 0
    for b, m in enumerate(mats):
        for s in range(R):
            a1, b1, _ = m.lanczos(start[off:off + m.dim, s], M, reorth=2, breakdown_tol=tol[b], complex_vectors=True)
            n = len(a1)
            assert steps[b, s] == n, (b, s, steps[b, s], n)
            scale = max(1.0, np.abs(a1).max())
            h = min(n, 12)
            assert np.abs(alpha[:h, b, s] - a1[:h]).max() <= 1e-12 * scale, (b, s)
            assert np.abs(beta[:h, b, s] - b1[1:h + 1]).max() <= 1e-12 * scale, (b, s)
            r0 = po.tridiag_eig(alpha[:n, b, s], np.concatenate([[0.0], beta[:n, b, s]]))[0]
            r1 = po.tridiag_eig(a1, b1[: n + 1])[0]
            assert abs(r0[0] - r1[0]) <= 1e-9 * scale and abs(r0[-1] - r1[-1]) <= 1e-9 * scale, (b, s)
        off += m.dim


def test_batch_lanczos_graph_equals_eager_bitwise():
    say("batch_lanczos graph vs eager")
    N, M, R = 12, 25, 3
    H = xxz(N)
    mats = blocks(H, N, [(6, 0), (5, 3), (4, 7)])
    rows = sum(m.dim for m in mats)
    rng = np.random.default_rng(5)
    start = rng.normal(size=(rows, R)) + 1j * rng.normal(size=(rows, R))
    tol = [EPS64 * H.norm_bound()] * len(mats)
    ag, bg, sg, _ = th.batch_lanczos(mats, R, M, start, tol, use_graph=True)
    ae, be_, se, _ = th.batch_lanczos(mats, R, M, start, tol, use_graph=False)
    ag2, bg2, sg2, _ = th.batch_lanczos(mats, R, M, start, tol, use_graph=True)
    for x, y in ((ag, ae), (bg, be_), (ag, ag2), (bg, bg2)):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    assert np.array_equal(sg, se) and np.array_equal(sg, sg2)


@pytest.mark.parametrize("base,first,count,n", [(11, 0, 3, 1), (11, 2, 2, 155), (7, 0, 4, 157), (2**63 + 5, 37, 3, 100003)])
def test_device_gaussian_matches_library_draw(base, first, count, n):
    """The FTLM start vectors drawn on the device against std::mt19937 -> std::normal_distribution
    on the host: every entry within 1e-14 relative (the device logarithm may differ from the host's
    in the last bit; every other step is exact), most of them bit for bit."""
    from qed_spin_amd import _lib
    from qed_spin_amd.sector import DeviceArray

    say("device gaussian", base, first, count, n)
    d = DeviceArray(count * n)
    _lib.call("qedh_gaussian_vectors", base, first, count, n, d.ptr)
    got = d.to_host().reshape(count, n)
    exact = 0
    for c in range(count):
        want = po.gaussian_raw(base, first + c, n)
        for part in (np.real, np.imag):
            g, w = part(got[c]), part(want)
            assert np.all(np.abs(g - w) <= 1e-14 * np.abs(w)), (c, np.max(np.abs(g - w)))
            exact += int(np.sum(g == w))
    assert exact >= count * n   # at least half of the 2 count n components bit for bit

@pytest.mark.parametrize("samples", [6, 70])
def test_ftlm_batched_matches_per_block_driver(samples):
    say("ftlm_batched vs ftlm, samples", samples)
    N = 10
    H = chiral_j1j2(N)
    mats = blocks(H, N, [(5, 0), (5, 3), (4, 1), (4, 6)])
    betas = 1.0 / np.linspace(0.2, 3.0, 7)
    krylov = 18
    seeds = [th.next_block_seed(7 + i) for i in range(len(mats))]
    tol = [EPS64 * H.norm_bound()] * len(mats)
    lnZ, E, V, emin, stats = th.ftlm_batched(mats, samples, krylov, seeds, betas, breakdown_tol=tol)
    assert stats["rows"] == sum(m.dim for m in mats)
    for b, m in enumerate(mats):
        l1, e1, v1, em1 = m.ftlm(samples, krylov, seeds[b], betas, breakdown_tol=tol[b])
        assert np.allclose(lnZ[b], l1, rtol=1e-8, atol=1e-8)
        assert np.allclose(E[b], e1, rtol=1e-8, atol=1e-8)
        assert np.allclose(V[b], v1, rtol=1e-8, atol=1e-8)
        assert abs(emin[b] - em1) <= 1e-8 * max(1.0, abs(em1))


def test_exact_thermal_matches_dense_eigensolve():
    say("exact_thermal vs numpy eigvalsh")
    N = 12
    H = chiral_j1j2(N)
    mats = blocks(H, N, [(6, 0), (6, 5), (3, 2), (1, 0)])
    betas = 1.0 / np.linspace(0.25, 3.0, 6)
    lnZ, E, V, emin = th.exact_thermal(mats, betas)
    for b, m in enumerate(mats):
        row_ptr, col, val = m.export_csr()
        A = np.zeros((m.dim, m.dim), np.complex128)
        for r in range(m.dim):
            for e in range(int(row_ptr[r]), int(row_ptr[r + 1])):
                A[r, col[e]] += val[e]
        assert np.allclose(A, A.conj().T, atol=1e-13)
        ev = np.linalg.eigvalsh(A)
        e0 = ev.min()
        assert abs(emin[b] - e0) <= 1e-10 * max(1.0, abs(e0))
        for i, bt in enumerate(betas):
            w = np.exp(-bt * (ev - e0))
            Z = w.sum()
            Em = (w * ev).sum() / Z
            Vm = (w * (ev - Em) ** 2).sum() / Z
            assert abs(lnZ[b, i] - (np.log(Z) - bt * e0)) <= 1e-10 * max(1.0, abs(lnZ[b, i]))
            assert abs(E[b, i] - Em) <= 1e-10 * max(1.0, abs(Em))
            assert abs(V[b, i] - Vm) <= 1e-10 * max(1.0, abs(Vm))


def test_thermal_reproduces_reference_ftlm_record():
    say("thermal ftlm none vs reference record")
    rec = records()
    r = rec["records"]["j1j2_chain12/api/thermal/ftlm/none"]["values"]
    res = th.thermal(j1j2(12), np.array(r["T"]), sym=qs.Symmetry.none(), samples=8, krylov=40, seed=11)
    assert len(res.blocks) == 1 and res.blocks[0]["dim"] == 4096 and not res.blocks[0]["exact"]
    for key, got in (("E", res.E), ("C", res.C), ("S", res.S)):
        err = np.abs(got - np.array(r[key])).max()
        print(f"  {key}: max |diff| {err:.3e}")
        assert np.allclose(got, r[key], rtol=1e-8, atol=1e-8), key


def test_thermal_reproduces_reference_exact_record():
    say("thermal exact over (Sz, k) blocks vs reference record")
    rec = records()
    r = rec["records"]["j1j2_chain12/api/thermal/exact/auto"]["values"]
    res = th.thermal(j1j2(12), np.array(r["T"]), sym=qs.Symmetry(), samples=4, krylov=20, seed=3)
    assert all(b["exact"] for b in res.blocks) and res.total_dim == 4096
    for key, got in (("E", res.E), ("C", res.C), ("S", res.S)):
        assert np.allclose(got, r[key], rtol=1e-10, atol=1e-10), key


def _oracle_block(H_products, N, n_up, k):
    """The block matrix from the dense oracle: E^dagger H E on the brute-force sector."""
    s = po.OracleSector(N, n_up, sym.translations(N), sym.momentum_chars(N, k))
    states = np.array([x for x in range(1 << N) if bin(x).count("1") == n_up], np.uint64)
    Hd = po.dense_matrix(H_products, N, states)
    Em = po.sector_basis_matrix(s, states)
    return s, Em.conj().T @ Hd @ Em


def test_thermal_mixed_walk_matches_oracle():
    say("thermal mixed exact/sampled walk vs oracle")
    N, dmax, samples, krylov, seed = 10, 12, 5, 15, 7
    H = xxz(N)
    T = np.linspace(0.3, 3.0, 5)
    betas = 1.0 / T
    res = th.thermal(H, T, sym=qs.Symmetry(), samples=samples, krylov=krylov, seed=seed, dense_max_dim=dmax)
    assert res.total_dim == 1 << N
    assert any(b["exact"] for b in res.blocks) and any(not b["exact"] for b in res.blocks)
    # the oracle, block by block in the same walk order with the same seed stream; H is real, so
    # the momentum-k block with 2k > N is the time reversal of the (N - k) block and folds into it
    seed_b = seed
    lnZ, E, V, wts = [], [], [], []
    for n_up in range(N + 1):
        for k in range(N):
            if 2 * k > N:
                continue
            s, A = _oracle_block(H.products, N, n_up, k)
            if s.dim == 0:
                continue
            wts.append(1.0 if k in (0, N // 2) else 2.0)
            seed_b = th.next_block_seed(seed_b)
            if s.dim <= dmax:
                ev = np.linalg.eigvalsh(A)
                e0 = ev.min()
                w = np.exp(-np.outer(betas, ev - e0))
                Z = w.sum(1)
                Em = (w * ev).sum(1) / Z
                lnZ.append(np.log(Z) - betas * e0)
                E.append(Em)
                V.append((w * (ev[None, :] - Em[:, None]) ** 2).sum(1) / Z)
                continue
            rp = np.zeros(s.dim + 1, np.int64)
            cols, vals = [], []
            for r_ in range(s.dim):
                nz = np.nonzero(A[r_])[0]
                cols += list(nz)
                vals += list(A[r_, nz])
                rp[r_ + 1] = len(cols)
            cols, vals = np.array(cols, np.int64), np.array(vals, np.complex128)
            mom = []
            for q in range(samples):
                v0 = po.gaussian_vector(seed_b, q, s.dim)
                a, b = po.lanczos(rp, cols, vals, v0, krylov, 2, EPS64 * H.norm_bound())
                ritz, Vv = po.tridiag_eig(a, b)
                mom.append(po.ftlm_moments(ritz, Vv[0] ** 2, betas))
            l_, e_, v_, _ = po.ftlm_combine(np.array(mom), s.dim, betas)
            lnZ.append(l_)
            E.append(e_)
            V.append(v_)
    lnZ, E, V, wts = np.array(lnZ), np.array(E), np.array(V), np.array(wts)
    assert lnZ.shape[0] == len(res.blocks)
    for i, b in enumerate(res.blocks):
        assert b["weight"] == wts[i], b["label"]
        assert np.allclose(b["lnZ"], lnZ[i], rtol=1e-8, atol=1e-8), b["label"]
        assert np.allclose(b["E"], E[i], rtol=1e-8, atol=1e-8), b["label"]
    mx = lnZ.max(0)
    p = wts[:, None] * np.exp(lnZ - mx)
    z = p.sum(0)
    Et = (p * E).sum(0) / z
    var = (p * (V + (E - Et) ** 2)).sum(0) / z
    assert np.allclose(res.lnZ, mx + np.log(z), rtol=1e-8, atol=1e-8)
    assert np.allclose(res.E, Et, rtol=1e-8, atol=1e-8)
    assert np.allclose(res.C, betas ** 2 * var, rtol=1e-8, atol=1e-8)


def test_cross_sector_matvec_matches_dense_oracle():
    """O from sector s to sector t (both lanes) against E_t^dagger O E_s of the dense oracle, for
    S^z(q) (same n_up) and S^+(q) (one more up spin); N^-1/2 sum_j e^{-i q j} O_j takes momentum
    index k to k - m (q = 2 pi m / N)."""
    from qed_spin_amd import dynamics as dy

    say("cross-sector matvec vs dense")
    N = 10
    states = np.arange(1 << N, dtype=np.uint64)
    splus = qs.Operator(N)
    for j in range(N):
        splus.add_product("+", (j,), np.exp(-2j * np.pi * 3 * j / N) / np.sqrt(N))
    cases = [(dy.sz_q(N, 2 * np.pi * 2 / N), 5, 1, 5, 9), (dy.sz_q(N, np.pi), 4, 0, 4, 5), (splus, 4, 2, 5, 9)]
    rng = np.random.default_rng(9)
    for O, nu_s, k_s, nu_t, k_t in cases:
        ss = qs.Sector(sym.chain_group(N, k_s), nu_s)
        st = qs.Sector(sym.chain_group(N, k_t), nu_t)
        Es = po.sector_basis_matrix(po.OracleSector(N, nu_s, sym.translations(N), sym.momentum_chars(N, k_s)), states)
        Et = po.sector_basis_matrix(po.OracleSector(N, nu_t, sym.translations(N), sym.momentum_chars(N, k_t)), states)
        A = Et.conj().T @ po.dense_matrix(O.products, N, states) @ Es
        x = rng.normal(size=ss.dim) + 1j * rng.normal(size=ss.dim)
        want = A @ x
        assert np.linalg.norm(want) > 1e-3
        for lane in (qs.LANE_GATHER, qs.LANE_CSR):
            X = qs.SectorMatvec(O, st, ss, lane=lane)
            got = X.apply_host(x, complex_vectors=True)
            assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want), (nu_s, k_s, nu_t, k_t, lane)


def _lehmann(Hd, O, omega, eta, window=1e-8):
    """The T = 0 spectral function of a dense H (the oracle): ground manifold average of
    sum_n |<n|O|g>|^2 eta / pi / ((omega - E_n + E0)^2 + eta^2)."""
    ev, U = np.linalg.eigh(Hd)
    e0 = ev[0]
    g = [i for i in range(len(ev)) if ev[i] <= e0 + window]
    S = np.zeros(len(omega))
    for i in g:
        amp = U.conj().T @ (O @ U[:, i])
        w = np.abs(amp) ** 2
        S += (w[None, :] * (eta / np.pi) / ((omega[:, None] - (ev[None, :] - e0)) ** 2 + eta ** 2)).sum(1)
    return S / len(g), e0, len(g)


def test_dynamics_T0_matches_dense_lehmann_sum():
    """A complex (chiral) ring, every (Sz, k) block, S^z(q) probes: the batched continued fractions
    against the Lehmann sum of the dense oracle (krylov above every target dimension)."""
    from qed_spin_amd import dynamics as dy

    say("dynamics T=0 vs dense Lehmann sum")
    N = 10
    H = chiral_j1j2(N, 0.25)
    omega = np.linspace(-0.5, 4.0, 61)
    eta = 0.1
    probes = [dy.sz_q(N, 2 * np.pi * m / N) for m in (0, 2, 5)]
    r = dy.dynamics(H, probes, omega, eta=eta, sym=qs.Symmetry(sz=[N // 2]), krylov=120)
    states = np.array([x for x in range(1 << N) if bin(x).count("1") == N // 2], np.uint64)
    Hd = po.dense_matrix(H.products, N, states)
    for p, O in enumerate(probes):
        Od = po.dense_matrix(O.products, N, states)
        want, e0, ng = _lehmann(Hd, Od, omega, eta)
        assert abs(r.e0 - e0) < 1e-10 and r.ground_manifold == ng
        assert np.abs(r.S[p] - want).max() <= 1e-8 * max(1.0, np.abs(want).max()), p


def test_dynamics_reproduces_reference_T0_records():
    from qed_spin_amd import dynamics as dy

    say("dynamics T=0 vs reference records")
    rec = records()
    dd = rec["dynamics"]
    omega = np.array(dd["omega"])
    N = 12
    O = qs.Operator(N)
    for i in range(N):
        O.add_product("z", (i,), np.exp(1j * np.pi * i) / np.sqrt(N))
    H = j1j2(N)
    for label, symm in (("none", qs.Symmetry.none()), ("auto", qs.Symmetry())):
        want = np.array(rec["records"][f"j1j2_chain12/api/dynamics/T=0/{label}/SzQ"]["values"]["S"])
        r = dy.dynamics(H, O, omega, eta=dd["eta"], sym=symm, krylov=dd["krylov_T0"])
        err = np.abs(r.S[0] - want).max()
        print(f"  {label}: ground manifold {r.ground_manifold}, E0 {r.e0:.12f}, max |diff| {err:.3e}")
        assert np.allclose(r.S[0], want, rtol=1e-8, atol=1e-8), label


def _small_stream(n=18, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(dict(kind=i % 3, n_sites=int(rng.integers(8, 15)), j1=float(rng.uniform(0.5, 1.5)),
                        j2=float(rng.uniform(0.0, 0.6)), chi=float(rng.uniform(-0.3, 0.3)), arrival_s=0.01 * i,
                        k=3, samples=6, depth=20, krylov=40, seed=100 + i))
    return out


def test_serve_results_same_alone_and_in_the_mix():
    """Every request's result served in a mix of concurrent requests equals the same request
    served alone (one slot) bit for bit, and the plain path (no batching, no sector cache) within
    1e-8; eigs levels against the dense oracle."""
    from qed_spin_amd import serve as sv

    say("serve: mix vs alone vs plain")
    reqs = _small_stream()
    T = np.linspace(0.3, 3.0, 5)
    omega = np.linspace(0.0, 4.0, 50)
    mix = sv.serve(reqs, slots=6, T=T, omega=omega, eta=0.1, realtime=True)
    alone = sv.serve(reqs, slots=1, T=T, omega=omega, eta=0.1, realtime=False)
    plain = sv.serve(reqs, slots=1, batched=False, cache_sectors=False, T=T, omega=omega, eta=0.1, realtime=False)
    print("  mix stats", mix.stats)
    assert mix.stats["completed"] == len(reqs) and mix.stats["peak_in_flight"] > 1
    for i, r in enumerate(reqs):
        assert mix.status[i] == 0 and alone.status[i] == 0 and plain.status[i] == 0, i
        assert np.array_equal(mix.values[i].view(np.uint64), alone.values[i].view(np.uint64)), i
        assert np.allclose(mix.values[i], plain.values[i], rtol=1e-8, atol=1e-8), i
    # eigs against the dense oracle of the sector
    for i, r in enumerate(reqs):
        if r["kind"] != 0:
            continue
        N = r["n_sites"]
        b = qs.HamiltonianBuilder(N)
        b.heisenberg(ring(N), r["j1"])
        b.heisenberg(ring(N, 2), r["j2"])
        b.chirality([(j, (j + 1) % N, (j + 2) % N) for j in range(N)], r["chi"])
        H = b.to_operator()
        s = po.OracleSector(N, N // 2, sym.translations(N), sym.momentum_chars(N, 0))
        states = np.array([x for x in range(1 << N) if bin(x).count("1") == N // 2], np.uint64)
        E = po.sector_basis_matrix(s, states)
        ev = np.linalg.eigvalsh(E.conj().T @ po.dense_matrix(H.products, N, states) @ E)
        k = len(mix.values[i])
        assert np.allclose(mix.values[i], ev[:k], rtol=1e-10, atol=1e-10), i


def test_dynamics_finite_T_reproduces_reference_record():
    """FTLM dynamics at T = 1 of j1j2_chain12 with no symmetry (one 4096-state source, 6 samples,
    krylov 40, seed 11: the source's seed is block_seed(11, 0)) against the reference's record."""
    from qed_spin_amd import dynamics as dy

    say("dynamics T=1 vs reference record")
    rec = records()
    dd = rec["dynamics"]
    omega = np.array(dd["omega"])
    N = 12
    O = qs.Operator(N)
    for i in range(N):
        O.add_product("z", (i,), np.exp(1j * np.pi * i) / np.sqrt(N))
    want = np.array(rec["records"]["j1j2_chain12/api/dynamics/T=1/none/SzQ"]["values"]["S"])
    r = dy.dynamics(j1j2(N), O, omega, eta=dd["eta"], T=[1.0], sym=qs.Symmetry.none(), krylov=dd["krylov_T"],
                    samples=dd["samples"], seed=dd["seed"])
    got = r.S[0][0]
    err = np.abs(got.real - want).max()
    print(f"  max |diff| {err:.3e}, max |imag| {np.abs(got.imag).max():.3e}")
    assert np.allclose(got.real, want, rtol=1e-8, atol=1e-8)


def test_dynamics_finite_T_symmetric_walk_matches_oracle():
    """The same estimator over every momentum sector of Sz = 0 (N = 10, S^z(q) probes) against the
    oracle's restatement of the kernel on dense block matrices with the same seeds."""
    from qed_spin_amd import dynamics as dy

    say("dynamics T>0 walk vs oracle")
    N, samples, krylov, seed, eta = 10, 3, 12, 5, 0.2
    H = chiral_j1j2(N, 0.2)
    omega = np.linspace(-1.0, 4.0, 21)
    Ts = np.array([0.7, 1.5])
    probes = [dy.sz_q(N, 2 * np.pi * 3 / N)]
    r = dy.dynamics(H, probes, omega, eta=eta, T=Ts, sym=qs.Symmetry(sz=[N // 2]), krylov=krylov, samples=samples,
                    seed=seed)
    states = np.array([x for x in range(1 << N) if bin(x).count("1") == N // 2], np.uint64)
    Hd = po.dense_matrix(H.products, N, states)
    Od = po.dense_matrix(probes[0].products, N, states)
    Es = [po.sector_basis_matrix(po.OracleSector(N, N // 2, sym.translations(N), sym.momentum_chars(N, k)), states)
          for k in range(N)]
    Hb = [E.conj().T @ Hd @ E for E in Es]
    tol = EPS64 * H.norm_bound()
    per = []
    for k in range(N):
        n = Es[k].shape[1]
        key = ((N // 2 + 1) * 3) * N + k
        bseed = dy.block_seed(seed, key)
        targets = []
        for t in range(N):
            A = Es[t].conj().T @ Od @ Es[k]
            if np.abs(A).max() > 1e-12:
                targets.append((t, A))
        Zs, Ss, smins = [], [], []
        for s in range(samples):
            v0 = po.gaussian_vector(bseed, s, n)
            a, b, V = _lanczos_basis(Hb[k], v0, min(n, krylov), tol)
            ev, U = np.linalg.eigh(np.diag(a) + np.diag(b[1:len(a)], 1) + np.diag(b[1:len(a)], -1))
            c = U[0]
            smin = ev.min()
            wt = np.exp(-np.outer(1.0 / Ts, ev - smin))
            Z = (wt * c ** 2).sum(1)
            S = np.zeros((len(Ts), omega.size), np.complex128)
            for t, A in targets:
                phi = A @ v0
                nphi = np.linalg.norm(phi)
                if nphi < 1e-14:
                    continue
                a2, b2, V2 = _lanczos_basis(Hb[t], phi / nphi, min(Hb[t].shape[0], krylov), tol)
                lam, US = np.linalg.eigh(np.diag(a2) + np.diag(b2[1:len(a2)], 1) + np.diag(b2[1:len(a2)], -1))
                W = (A @ V).conj().T @ V2                  # W[a, b] = <A v_a | w_b>
                Tm = U.T @ W                                 # [i, b]
                obar = Tm @ US                               # [i, j]
                w_ij = c[:, None] * obar * (nphi * US[0][None, :])
                E_ij = lam[None, :] - ev[:, None]
                L = (eta / np.pi) / ((omega[None, None, :] - E_ij[:, :, None]) ** 2 + eta ** 2)
                si = (w_ij[:, :, None] * L).sum(1)           # [i, w]
                S += wt @ si
            Zs.append(Z)
            Ss.append(S)
            smins.append(smin)
        E_min = min(smins)
        Zt = sum(np.exp(-(sm - E_min) / Ts) * Z for sm, Z in zip(smins, Zs)) * n / samples
        St = sum(np.exp(-(sm - E_min) / Ts)[:, None] * S for sm, S in zip(smins, Ss)) * n / samples
        per.append((E_min, Zt, St))
    ref = min(e for e, _, _ in per)
    want = np.zeros((len(Ts), omega.size), np.complex128)
    for it, Tv in enumerate(Ts):
        Zp = sum(np.exp(-(e - ref) / Tv) * Z[it] for e, Z, _ in per)
        want[it] = sum(np.exp(-(e - ref) / Tv) * S[it] for e, _, S in per) / Zp
    got = r.S[0]
    print(f"  max |diff| {np.abs(got - want).max():.3e}")
    assert np.allclose(got, want, rtol=1e-8, atol=1e-8)


def _lanczos_basis(A, v0, m, tol):
    """Lanczos with full classical Gram-Schmidt twice (the dense restatement): alpha, beta (beta[0]
    = 0), the basis as columns."""
    n = A.shape[0]
    V = np.zeros((n, m), np.complex128)
    alpha, beta = [], [0.0]
    v = v0 / np.linalg.norm(v0)
    V[:, 0] = v
    for j in range(m):
        w = A @ V[:, j]
        if j > 0:
            w = w - beta[j] * V[:, j - 1]
        a = np.vdot(V[:, j], w).real
        alpha.append(a)
        w = w - a * V[:, j]
        for _ in range(2):
            w = w - V[:, : j + 1] @ (V[:, : j + 1].conj().T @ w)
        b = np.linalg.norm(w)
        beta.append(b)
        if b < tol or j + 1 >= m:
            break
        V[:, j + 1] = w / b
    k = len(alpha)
    return np.array(alpha), np.array(beta), V[:, :k]

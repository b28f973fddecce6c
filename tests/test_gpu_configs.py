"""The benchmark configurations at their own sizes, checked by what the mathematics fixes, and the
reference's own full-size recorded results.

- FTLM of the XXZ ring N = 24 (Jxy = 1, Delta = 0.5) over every (Sz, k) block, 40 samples, depth 100,
  seed 7: the block dimensions (times their weights) add up to 2^24; every exactly solved block's
  curves equal the dense eigensolve of its CSR (1e-10); at T = 1e8 (beta -> 0) FTLM is exact:
  lnZ = 24 ln 2 up to beta <E>; C >= 0 everywhere.
- T = 0 S^z(q) of the Heisenberg ring N = 28 for all 28 q, krylov 200: S^z(0) annihilates the
  Sz = 0 ground state (no weight at q = 0); sum_q ||S^z(q)|0>||^2 = sum_j <(S^z_j)^2> = N / 4 = 7
  (1e-12); S >= 0.
- The served stream: the first 200 requests of the benchmark stream (N in 16..30, seed 3); every
  result in the mix equals the same request served alone bit for bit; eigs requests of N <= 18
  against the dense eigensolve of the oracle's CSR of their sector (1e-10).
- The reference's recorded bench results (bench/baseline/v1_cd08f64.jsonl, copied by
  tests/golden/extract_bench_baseline.py): chain32_abelian_eigs E0 (1e-10) and chain30_dyn0's
  spectral weight (2e-8: see its test).

Every case prints its name before it touches the device."""

import json
import os
import sys

import numpy as np
import pytest

from conftest import FIXTURES
from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import dynamics as dy
from qed_spin_amd import symmetry as sym
from qed_spin_amd import thermal as th

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ring(N, d=1):
    return [(i, (i + d) % N) for i in range(N)]


def _bench_records():
    with open(os.path.join(FIXTURES, "reference_bench_v1.json")) as f:
        return json.load(f)


def _ref_chain(N, J2=0.35):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg(ring(N), 1.0)
    b.heisenberg(ring(N, 2), J2)
    return b.to_operator()


def test_ftlm_xxz24_every_block():
    say("config: FTLM XXZ N=24, 40 samples, depth 100, seed 7")
    N = 24
    H = qs.HamiltonianBuilder(N).xxz(ring(N), 1.0, 0.5).to_operator()
    T = np.concatenate([np.linspace(0.1, 4.0, 40), [1e8]])
    r = th.thermal(H, T, sym=qs.Symmetry(), samples=40, krylov=100, seed=7)
    assert r.total_dim == 2 ** N
    assert sum(b["dim"] * b["weight"] for b in r.blocks) == 2 ** N
    assert np.all(r.C >= -1e-12)
    beta = 1.0 / T[-1]
    assert abs(r.lnZ[-1] - (N * np.log(2.0) - beta * r.E[-1])) <= 1e-9
    # the exactly solved blocks against numpy's dense eigensolve of their own CSR
    blocks = th.walk_blocks(H, qs.Symmetry(), tune=False, fold=True)
    exact = [i for i, b in enumerate(r.blocks) if b["exact"]]
    assert len(exact) > 50
    betas = 1.0 / T
    for i in exact:
        blk, sec, m = blocks[i]
        assert r.blocks[i]["dim"] == sec.dim
        row_ptr, col, val = m.export_csr()
        A = np.zeros((sec.dim, sec.dim), complex)
        for rr in range(sec.dim):
            A[rr, col[row_ptr[rr]:row_ptr[rr + 1]]] += val[row_ptr[rr]:row_ptr[rr + 1]]
        ev = np.linalg.eigvalsh(A)
        e0 = ev[0]
        w = np.exp(-np.outer(betas, ev - e0))
        lnZ = np.log(w.sum(axis=1)) - betas * e0
        assert np.allclose(r.blocks[i]["lnZ"], lnZ, rtol=1e-10, atol=1e-10), i
    th.release_later(blocks)


def test_dynamics_sz_q_heisenberg28_all_q():
    say("config: T=0 S^z(q) Heisenberg N=28, all 28 q, krylov 200")
    N = 28
    H = qs.HamiltonianBuilder(N).heisenberg(ring(N), 1.0).to_operator()
    omega = np.linspace(0.0, 4.0, 400)
    blocks = th.walk_blocks(H, qs.Symmetry(sz=[N // 2]))
    k0 = [i for i, (b, s, m) in enumerate(blocks) if b.momentum == 0]
    probes = [dy.sz_q(N, 2 * np.pi * m / N) for m in range(N)]
    r = dy.dynamics(H, probes, omega, eta=0.05, krylov=200, blocks=blocks, ground_blocks=k0)
    th.release_later(blocks)
    assert r.ground_manifold == 1
    assert r.norms[0] == 0.0 and not np.any(r.S[0])
    assert abs(float(np.sum(r.norms)) - N / 4) <= 1e-12 * N
    for m in range(N):
        assert np.all(r.S[m] >= -1e-12), m
    # S^z(q) and S^z(-q) carry the same weight in a reflection-symmetric ring
    for m in range(1, N // 2):
        assert abs(r.norms[m] - r.norms[N - m]) <= 1e-12


def test_served_stream_first_200_requests():
    from qed_spin_amd import serve as sv

    say("config: served stream, first 200 requests of seed 3")
    reqs = sv.mixed_stream(200, rate=5.0, seed=3)
    mix = sv.serve(reqs, slots=8, realtime=False)
    alone = sv.serve(reqs, slots=1, realtime=False)
    print("  mix stats", mix.stats)
    assert mix.stats["completed"] == len(reqs) and mix.stats["peak_in_flight"] > 1
    for i in range(len(reqs)):
        assert mix.status[i] == 0 and alone.status[i] == 0, i
        assert np.array_equal(mix.values[i].view(np.uint64), alone.values[i].view(np.uint64)), i
    checked = 0
    for i, q in enumerate(reqs):
        if q["kind"] != 0 or q["n_sites"] > 18:
            continue
        N = q["n_sites"]
        b = qs.HamiltonianBuilder(N)
        b.heisenberg(ring(N), q["j1"])
        b.heisenberg(ring(N, 2), q["j2"])
        b.chirality([(j, (j + 1) % N, (j + 2) % N) for j in range(N)], q["chi"])
        H = b.to_operator()
        o = po.OracleSector(N, N // 2, sym.translations(N), sym.momentum_chars(N, 0))
        row_ptr, col, val = po.sector_csr(qs.Program(H).export(), o)
        A = np.zeros((o.dim, o.dim), complex)
        for rr in range(o.dim):
            A[rr, col[row_ptr[rr]:row_ptr[rr + 1]]] += val[row_ptr[rr]:row_ptr[rr + 1]]
        ev = np.linalg.eigvalsh(A)
        k = len(mix.values[i])
        assert np.allclose(mix.values[i], ev[:k], rtol=1e-10, atol=1e-10), (i, N)
        checked += 1
    assert checked >= 10


def test_reference_chain32_abelian_eigs_e0():
    from qed_spin_amd import eigs as ev

    rec = _bench_records()["chain32_abelian_eigs_gpu"]["result"]["E0"]
    say("reference record: chain32_abelian_eigs E0", rec)
    r = ev.eigs(_ref_chain(32), 1, sym=qs.Symmetry(sz=[16], spin_flip="all"), stream=True)
    assert abs(r.energies[0] - rec) <= 1e-10 * abs(rec)


def test_reference_chain30_dyn0_weight():
    """The recorded weight within 2e-8. The weight is first order in the ground-state vector, which
    both codes converge to a residual of 1e-10 s_H (the lock tolerance), i.e. to
    about 1e-8 relative in the vector at this block's gap: two different eigensolvers of this
    package (a Krylov-Schur with host-drawn starts and the locking Lanczos with device starts) gave
    0.30024231633 and 0.30024231675, 1.4e-9 apart, and both lie 0.9e-8 .. 1.0e-8 above the record."""
    rec = _bench_records()["chain30_dyn0_cpu"]["result"]["weight"]
    say("reference record: chain30_dyn0 weight", rec)
    N = 30
    w = np.linspace(-1, 7, 401)
    r = dy.dynamics(_ref_chain(N), [dy.sz_q(N, 2 * np.pi * 10 / N)], w, eta=0.05, krylov=200,
                    sym=qs.Symmetry(sz=[15]))
    weight = float(np.trapezoid(r.S[0], w))
    print("  weight", weight, "relative difference", abs(weight - rec) / abs(rec))
    assert abs(weight - rec) <= 2e-8 * abs(rec)

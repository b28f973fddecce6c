"""GPU parity of the eigs verb: the k = 4 lowest levels of j1j2_chain12 against the reference's own
recorded levels (tests/python/golden/refs/api-2026-09/cpu.json.gz, copied to
tests/golden/fixtures/reference_records.json) within 1e-10, over the whole space, every (Sz, k)
block, the Sz blocks without spatial symmetry and the Sz = 0 block alone; each with the automatic
dense floor (small blocks solved densely on the host) and with every block above two states on the
device Krylov lanes (dense_max_dim = 0). Every case prints its name before it touches the device."""

import json
import os
import sys

import numpy as np
import pytest

from conftest import FIXTURES

import qed_spin_amd as qs
from qed_spin_amd import eigs as ev

pytestmark = pytest.mark.gpu

CASES = {
    "none": qs.Symmetry.none(),
    "auto": qs.Symmetry(),
    "spatial=none": qs.Symmetry(translations=False),
    "sz=6": qs.Symmetry(sz=[6], translations=False),
    "auto/sz=6": qs.Symmetry(sz=[6]),
}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _records():
    with open(os.path.join(FIXTURES, "reference_records.json")) as f:
        return json.load(f)["records"]


def _h12():
    n = 12
    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    b.heisenberg([(i, (i + 2) % n) for i in range(n)], 0.5)
    return b.to_operator()


@pytest.mark.parametrize("prune", [True, False])
@pytest.mark.parametrize("dense_max_dim", [None, 0])
@pytest.mark.parametrize("key", list(CASES))
def test_eigs_reproduces_reference_records(key, dense_max_dim, prune):
    say("eigs", key, "dense_max_dim", dense_max_dim, "prune", prune)
    want = _records()[f"j1j2_chain12/api/eigs/{key}/k=4"]["values"]["energies"]
    r = ev.eigs(_h12(), 4, sym=CASES[key], dense_max_dim=dense_max_dim, prune=prune)
    assert r.converged
    assert np.allclose(r.energies, want, rtol=1e-10, atol=1e-10), (r.energies, want)
    lanes = {b["lane"] for b in r.blocks}
    if dense_max_dim == 0:
        assert "device-krylov" in lanes
    if not prune:
        assert r.timings["pruned_blocks"] == 0 and "pruned" not in lanes
    elif key == "spatial=none":
        # the Sz blocks of 220-924 states far from the lowest levels are estimated and skipped
        assert r.timings["pruned_blocks"] > 0


@pytest.mark.parametrize("device_min_dim", [None, 0])
@pytest.mark.parametrize("key", list(CASES))
def test_spectrum_reproduces_reference_records(key, device_min_dim):
    """The whole spectrum over the walk (dense blocks from 1024 states on the device by default,
    every block on the device with device_min_dim = 0) against the reference's recorded spectrum."""
    say("spectrum", key, "device_min_dim", device_min_dim)
    want = np.sort(np.array(_records()[f"j1j2_chain12/api/spectrum/{key}"]["values"]["eigenvalues"]))
    r = ev.spectrum(_h12(), sym=CASES[key], device_min_dim=device_min_dim)
    assert r.eigenvalues.size == want.size
    assert np.allclose(r.eigenvalues, want, rtol=1e-10, atol=1e-10), np.max(np.abs(r.eigenvalues - want))
    if device_min_dim == 0 or key == "none":
        assert any(b["lane"] == "device-dense" for b in r.blocks)


def test_exact_thermal_device_lane_matches_host_lane():
    """Exact thermodynamics of the 4096-state block (device lane: 4096 >= 1024) against the reference's
    recorded exact curves of the whole space."""
    say("exact thermal on the device lane")
    from qed_spin_amd import thermal as th

    with open(os.path.join(FIXTURES, "reference_records.json")) as f:
        rec = json.load(f)
    r = rec["records"]["j1j2_chain12/api/thermal/exact/none"]["values"]
    T = np.array(r["T"])
    res = th.thermal(_h12(), T, sym=qs.Symmetry.none(), dense_max_dim=8192, seed=1)
    assert res.blocks[0]["exact"] and res.blocks[0]["dim"] == 4096
    for key in ("E", "C", "S"):
        assert np.allclose(getattr(res, key), r[key], rtol=1e-10, atol=1e-10), key


def test_expect_reproduces_reference_record():
    """<S_0 . S_1> in the eigenvectors of the 4 lowest levels over every (Sz, k) block against the
    reference's recorded expect of j1j2_chain12: each recorded (energy, bond) pair matched by one
    level (energies within 1e-10; bonds within 1e-8, the accuracy of eigenvectors locked at a 1e-10
    residual), the two ground states (momentum 0 and pi) told apart by their bonds."""
    say("expect S0.S1")
    rec = _records()["j1j2_chain12/api/expect/auto/S0.S1/k=4"]["values"]
    bond = qs.HamiltonianBuilder(12).heisenberg([(0, 1)], 1.0).to_operator()
    r = ev.expect(_h12(), [bond], 4, sym=qs.Symmetry())
    assert np.allclose(r.energies, rec["energy"], rtol=1e-10, atol=1e-10)
    used = set()
    for e, b in zip(rec["energy"], rec["bond"]):
        hit = [i for i in range(4) if i not in used and abs(r.energies[i] - e) < 1e-9 and abs(r.values[i, 0] - b) < 1e-8]
        assert hit, (e, b, r.energies, r.values[:, 0])
        used.add(hit[0])


@pytest.mark.parametrize("dense_max_dim", [None, 0])
@pytest.mark.parametrize("sym", ["spatial=none", "auto"])
def test_total_spin_min_reproduces_the_record(sym, dense_max_dim):
    """eigs(total_spin="min"): the 4 lowest singlets of j1j2_chain12 (the reference records them over
    the Sz = 0 sector without spatial symmetry) within 1e-10 -- in that sector and over its momentum
    blocks, by the dense lanes and by Krylov-Schur on the penalised operator."""
    say("total_spin=min", sym, "dense_max_dim", dense_max_dim)
    want = _records()["j1j2_chain12/api/eigs/total_spin=min/k=4"]["values"]["energies"]
    s = qs.Symmetry(translations=False) if sym == "spatial=none" else qs.Symmetry()
    r = ev.eigs(_h12(), 4, sym=s, total_spin="min", dense_max_dim=dense_max_dim)
    assert r.converged
    assert np.allclose(r.energies, want, rtol=1e-10, atol=1e-10), (r.energies, want)
    assert all(L["n_up"] == 6 and L["total_spin"] == 0.0 for L in r.levels)
    if dense_max_dim == 0:
        assert "device-krylov" in {b["lane"] for b in r.blocks}


def test_vectors_record_count_and_rayleigh():
    """eigs(k = 3, vectors=True).vectors() over every block: the reference records 3 vectors for
    j1j2_chain12; each is a unit eigenvector whose Rayleigh quotient is its level's energy."""
    say("vectors record")
    want = _records()["j1j2_chain12/api/eigs/auto/vectors/k=3"]["values"]["count"]
    H = _h12()
    r = ev.eigs(H, 3, vectors=True)
    assert len(r.vectors()) == want
    vs = [v for n in sorted({L["n_up"] for L in r.levels}) for v in r.vectors("sz", n_up=n)]
    assert len(vs) == want
    for v in vs:
        assert abs(np.linalg.norm(v) - 1.0) < 1e-12
    ex = r.expect([H])
    assert np.allclose(ex[:, 0].real, r.energies, atol=1e-10)


@pytest.mark.parametrize("prune", [True, False])
@pytest.mark.parametrize("key", ["auto", "sz=6", "auto/sz=6"])
def test_streamed_walk_reproduces_reference_records(key, prune):
    """eigs(stream=True) -- one block per build thread on the device at a time, estimates and pruning
    as the resident walk -- gives the recorded levels of j1j2_chain12 (1e-10) and the same pruned
    count as the resident walk."""
    say("eigs streamed", key, "prune", prune)
    want = _records()[f"j1j2_chain12/api/eigs/{key}/k=4"]["values"]["energies"]
    r = ev.eigs(_h12(), 4, sym=CASES[key], prune=prune, stream=True, dense_max_dim=0)
    assert r.converged
    assert np.allclose(r.energies, want, rtol=1e-10, atol=1e-10), (r.energies, want)
    w = ev.eigs(_h12(), 4, sym=CASES[key], prune=prune, dense_max_dim=0)
    assert r.timings["pruned_blocks"] == w.timings["pruned_blocks"]
    assert [b["lane"] for b in r.blocks] == [b["lane"] for b in w.blocks]


def test_spin_flip_blocks_reproduce_the_records():
    """Translations times the spin flip at Sz = 0 (both parities): the recorded Sz = 0 levels of
    j1j2_chain12, resident and streamed, with vectors (unit, Rayleigh quotient = energy)."""
    say("spin flip blocks")
    want = _records()["j1j2_chain12/api/eigs/auto/sz=6/k=4"]["values"]["energies"]
    H = _h12()
    for stream in (False, True):
        r = ev.eigs(H, 4, sym=qs.Symmetry(sz=[6], spin_flip="all"), stream=stream, vectors=True)
        assert np.allclose(r.energies, want, rtol=1e-10, atol=1e-10), (stream, r.energies, want)
        assert len(r.blocks) == 24 and all("Z=" in b["label"] for b in r.blocks)
        ex = r.expect([H])[:, 0].real
        assert np.allclose(ex, r.energies, atol=1e-10)

"""GPU parity on the reference's small audit models (tests/golden/audit_models.py): every recorded
eigs k = 4 and spectrum of dimer, triangle3, chain4, chain5_odd, open_chain10, xy_chain10,
staggered_field10, tfim10 (no U(1)), random_real8, random_complex8 and tri_chiral4x3 (complex
chiral couplings) whose symmetry selection maps onto this backend's -- the whole space, the Sz
sectors with and without momentum blocks, total_spin=min -- within 1e-10 of the reference's recorded
values (tests/python/golden/refs/api-2026-09/cpu.json.gz, copied to
tests/golden/fixtures/audit_records.json). eigs runs with the automatic dense floor and with every
block above two states on the device Krylov lanes."""

import json
import os
import sys

import numpy as np
import pytest

from conftest import FIXTURES
from qed_spin_amd import eigs as ev
from tests.golden import audit_models as am

pytestmark = pytest.mark.gpu


def _records():
    with open(os.path.join(FIXTURES, "audit_records.json")) as f:
        return json.load(f)["records"]


RECS = _records()
CASES = [(m, verb, label, key) for m in am.MODELS for verb, label, key in am.record_keys(m, RECS) if key in RECS]


@pytest.mark.parametrize("name,verb,label,key", CASES, ids=[c[3] for c in CASES])
def test_audit_record(name, verb, label, key):
    print("[case]", key, flush=True)
    sys.stdout.flush()
    H = am.operator(name)
    sym, total_spin = am.symmetry(name, label)
    v = RECS[key]["values"]
    if verb == "spectrum":
        got = ev.spectrum(H, sym=sym).eigenvalues
        want = np.sort(v["eigenvalues"])
        assert got.size == want.size
        assert np.abs(got - want).max() < 1e-10
        return
    want = np.asarray(v["energies"])
    for dmd in (None, 0):
        r = ev.eigs(H, 4, sym=sym, total_spin=total_spin, dense_max_dim=dmd)
        assert r.converged
        assert r.energies.size == want.size, (dmd, r.energies, want)
        assert np.allclose(r.energies, want, rtol=1e-10, atol=1e-10), (dmd, r.energies, want)


THERMAL = [m for m in ("tfim10", "tri_chiral4x3") if f"{m}/api/thermal/exact/none" in RECS]


@pytest.mark.parametrize("name", THERMAL)
def test_audit_thermal_records(name):
    """Exact thermodynamics over the whole space and over this backend's blocks (the reference's
    'auto' records the same curves) within 1e-10 (C 1e-9); FTLM over the whole space with the
    recorded samples 8, krylov 40, seed 11 within 1e-8 -- for tri_chiral4x3 a complex H, whose
    sample vectors are complex."""
    from qed_spin_amd import thermal as th

    print("[case] thermal", name, flush=True)
    H = am.operator(name)
    u1 = am.MODELS[name][2]
    for label, sym in (("none", am.symmetry(name, "none")[0]), ("auto", am.symmetry(name, "auto")[0])):
        r = RECS[f"{name}/api/thermal/exact/{label}"]["values"]
        res = th.thermal(H, np.array(r["T"]), sym=sym, method="exact")
        for key, got, tol in (("E", res.E, 1e-10), ("C", res.C, 1e-9), ("S", res.S, 1e-10)):
            assert np.abs(got - np.array(r[key])).max() < tol, (label, key)
    r = RECS[f"{name}/api/thermal/ftlm/none"]["values"]
    res = th.thermal(H, np.array(r["T"]), sym=am.symmetry(name, "none")[0], samples=8, krylov=40, seed=11)
    assert len(res.blocks) == 1 and not res.blocks[0]["exact"]
    for key, got in (("E", res.E), ("C", res.C), ("S", res.S)):
        err = np.abs(got - np.array(r[key])).max()
        print(f"  {key}: max |diff| {err:.3e}", flush=True)
        assert np.allclose(got, r[key], rtol=1e-8, atol=1e-8), key
    assert u1 or name == "tfim10"


DYN = [k for k in RECS if "/dynamics/" in k]


@pytest.mark.parametrize("key", DYN)
def test_audit_dynamics_records(key):
    """S^z(Q) = N^-1/2 sum_i e^{i pi i} S^z_i (site i = x + Lx y) of square4x3 over the whole space:
    T = 0 (krylov 80; the reference's 'auto' walk records the same curve) and T = 1 (FTLM, samples
    6, krylov 40, seed 11) within 1e-8 of the records, omega = linspace(-1, 4, 41), eta = 0.15."""
    from qed_spin_amd import dynamics as dy
    from qed_spin_amd.operator import Operator

    print("[case]", key, flush=True)
    name = key.split("/api/")[0]
    H = am.operator(name)
    N = H.num_sites
    O = Operator(N)
    for i in range(N):
        O.add_product("z", (i,), np.exp(1j * np.pi * i) / np.sqrt(N))
    omega = np.linspace(-1.0, 4.0, 41)
    want = np.array(RECS[key]["values"]["S"])
    none = am.symmetry(name, "none")[0]
    if "/T=0/" in key:
        r = dy.dynamics(H, O, omega, eta=0.15, sym=none, krylov=80)
    else:
        r = dy.dynamics(H, O, omega, eta=0.15, T=[1.0], sym=none, krylov=40, samples=6, seed=11)
    got = r.S[0][0] if "/T=1/" in key else r.S[0]
    got = np.asarray(got).real.ravel()
    err = np.abs(got - want).max()
    print(f"  max |diff| {err:.3e}", flush=True)
    assert np.allclose(got, want, rtol=1e-8, atol=1e-8)

"""The reference's small audit models restated (tests/golden/audit_models.py) and pinned on the CPU:
the dense spectrum of each model's whole space (the oracle's dense matrix, numpy eigvalsh) equals the
reference's recorded spectrum/none within 1e-10, and every recorded eigs k = 4 and Sz-sector spectrum
is consistent with it (the k lowest, multiplicities repeated; the Sz sector's eigenvalues a
sub-multiset)."""

import json
import os

import numpy as np
import pytest

from conftest import FIXTURES
from oracle import pyoracle as po
from qed_spin_amd import eigs as ev
from tests.golden import audit_models as am


def _records():
    with open(os.path.join(FIXTURES, "audit_records.json")) as f:
        return json.load(f)["records"]


@pytest.mark.parametrize("name", list(am.MODELS))
def test_restated_model_has_the_recorded_spectrum(name):
    recs = _records()
    H = am.operator(name)
    N = H.num_sites
    if N > 12:
        pytest.skip("2^N too large for the dense oracle here: the model is checked by its eigs records on the device")
    M = po.dense_matrix(H.products, N)
    assert np.abs(M - M.conj().T).max() < 1e-14
    w = np.linalg.eigvalsh(M)
    want = np.asarray(recs[f"{name}/api/spectrum/none"]["values"]["eigenvalues"])
    assert want.size == 1 << N
    assert np.abs(w - np.sort(want)).max() < 1e-10
    def sector(n_up):
        states = [x for x in range(1 << N) if bin(x).count("1") == n_up]
        return states, M[np.ix_(states, states)]

    for verb, label, key in am.record_keys(name, recs):
        if key not in recs:
            continue
        v = recs[key]["values"]
        if label.startswith("sz=") or label.startswith("auto/sz="):
            ws = np.linalg.eigvalsh(sector(int(label.split("=")[-1]))[1])
        elif label == "total_spin=min":
            two_S = N % 2
            states, Ms = sector((N + two_S) // 2)
            S2 = po.dense_matrix(ev.spin_squared(N).products, N, states)
            lam, Q = np.linalg.eigh(S2)
            Q = Q[:, np.abs(lam - 0.25 * two_S * (two_S + 2)) < 1e-8]
            ws = np.repeat(np.linalg.eigvalsh(Q.conj().T @ Ms @ Q), two_S + 1)   # a multiplet counts 2S + 1
        else:
            ws = w
        if verb == "eigs":
            got = np.asarray(v["energies"])
            assert got.size == min(4, ws.size), key
            assert np.allclose(got, ws[: got.size], atol=1e-10), key
        else:
            assert np.abs(ws - np.sort(v["eigenvalues"])).max() < 1e-10, key


THERMAL = ["tfim10", "tri_chiral4x3"]


@pytest.mark.parametrize("name", THERMAL)
def test_thermal_records_are_the_dense_thermodynamics(name):
    """The recorded exact curves (whole system, not per site) are E = <E>, C = (<E^2> - <E>^2) / T^2,
    S = ln Z + <E> / T of the dense spectrum (1e-10; C 1e-9, a difference of large moments)."""
    recs = _records()
    H = am.operator(name)
    w = np.linalg.eigvalsh(po.dense_matrix(H.products, H.num_sites))
    for label in ("none", "auto"):
        r = recs[f"{name}/api/thermal/exact/{label}"]["values"]
        b = 1.0 / np.asarray(r["T"])
        x = np.exp(-np.outer(b, w - w[0]))
        Z = x.sum(1)
        E = (x * w).sum(1) / Z
        C = ((x * w * w).sum(1) / Z - E * E) * b * b
        S = np.log(Z) + b * (E - w[0])
        assert np.abs(E - r["E"]).max() < 1e-10 and np.abs(S - r["S"]).max() < 1e-10
        assert np.abs(C - r["C"]).max() < 1e-9


def test_complex_ftlm_record_pins_the_draw_order():
    """tri_chiral4x3's recorded FTLM (a complex H over the whole 4096-state space, samples 8,
    krylov 40, seed 11) through the oracle's FTLM within 1e-10. Its sample vectors take each entry's
    first normal as the imaginary part (the reference's Complex(nd(gen), nd(gen)) under GCC's
    right-to-left argument evaluation); drawn the other way round the curves miss by ~5e-2. A real
    H cannot tell the two orders apart."""
    import qed_spin_amd as qs

    r = _records()["tri_chiral4x3/api/thermal/ftlm/none"]["values"]
    H = am.operator("tri_chiral4x3")
    s = po.OracleSector(12, -1, np.arange(12, dtype=np.int32)[None, :], [1.0])
    row_ptr, col, val = po.sector_csr(qs.Program(H, adjoint=True).export(), s)
    betas = 1.0 / np.array(r["T"])
    seed = (11 * 6364136223846793005 + 1442695040888963407) % (1 << 64)
    E = {}
    for order in ("im first", "re first"):
        moments = []
        for k in range(8):
            v0 = po.gaussian_vector(seed, k, s.dim)
            if order == "re first":
                v0 = v0.imag + 1j * v0.real
            alpha, beta = po.lanczos(row_ptr, col, val, v0, 40, 2, 64 * np.finfo(float).eps * H.norm_bound())
            ritz, V = po.tridiag_eig(alpha, beta)
            moments.append(po.ftlm_moments(ritz, np.abs(V[0]) ** 2, betas))
        E[order] = po.ftlm_combine(np.array(moments), s.dim, betas)[1]
    assert np.abs(E["im first"] - r["E"]).max() < 1e-10
    assert np.abs(E["re first"] - r["E"]).max() > 1e-3

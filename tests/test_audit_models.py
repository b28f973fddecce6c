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

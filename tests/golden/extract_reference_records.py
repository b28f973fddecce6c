"""Copy the reference's own recorded values for the thermal and dynamics verbs into a fixture.

The reference keeps golden-master records of every public verb on small models
(tests/python/golden/refs/api-2026-09/cpu.json.gz, written by tests/python/golden/golden.py from
the cases of tests/python/golden/cases.py). This script reads that data file (JSON, no code) and
writes the records this backend is checked against, with the inputs that produced them, to
tests/golden/fixtures/reference_records.json:

    python tests/golden/extract_reference_records.py [/root/reference]

Model j1j2_chain12 (tests/python/support/models.py:527-535): the J1-J2 Heisenberg ring, N = 12,
J1 = 1 on (i, i+1), J2 = 0.5 on (i, i+2). Cases (cases.py:138-206): thermal at
T = linspace(0.25, 3, 8), method exact / ftlm (samples 8, krylov 40, seed 11), symmetry auto / none;
dynamics of S^z(pi) = N^-1/2 sum_i e^{i pi i} S^z_i at omega = linspace(-1, 4, 41), eta = 0.15,
T = 0 with krylov 80, T = 1 with krylov 40, samples 6, seed 11; eigs k = 4 (cases.py:95-110) over
the whole space, every symmetry block, the Sz blocks without spatial symmetry, and Sz = 0 alone;
the whole spectrum over the same selections; expect of S_0 . S_1 in the 4 lowest levels.
"""

from __future__ import annotations

import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = [
    "j1j2_chain12/api/thermal/exact/auto",
    "j1j2_chain12/api/thermal/exact/none",
    "j1j2_chain12/api/thermal/ftlm/none",
    "j1j2_chain12/api/thermal/ftlm/auto",
    "j1j2_chain12/api/dynamics/T=0/none/SzQ",
    "j1j2_chain12/api/dynamics/T=0/auto/SzQ",
    "j1j2_chain12/api/dynamics/T=1/none/SzQ",
    "j1j2_chain12/api/eigs/none/k=4",
    "j1j2_chain12/api/eigs/auto/k=4",
    "j1j2_chain12/api/eigs/spatial=none/k=4",
    "j1j2_chain12/api/eigs/sz=6/k=4",
    "j1j2_chain12/api/eigs/auto/sz=6/k=4",
    "j1j2_chain12/api/spectrum/none",
    "j1j2_chain12/api/spectrum/auto",
    "j1j2_chain12/api/spectrum/spatial=none",
    "j1j2_chain12/api/spectrum/sz=6",
    "j1j2_chain12/api/spectrum/auto/sz=6",
    "j1j2_chain12/api/expect/auto/S0.S1/k=4",
    "j1j2_chain12/api/eigs/total_spin=min/k=4",
    "j1j2_chain12/api/eigs/auto/vectors/k=3",
]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "tests/python/golden/refs/api-2026-09/cpu.json.gz")
    with gzip.open(src, "rt") as f:
        d = json.load(f)
    recs = d["records"]
    out = {
        "source": "tests/python/golden/refs/api-2026-09/cpu.json.gz",
        "reference_sha": d["meta"].get("sha"),
        "model": {"name": "j1j2_chain12", "N": 12, "J1": 1.0, "J2": 0.5},
        "thermal": {"T": [0.25 + i * (2.75 / 7) for i in range(8)], "samples": 8, "krylov": 40, "seed": 11,
                    "dense_max_dim": 512},
        "dynamics": {"omega": [-1.0 + i * 0.125 for i in range(41)], "eta": 0.15, "krylov_T0": 80,
                     "krylov_T": 40, "samples": 6, "seed": 11, "probe": "Sz(pi)"},
        "records": {k: recs[k] for k in KEYS},
    }
    dst = os.path.join(HERE, "fixtures", "reference_records.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)
    # the small audit models: every eigs / spectrum record whose symmetry selection maps onto this
    # backend's (audit_models.record_keys)
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from tests.golden import audit_models as am

    keys = sorted(k for m in am.MODELS for verb, label, k in am.record_keys(m, recs)
                  if verb == "eigs" or (am.MODELS[m][0] <= 12 and (label == "none" or label.startswith("sz="))))
    keys += sorted(k for m in am.MODELS for k in (f"{m}/api/thermal/exact/none", f"{m}/api/thermal/exact/auto",
                                                 f"{m}/api/thermal/ftlm/none", f"{m}/api/dynamics/T=0/none/SzQ",
                                                 f"{m}/api/dynamics/T=0/auto/SzQ", f"{m}/api/dynamics/T=1/none/SzQ")
                   if k in recs)
    audit = {"source": out["source"], "reference_sha": out["reference_sha"], "records": {k: recs[k] for k in keys}}
    dst = os.path.join(HERE, "fixtures", "audit_records.json")
    with open(dst, "w") as f:
        json.dump(audit, f, separators=(",", ":"))
    print("wrote", dst, len(keys), "records")


if __name__ == "__main__":
    main()

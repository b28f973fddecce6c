"""Copies the deterministic full-size results of the reference's recorded bench runs
(bench/baseline/v1_cd08f64.jsonl: chain32_abelian_eigs E0, chain30_dyn0 weight) into
fixtures/reference_bench_v1.json, with the case definitions they belong to (bench/cases.py:62-65,
105-116). Run here, where the reference checkout exists; the GPU tests read only the fixture."""

import json
import os

REF = "/root/reference/bench/baseline/v1_cd08f64.jsonl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "reference_bench_v1.json")

CASES = {
    "chain32_abelian_eigs_gpu": {
        "source": "bench/baseline/v1_cd08f64.jsonl:3; case bench/cases.py:62-65",
        "model": "chain(32): J1 = 1 on (i, i+1), J2 = 0.35 on (i, i+2), periodic",
        "spec": "eigs k = 1, Sz = 0 (n_up = 16), translations and the spin flip, every block",
    },
    "chain30_dyn0_cpu": {
        "source": "bench/baseline/v1_cd08f64.jsonl:7; case bench/cases.py:105-116",
        "model": "chain(30): J1 = 1, J2 = 0.35, periodic",
        "spec": "T = 0 S^z(q = 2 pi 10 / 30), n_up = 15, every momentum block, krylov 200, "
                "omega = linspace(-1, 7, 401), eta 0.05; weight = trapezoid(S, omega)",
    },
}


def main():
    out = {}
    with open(REF) as f:
        for line in f:
            rec = json.loads(line)
            if rec.get("case") in CASES and rec.get("result"):
                out[rec["case"]] = dict(CASES[rec["case"]], result=rec["result"], commit=rec["commit"])
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {OUT}: {sorted(out)}")


if __name__ == "__main__":
    main()

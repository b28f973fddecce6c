#!/usr/bin/env python3
"""Headline benchmark: the FP64 symmetry-sector matvec of the N = 32 J1-J2 ring (J2 = 1/2, periodic,
Sz = 0, k = 0, reflection-even; dim 9,398,520, real) on one MI355X, plus the time to E0 of its
300-step fully reorthogonalised Lanczos.

A step is one apply of the sector operator (the merged reduced CSR built on the device) to one
vector already resident in HBM. value = algorithmic bytes per apply / apply time, with
    B(1) = D*8 + nnz*(4 + w_id) + (nnz*8 + D*8)
(row descriptor, column + value id, gathered x and written y; w_id the reference's id width).
The apply time is measured with HIP events on the stream the SpMV runs on, over exactly --steps
applies after --warmup untimed ones.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--no-lanczos] [--no-cpu]
                    [--workload matvec,serve,ftlm,dynamics] [--serve-requests 2000 --serve-plain-requests 200]

By default the line also carries a sample of the served stream ("serve": the first 200 requests of
the mixed stream served at saturation, and the plain one-request-at-a-time path against the served
path on its first 40).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 32
J2 = 0.5


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(m, D, w_id, target_s=10.0):
    """The oracle's real-block SpMV (the reference's RealCsrView::spmv restated) on the host cores:
    over the whole of this sector's CSR (downloaded) with one OpenMP thread per host core this
    process may use (at most 16, a static row schedule as the reference's host lane), repeated for
    about target_s seconds; and on one thread over the first 10^6 rows, for the record. Reported in
    the metric's unit over the algorithmic bytes of what was applied."""
    import numpy as np

    from oracle import pyoracle as po
    from qed_spin_amd import _lib

    L = po.lib()
    P64, P32, PD = C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    L.orc_csr_matvec_real.restype = None
    L.orc_csr_matvec_real.argtypes = [C.c_int64, P64, P32, PD, PD, PD]
    L.orc_csr_matvec_real_mt.restype = None
    L.orc_csr_matvec_real_mt.argtypes = [C.c_int64, P64, P32, PD, PD, PD, C.c_int]

    def download(rows):
        rp = np.zeros(rows + 1, np.uint64)
        _lib.call("qedh_matvec_export_csr_rows", m.h, 0, rows, _lib.ptr(rp, C.c_uint64), None, None)
        nnz = int(rp[-1] - rp[0])
        col = np.zeros(nnz, np.uint32)
        val = np.zeros(2 * nnz, np.float64)
        _lib.call("qedh_matvec_export_csr_rows", m.h, 0, rows, _lib.ptr(rp, C.c_uint64), _lib.ptr(col, C.c_uint32),
                  _lib.ptr(val, C.c_double))
        re = np.ascontiguousarray(val[0::2])
        del val
        return np.ascontiguousarray(rp - rp[0]), col, re, nnz

    def timed(fn, budget):
        reps, t0 = 0, time.perf_counter()
        while True:
            fn()
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return el / reps, reps

    x = np.random.default_rng(3).normal(size=D)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rp, col, re, nnz = download(D)
    y = np.zeros(D)
    args = (D, rp.ctypes.data_as(P64), col.ctypes.data_as(P32), re.ctypes.data_as(PD), x.ctypes.data_as(PD),
            y.ctypes.data_as(PD))
    per_mt, reps_mt = timed(lambda: L.orc_csr_matvec_real_mt(*args, threads), target_s)
    B = D * 8 + nnz * (4 + w_id) + (nnz * 8 + D * 8)
    del rp, col, re, y
    rows = min(D, 1_000_000)
    rp1, col1, re1, nnz1 = download(rows)
    y1 = np.zeros(rows)
    args1 = (rows, rp1.ctypes.data_as(P64), col1.ctypes.data_as(P32), re1.ctypes.data_as(PD), x.ctypes.data_as(PD),
             y1.ctypes.data_as(PD))
    per_1, reps_1 = timed(lambda: L.orc_csr_matvec_real(*args1), target_s / 2)
    B1 = rows * 8 + nnz1 * (4 + w_id) + (nnz1 * 8 + rows * 8)
    return {
        "value": round(B / per_mt / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole sector CSR ({D} rows, {nnz} entries) on {threads} OpenMP threads, {reps_mt} applies, "
                  f"{per_mt * 1e3:.1f} ms each",
        "single_core": {"value": round(B1 / per_1 / 1e9, 3), "cores": 1,
                        "sample": f"rows 0..{rows} ({nnz1} entries), {reps_1} applies, {per_1 * 1e3:.1f} ms each"},
        "cpu": _cpu_model(),
    }


def _oracle_lanczos_rate(m, reorth, steps, target_s, max_rows=None):
    """The oracle's Lanczos (the reference's kernel restated, one thread, complex CSR) on the CSR of
    matvec m (its first max_rows rows when the whole block would take too long): rows x steps per
    second over repeated runs for about target_s seconds."""
    import numpy as np

    from oracle import pyoracle as po

    row_ptr, col, val = m.export_csr()
    rp = row_ptr.astype(np.int64)
    n = m.dim
    v0 = np.random.default_rng(7).normal(size=n) + 1j * np.random.default_rng(8).normal(size=n)
    reps, t0 = 0, time.perf_counter()
    done = 0
    while True:
        a, _ = po.lanczos(rp, col.astype(np.int64), val, v0, steps, reorth=reorth)
        done += len(a)
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_s:
            break
    return n * done / el, {"rows": n, "nnz": int(rp[-1]), "runs": reps, "steps_per_run": steps, "seconds": el}


def run_ftlm(args):
    """FTLM thermodynamics of the XXZ ring N=24 (Jxy = 1, Delta = 0.5) over every (Sz, k) block:
    40 samples per block (seed 7), Lanczos depth 100, 40 temperatures linspace(0.1, 4); blocks of at
    most 512 states exactly, all others through the batched engine in one packed multi-vector matvec
    per step. roofline: the packed step's algorithmic bytes (rows*8 + nnz*(4 + w) + R*(nnz*16 +
    rows*16), w the stored value bytes) over the measured step time; cpu_baseline: the oracle's
    LocalDGKS3 Lanczos (the reference's FTLM sample kernel restated) on one thread over the block
    with the most rows, in the same unit (row-vector Lanczos steps per second)."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import thermal as th

    n = args.ftlm_sites
    H = qs.HamiltonianBuilder(n).xxz([(i, (i + 1) % n) for i in range(n)], 1.0, 0.5).to_operator()
    T = np.linspace(0.1, 4.0, 40)
    t0 = time.perf_counter()
    r = th.thermal(H, T, sym=qs.Symmetry(), samples=40, krylov=100, seed=7)
    wall = time.perf_counter() - t0
    e = r.engine
    R = 40
    rows, nnz = e.get("rows", 0.0), e.get("nnz", 0.0)
    wval = e.get("value_bytes", 16.0) or 16.0
    step_bytes = rows * 8 + nnz * (4 + wval) + R * (nnz * 16 + rows * 16)
    steps = e.get("steps", 100.0) or 100.0
    gbs = step_bytes * steps / e["steps_s"] / 1e9 if e.get("steps_s") else None
    row_steps = rows * R * steps / e["steps_s"] if e.get("steps_s") else None
    cpu = None
    if not args.no_cpu:
        blocks = th.walk_blocks(H, qs.Symmetry(sz=[n // 2], momenta=[0]), tune=False)
        rate, sample = _oracle_lanczos_rate(blocks[0][2], 2, 100, 10.0)
        th.release_later(blocks)
        cpu = {"value": round(rate, 1), "unit": "row-vector Lanczos steps/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's LocalDGKS3 Lanczos, depth 100, on the (Sz=0, k=0) block ({sample['rows']} "
                         f"rows, {sample['nnz']} entries), {sample['runs']} runs in {sample['seconds']:.1f} s",
               "cpu": _cpu_model()}
    out = {
        "throughput": {"value": round(row_steps, 1) if row_steps else None, "unit": "row-vector Lanczos steps/s"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1) if gbs else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "traffic": None,
                     "kernel": "packed SpMV + recurrence per step (batched engine)"},
        "cpu_baseline": cpu,
        "workload": f"FTLM XXZ ring N={n}, Delta=0.5, every (Sz, k) block, 40 samples, depth 100, 40 T",
        "blocks": len(r.blocks),
        "sampled_blocks": sum(1 for b in r.blocks if not b["exact"]),
        "exact_blocks": sum(1 for b in r.blocks if b["exact"]),
        "total_dim": r.total_dim,
        "wall_s": round(wall, 3),
        "timings": {k: round(v, 3) for k, v in r.timings.items()},
        "engine": {k: round(v, 4) for k, v in e.items()},
        "batched_step_GBps_algorithmic": round(gbs, 1) if gbs else None,
        "E_T0.1": float(r.E[0]),
        "C_max": float(r.C.max()),
    }
    log(f"ftlm: {out}")
    return out


def run_config1(args):
    """The Heisenberg ring N=20 (J = 1), Sz = 0, k = 0 (dim 9252, real): its 4 lowest levels through
    the device Krylov-Schur lane, and a 300-step FullCGS2 Lanczos from v0 = N(0,1)^D (seed 1)
    normalised, each timed from the sector build on. cpu_baseline: the oracle's whole path on one
    thread (orbit scan, merged CSR, the same 300-step FullCGS2 Lanczos from the same v0), whose
    lowest Ritz values must agree with the device's."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import _lib
    from qed_spin_amd import symmetry as sym
    from qed_spin_amd.krylov import tridiag_lowest

    n = 20
    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    op = b.to_operator()
    g = sym.chain_group(n, 0)
    _lib.call("qedh_synchronize")
    t0 = time.perf_counter()
    s = qs.Sector(g, n // 2)
    m = qs.SectorMatvec(op, s, lane=qs.LANE_CSR)
    t_build = time.perf_counter() - t0
    t1 = time.perf_counter()
    vals, conv, applies = m.lowest(4)
    t_levels = time.perf_counter() - t1
    D = s.dim
    v0 = np.random.default_rng(1).normal(size=D)
    v0 /= np.linalg.norm(v0)
    t2 = time.perf_counter()
    a, bb, tm = m.lanczos(v0, 300, reorth=1)
    t_lz = time.perf_counter() - t2
    T = np.diag(a) + np.diag(bb[1:len(a)], 1) + np.diag(bb[1:len(a)], -1)
    ritz = np.linalg.eigvalsh(T)[:4]
    info = m.info
    nnz, w_id = int(info["nnz"]), int(info["id_width_ref"])
    B1 = D * 8 + nnz * (4 + w_id) + (nnz * 8 + D * 8)
    avg = C.c_double()
    _lib.call("qedh_matvec_benchmark", m.h, 1, 0, 0, 20, 200, C.byref(avg))
    gbs = B1 / avg.value / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po

        c0 = time.perf_counter()
        o = po.OracleSector(n, n // 2, g.perms, g.chars)
        rp, col, val = po.sector_csr(qs.Program(op).export(), o)
        c1 = time.perf_counter()
        ca, cb = po.lanczos(rp, col, val, v0.astype(np.complex128), 300, reorth=1, breakdown_tol=0.0)
        c2 = time.perf_counter()
        Tc = np.diag(ca) + np.diag(cb[1:len(ca)], 1) + np.diag(cb[1:len(ca)], -1)
        cpu_ritz = np.linalg.eigvalsh(Tc)[:4]
        cpu = {"value": round(c2 - c0, 3), "unit": "s (sector, CSR and 300 FullCGS2 Lanczos steps)", "cores": 1,
               "kind": "port",
               "sample": f"the whole workload on one thread: orbit scan and merged CSR {c1 - c0:.2f} s, Lanczos "
                         f"{c2 - c1:.2f} s",
               "lowest_ritz": [float(x) for x in cpu_ritz],
               "max_abs_difference_to_device": float(np.max(np.abs(cpu_ritz - ritz))),
               "cpu": _cpu_model()}
    out = {
        "workload": "Heisenberg ring N=20, J=1, Sz=0, k=0 (translations only), real",
        "dim": D, "nnz": nnz,
        "build_s": round(t_build, 4),
        "lowest4": [float(x) for x in vals], "lowest4_converged": bool(conv), "lowest4_applies": applies,
        "time_to_lowest4_s": round(t_build + t_levels, 4),
        "lanczos300": {"seconds": round(t_lz, 4), "apply_s": round(tm["apply_s"], 4),
                       "reorth_s": round(tm["reorth_s"], 4), "lowest_ritz": [float(x) for x in ritz],
                       "E0": tridiag_lowest(a, bb)},
        "time_to_lanczos300_s": round(t_build + t_lz, 4),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"the sector SpMV, {B1 / 1e6:.2f} MB algorithmic per apply (cache-resident: "
                               f"this size is launch-bound, not bandwidth-bound)",
                     "apply_us": round(avg.value * 1e6, 2)},
        "cpu_baseline": cpu,
    }
    log(f"config1: {out}")
    return out


def run_dynamics(args):
    """T = 0 S^z(q) of the Heisenberg ring N=28 for all 28 q: the ground state of (Sz=0, k=0), the
    probe images in the 28 momentum sectors, 28 continued fractions of depth 200 through the batched
    engine (omega = linspace(0, 4, 400), eta = 0.05)."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import dynamics as dy
    from qed_spin_amd import thermal as th

    n = args.dyn_sites
    H = qs.HamiltonianBuilder(n).heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0).to_operator()
    omega = np.linspace(0.0, 4.0, 400)
    t0 = time.perf_counter()
    blocks = th.walk_blocks(H, qs.Symmetry(sz=[n // 2]))
    t_build = time.perf_counter() - t0
    k0 = [i for i, (b, s, m) in enumerate(blocks) if b.momentum == 0]
    probes = [dy.sz_q(n, 2 * np.pi * m / n) for m in range(n)]
    r = dy.dynamics(H, probes, omega, eta=0.05, krylov=200, blocks=blocks, ground_blocks=k0)
    wall = time.perf_counter() - t0
    e = r.engine
    rows, nnz = e.get("rows", 0.0), e.get("nnz", 0.0)
    wval = e.get("value_bytes", 2.0) or 2.0
    step_bytes = rows * 8 + nnz * (4 + wval) + (nnz * 16 + rows * 16)
    steps = e.get("steps", 200.0) or 200.0
    gbs = step_bytes * steps / e["steps_s"] / 1e9 if e.get("steps_s") else None
    row_steps = rows * steps / e["steps_s"] if e.get("steps_s") else None
    cpu = None
    if not args.no_cpu:
        tgt = [b for b in blocks if b[0].momentum == n // 2][0]
        rate, sample = _oracle_lanczos_rate(tgt[2], 0, 20, 10.0)
        cpu = {"value": round(rate, 1), "unit": "row-vector Lanczos steps/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's Lanczos without reorthogonalisation (the continued fraction's), 20 steps, on "
                         f"the (Sz=0, k=pi) block ({sample['rows']} rows, {sample['nnz']} entries), "
                         f"{sample['runs']} runs in {sample['seconds']:.1f} s",
               "cpu": _cpu_model()}
    th.release_later(blocks)
    out = {
        "throughput": {"value": round(row_steps, 1) if row_steps else None, "unit": "row-vector Lanczos steps/s"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1) if gbs else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "traffic": None,
                     "kernel": "packed SpMV + recurrence per step (batched engine, 28 right-hand sides)"},
        "cpu_baseline": cpu,
        "norm_sum": float(np.sum(r.norms)),
        "workload": f"T=0 S^z(q) of the Heisenberg ring N={n}, all {n} q, krylov 200, 400 omega, eta 0.05",
        "e0": r.e0,
        "ground_manifold": r.ground_manifold,
        "target_sectors": r.target_sectors,
        "wall_s": round(wall, 3),
        "build_s": round(t_build, 3),
        "timings": {k: round(v, 3) for k, v in r.timings.items()},
        "engine": {k: round(v, 4) for k, v in r.engine.items()},
        "S_pi_max": float(max(r.S[n // 2])),
    }
    log(f"dynamics: {out}")
    return out


def run_dynamics_finite_T(args):
    """A bounded sample of the T > 0 half of the dynamics configuration (the whole half -- every
    (Sz, k) source of the N = 28 ring, 28 probes each, 40 samples, krylov 200 -- keeps ~2e17
    row-steps of work): the Heisenberg ring N = 28, source (Sz = 0, k = 0), probe S^z(pi), 40
    samples (seed 11), krylov 200, T = 0.5 and 1, omega = linspace(0, 4, 400), eta 0.05, through the
    batched kept-basis engine (source and target Lanczos of all samples in lockstep, overlaps as one
    GEMM per sample)."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import dynamics as dy
    from qed_spin_amd import thermal as th

    n = args.dyn_sites
    H = qs.HamiltonianBuilder(n).heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0).to_operator()
    omega = np.linspace(0.0, 4.0, 400)
    blocks = th.walk_blocks(H, qs.Symmetry(sz=[n // 2]), tune=False)
    k0 = [i for i, (b, s, m) in enumerate(blocks) if b.momentum == 0]
    t0 = time.perf_counter()
    r = dy.dynamics(H, [dy.sz_q(n, np.pi)], omega, eta=0.05, T=[0.5, 1.0], samples=40, seed=11, krylov=200,
                    blocks=blocks, sources=k0)
    wall = time.perf_counter() - t0
    dim = blocks[k0[0]][1].dim
    th.release_later(blocks)
    e = r.engine
    out = {"workload": f"FTLM S^z(pi) of the Heisenberg ring N={n}, source (Sz=0, k=0) ({dim} states), 40 samples, "
                       f"krylov 200, T = 0.5, 1",
           "wall_s": round(wall, 3), "engine": {k: round(v, 3) for k, v in e.items()},
           "row_steps_per_s": round(2 * 40 * 200 * dim / e["lanczos_s"], 1) if e.get("lanczos_s") else None,
           "weight_T1": float(np.trapezoid(r.S[0][1].real, omega))}
    log(f"dynamics T>0 sample: {out}")
    return out


def _ref_chain(n, J2=0.35):
    """The reference's bench model chain(N) (tests/python/support/models.py:92-110, imported by
    bench/cases.py through grid.models): the J1-J2 ring, J1 = 1, J2 = 0.35."""
    import qed_spin_amd as qs

    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    b.heisenberg([(i, (i + 2) % n) for i in range(n)], J2)
    return b.to_operator()


def run_observe(args):
    """The fused observable sweep on the Heisenberg ring N=28 (sector Sz=0, k=0: 1,432,860 states) for
    a random unit vector: S_0.S_j (j = 1..14), S^z_j (28) and S^z_i S^z_j (i < j, 378): 420 observables
    in one walk (qedh_matrix_elements), against one apply and one dot per observable on the device."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import _lib
    from qed_spin_amd import symmetry as sym
    from qed_spin_amd.sector import LANE_GATHER, projected_program

    N = 28
    g = sym.chain_group(N, 0)
    sec = qs.Sector(g, N // 2)
    reps, inv = sec.export()
    twin = qs.Sector(g, N // 2, reps=reps, inv_norms=inv)
    ops = []
    for j in range(1, N // 2 + 1):
        O = qs.Operator(N)
        O.add_product("zz", (0, j), 1.0)
        O.add_product("+-", (0, j), 0.5)
        O.add_product("-+", (0, j), 0.5)
        ops.append(O)
    ops += [qs.Operator(N).add_product("z", (j,), 1.0) for j in range(N)]
    ops += [qs.Operator(N).add_product("zz", (i, j), 1.0) for i in range(N) for j in range(i + 1, N)]
    rng = np.random.default_rng(2)
    v = rng.normal(size=sec.dim) + 1j * rng.normal(size=sec.dim)
    v /= np.linalg.norm(v)
    dv = qs.DeviceArray.from_host(v)
    t0 = time.perf_counter()
    progs = [projected_program(O, twin, sec) for O in ops]
    t_compile = time.perf_counter() - t0
    qs.matrix_elements(ops, twin, sec, [dv], [dv], programs=progs)   # warm
    _lib.call("qedh_synchronize")
    t1 = time.perf_counter()
    got = qs.matrix_elements(ops, twin, sec, [dv], [dv], programs=progs)[0]
    t_sweep = time.perf_counter() - t1
    y = qs.DeviceArray(sec.dim)
    mvs = [qs.SectorMatvec(O, twin, sec, lane=LANE_GATHER, program=p) for O, p in zip(ops, progs)]
    _lib.call("qedh_synchronize")
    t2 = time.perf_counter()
    ref = np.zeros(len(ops), np.complex128)
    out = np.zeros(2)
    for o, m in enumerate(mvs):
        m.apply(dv, y, complex_vectors=True)
        _lib.call("qedh_vdot", dv.ptr, y.ptr, C.c_int64(sec.dim), 1, _lib.ptr(out, C.c_double))
        ref[o] = complex(out[0], out[1])
    t_loop = time.perf_counter() - t2
    diff = float(np.max(np.abs(got - ref)) / max(1e-300, np.max(np.abs(ref))))
    return {"workload": f"Heisenberg ring N={N}, Sz=0, k=0 ({sec.dim} states), {len(ops)} observables, one vector",
            "sweep_s": round(t_sweep, 4), "per_observable_s": round(t_loop, 4), "speedup": round(t_loop / t_sweep, 1),
            "compile_s": round(t_compile, 3), "max_rel_difference": diff}


def run_reference_dynamics(args):
    """The reference's two published dynamics cases, same inputs (its bench/cases.py:105-132, model
    chain(N): the J1-J2 ring with J2 = 0.35): chain30_dyn0 -- T = 0 S^z(q = 2 pi 10 / 30), N = 30,
    Sz = 0, every momentum block, krylov 200, omega = linspace(-1, 7, 401), eta 0.05 (published:
    57.6 s on a 32-core node); chain24_dynT -- T = 1 FTLM S^z(q = 2 pi 8 / 24), N = 24, every
    (Sz, k) source block, 5 samples, krylov 80, omega = linspace(-3, 7, 301), eta 0.1 (published:
    ~85 s on a 32-core node). The published numbers are context: another machine."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import dynamics as dy

    out = {}
    n = 30
    H = _ref_chain(n)
    w = np.linspace(-1, 7, 401)
    t0 = time.perf_counter()
    r = dy.dynamics(H, [dy.sz_q(n, 2 * np.pi * 10 / n)], w, eta=0.05, krylov=200, sym=qs.Symmetry(sz=[15]))
    out["chain30_dyn0"] = {"wall_s": round(time.perf_counter() - t0, 3), "weight": float(np.trapezoid(r.S[0], w)),
                           "e0": r.e0, "timings": {k: round(v, 3) for k, v in r.timings.items()},
                           "published_cpu_s": 57.6}
    log(f"chain30_dyn0: {out['chain30_dyn0']}")
    n = 24
    H = _ref_chain(n)
    w = np.linspace(-3, 7, 301)
    t0 = time.perf_counter()
    r = dy.dynamics(H, [dy.sz_q(n, 2 * np.pi * 8 / n)], w, eta=0.1, krylov=80, T=[1.0], samples=5, seed=11,
                    sym=qs.Symmetry())
    out["chain24_dynT"] = {"wall_s": round(time.perf_counter() - t0, 3),
                           "weight": float(np.trapezoid(r.S[0][0].real, w)),
                           "timings": {k: round(v, 3) for k, v in r.timings.items()},
                           "engine": {k: round(float(v), 3) for k, v in r.engine.items()}, "published_cpu_s": 85.0}
    log(f"chain24_dynT: {out['chain24_dynT']}")
    return out


def run_reference_ftlm(args):
    """The reference's published GPU thermal case chain28_ftlm_gpu, same inputs (bench/cases.py:68-88):
    FTLM of chain(28) (J1-J2 ring, J2 = 0.35) over every (Sz, k) block, 4 samples, krylov 60, seed
    7, T = linspace(0.1, 4, 20) (published: 24.3 s on one H100)."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import thermal as th

    from qed_spin_amd import _lib

    n = 28
    H = _ref_chain(n)
    T = np.linspace(0.1, 4.0, 20)
    _lib.device_count()     # the runtime's start-up is not part of the solve
    _lib.call("qedh_synchronize")
    t0 = time.perf_counter()
    r = th.thermal(H, T, sym=qs.Symmetry(), samples=4, krylov=60, seed=7)
    wall = time.perf_counter() - t0
    out = {"chain28_ftlm": {"wall_s": round(wall, 3), "E(Tmin)": float(r.E[0]), "C_max": float(np.max(r.C)),
                            "build_split": {k: round(v, 3) for k, v in th.WALK_TIMES.items()},
                            "blocks": len(r.blocks), "sampled_blocks": sum(1 for b in r.blocks if not b["exact"]),
                            "timings": {k: round(v, 3) for k, v in r.timings.items()},
                            "engine": {k: round(v, 4) for k, v in r.engine.items()},
                            "published_gpu_s": 24.3, "published_hw": "1x H100"}}
    log(f"chain28_ftlm: {out['chain28_ftlm']}")
    return out


def run_reference_eigs(args):
    """The reference's published GPU eigs case chain32_abelian_eigs_gpu, same inputs
    (bench/cases.py:62-65): the lowest level of chain(32) (J1-J2 ring, J2 = 0.35) at Sz = 0 over every
    momentum and spin-flip block (64 blocks of ~9.4e6 states), pruned as the reference's walk prunes
    (published: 64.9 s on one H100). The blocks are streamed: one per build thread on the device at
    a time."""
    import qed_spin_amd as qs
    from qed_spin_amd import _lib
    from qed_spin_amd import eigs as ev

    n = 32
    H = _ref_chain(n)
    _lib.device_count()
    _lib.call("qedh_synchronize")
    t0 = time.perf_counter()
    r = ev.eigs(H, 1, sym=qs.Symmetry(sz=[16], spin_flip="all"), stream=True)
    wall = time.perf_counter() - t0
    solved = [b for b in r.blocks if b["lane"] != "pruned"]
    out = {"chain32_abelian_eigs": {"wall_s": round(wall, 3), "E0": float(r.energies[0]), "blocks": len(r.blocks),
                                    "solved_blocks": len(solved), "pruned_blocks": r.timings["pruned_blocks"],
                                    "max_block_dim": max(b["dim"] for b in r.blocks),
                                    "timings": {k: round(float(v), 3) for k, v in r.timings.items()},
                                    "published_gpu_s": 64.9, "published_hw": "1x H100"}}
    log(f"chain32_abelian_eigs: {out['chain32_abelian_eigs']}")
    return out


def run_spectrum(args):
    """The whole spectrum of the Heisenberg ring over every (Sz, k) block: the dense lanes (blocks
    from 1024 states on the device, rocSOLVER zheevd batched per dimension; smaller ones on the host
    pool). Reported: wall time, the blocks on each lane and the ground-state energy."""
    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import eigs as ev

    n = args.spectrum_sites
    H = qs.HamiltonianBuilder(n).heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0).to_operator()
    t0 = time.perf_counter()
    r = ev.spectrum(H, sym=qs.Symmetry())
    wall = time.perf_counter() - t0
    dev = [b["dim"] for b in r.blocks if b["lane"] == "device-dense"]
    host = [b["dim"] for b in r.blocks if b["lane"] == "host-dense"]
    out = {
        "workload": f"whole spectrum of the Heisenberg ring N={n} over every (Sz, k) block",
        "eigenvalues": int(r.eigenvalues.size),
        "blocks_device": len(dev),
        "blocks_host": len(host),
        "largest_block": int(max(dev + host)),
        "wall_s": round(wall, 3),
        "timings": {k: round(v, 3) for k, v in r.timings.items()},
        "E0": float(r.eigenvalues[0]),
        "trace": float(np.sum(r.eigenvalues)),
    }
    log(f"spectrum: {out}")
    return out


def run_serve(args):
    """The mixed request stream: served at saturation (every request there at 0: solves/s), served
    with Poisson arrivals (latency p50/p99), and the plain path (one request after another, no
    batching, no sector cache) on the same stream's first requests against the served path on the
    same prefix."""
    from qed_spin_amd import serve as sv

    def brief(st):
        keys = ("wall_s", "solves_per_s", "p50_s", "p99_s", "mean_s", "completed", "failed", "peak_in_flight",
                "peak_reserved", "sector_builds", "sector_hits", "build_s", "service_eigs_s", "service_thermal_s",
                "service_dynamics_s")
        return {k: (round(st[k], 4) if isinstance(st[k], float) else st[k]) for k in keys}

    reqs = sv.mixed_stream(n=args.serve_requests, rate=args.serve_rate, seed=3, n_max=args.serve_nmax)
    out = {"workload": f"{len(reqs)} mixed requests (eigs k=4 / thermal FTLM 20x80 / dynamics S^z(pi) 150, "
                       f"50/30/20), N in 16..{args.serve_nmax}, seed 3",
           "slots": args.serve_slots}
    sat = sv.serve(reqs, slots=args.serve_slots, realtime=False, progress_s=20.0)
    out["served_saturated"] = brief(sat.stats)
    log(f"serve saturated: {out['served_saturated']}")
    if args.serve_realtime_requests > 0:
        rt = sv.serve(reqs[: args.serve_realtime_requests], slots=args.serve_slots, realtime=True, progress_s=20.0)
        out["served_poisson"] = dict(brief(rt.stats), rate_per_s=args.serve_rate, requests=args.serve_realtime_requests)
        log(f"serve poisson: {out['served_poisson']}")
    if args.serve_plain_requests > 0:
        m = args.serve_plain_requests
        plain = sv.serve(reqs[:m], slots=1, batched=False, cache_sectors=False, realtime=False, progress_s=20.0)
        same = sv.serve(reqs[:m], slots=args.serve_slots, realtime=False, progress_s=20.0)
        out["plain_prefix"] = dict(brief(plain.stats), requests=m)
        out["served_prefix"] = dict(brief(same.stats), requests=m)
        out["speedup_solves_per_s"] = round(same.stats["solves_per_s"] / plain.stats["solves_per_s"], 3)
        log(f"serve plain vs served on {m}: {out['plain_prefix']} / {out['served_prefix']}")
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lanczos-steps", type=int, default=300)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--workload", default="matvec,config1,ftlm,dynamics,dynT,serve,refftlm,refeigs",
                    help="matvec (the headline) and a comma list of ftlm, dynamics, dynT, serve, spectrum, refdyn, refftlm, refeigs")
    ap.add_argument("--ftlm-sites", type=int, default=24)
    ap.add_argument("--dyn-sites", type=int, default=28)
    ap.add_argument("--spectrum-sites", type=int, default=18)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-nmax", type=int, default=30)
    ap.add_argument("--serve-plain-requests", type=int, default=50)
    ap.add_argument("--serve-realtime-requests", type=int, default=300)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device to device (the known-size kernel the HBM counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box here has one; run --gpus 1"}))
        sys.exit(2)

    import numpy as np

    import qed_spin_amd as qs
    from qed_spin_amd import _lib
    from qed_spin_amd import symmetry as sym

    if _lib.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.calibrate_copy:
        src, dst = qs.DeviceArray(1 << 27, np.float64), qs.DeviceArray(1 << 27, np.float64)
        src.zero()
        _lib.call("qedh_memcpy_dtod", dst.ptr, src.ptr, C.c_size_t(1 << 30))
        _lib.call("qedh_synchronize")
        del src, dst
        log("calibration copy: 1 GiB device to device")

    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    b.heisenberg([(i, (i + 2) % N) for i in range(N)], J2)
    op = b.to_operator()

    t0 = time.perf_counter()
    sector = qs.Sector(sym.dihedral_group(N, 0, +1), N // 2)
    t_sector = time.perf_counter() - t0
    log(f"sector dim={sector.dim} orbits={sector.n_orbits} sublattice_L={sector.sublattice_block} "
        f"build {t_sector:.2f}s")
    t1 = time.perf_counter()
    m = qs.SectorMatvec(op, sector, lane=qs.LANE_CSR)
    t_csr = time.perf_counter() - t1
    if args.lanes:
        m.set_lanes(args.lanes)
    info = m.info
    D, nnz = sector.dim, int(info["nnz"])
    w_id = int(info["id_width_ref"])
    s = 8
    B1 = D * 8 + nnz * (4 + w_id) + (nnz * s + D * s)
    log(f"csr nnz={nnz} ({nnz / D:.2f}/row) dict={info['dict_size']} bytes/value={info['value_bytes']} "
        f"w_id(ref)={w_id} lanes={info['lanes']} build {t_csr:.2f}s; B(1)={B1 / 1e9:.3f} GB")

    avg = C.c_double()
    _lib.call("qedh_matvec_benchmark", m.h, 1, 0, 0, int(args.warmup), int(args.steps), C.byref(avg))
    t_apply = float(avg.value)
    gbs = B1 / t_apply / 1e9
    log(f"apply {t_apply * 1e3:.3f} ms -> {gbs:.1f} GB/s algorithmic ({100 * gbs / HBM_PEAK_GBS:.1f}% of peak)")

    # multi-vector applies (the batched regime), for the record
    multi = {}
    for k in (2, 4, 8):
        _lib.call("qedh_matvec_benchmark", m.h, k, 1, 0, 3, 10, C.byref(avg))
        Bk = D * 8 + nnz * (4 + w_id) + k * (nnz * s + D * s)
        multi[str(k)] = {"ms": avg.value * 1e3, "GBps": Bk / avg.value / 1e9}
        log(f"block k={k}: {avg.value * 1e3:.3f} ms -> {Bk / avg.value / 1e9:.1f} GB/s")

    # the on-the-fly (gather) lane on the same sector and program, scored by the same B(1): the lane
    # the engine falls back on when a CSR does not fit its budget
    mg = qs.SectorMatvec(op, sector, lane=qs.LANE_GATHER)
    _lib.call("qedh_matvec_benchmark", mg.h, 1, 0, 0, 2, 10, C.byref(avg))
    gather = {"ms": round(avg.value * 1e3, 3), "GBps": round(B1 / avg.value / 1e9, 1),
              "frac": round(B1 / avg.value / 1e9 / HBM_PEAK_GBS, 4),
              "note": "on-the-fly lane (no CSR), by the merged CSR's B(1)"}
    log(f"gather lane: {gather}")
    del mg

    lanczos = None
    if not args.no_lanczos:
        rng = np.random.default_rng(1)
        v0 = rng.normal(size=D)
        v0 /= np.linalg.norm(v0)
        t2 = time.perf_counter()
        a, bb, tm = m.lanczos(v0, args.lanczos_steps, reorth=1)
        t_lz = time.perf_counter() - t2
        from qed_spin_amd.krylov import tridiag_lowest
        e0 = tridiag_lowest(a, bb)
        lanczos = {
            "steps": int(len(a)),
            "reorth": "FullCGS2",
            "E0": e0,
            "E0_exact": -3.0 * N / 8.0,
            "E0_rel_err": abs(e0 + 3.0 * N / 8.0) / (3.0 * N / 8.0),
            "seconds": t_lz,
            "apply_s": tm["apply_s"],
            "reorth_s": tm["reorth_s"],
            "vector_s": tm["vector_s"],
            "time_to_E0_s": t_sector + t_csr + t_lz,
        }
        log(f"lanczos {len(a)} steps: E0={e0:.12f} in {t_lz:.2f}s (apply {tm['apply_s']:.2f}s, "
            f"reorth {tm['reorth_s']:.2f}s); time to E0 {t_sector + t_csr + t_lz:.2f}s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, D, w_id)
        log(f"cpu baseline: {cpu}")

    # HBM bytes per launch of this SpMV from the committed PMC summary (read requests by size and
    # WRITE_SIZE, calibrated on a 1 GiB copy), valid for the layout it was taken on (lanes per row:
    # 1 is SELL-64); not measured in this run
    traffic, traffic_from = None, None
    tf = os.path.join(ROOT, "profiles", "spmv_n32_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            rec = json.load(f)
        if int(rec.get("lanes_per_row", -1)) == int(info["lanes"]):
            traffic = rec.get("hbm_bytes_per_launch")
            traffic_from = rec.get("profile")

    out = {
        "metric": "FP64 sector matvec GB/s (% HBM peak)",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": int(args.steps),
        "warmup": int(args.warmup),
        "ms_per_step": round(t_apply * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the J1-J2 ring Hamiltonian itself; random unit input vector)",
        "config": {
            "workload": "J1-J2 Heisenberg ring N=32, J1=1, J2=0.5, periodic, Sz=0, k=0, reflection-even",
            "dim": D,
            "nnz": nnz,
            "bytes_per_apply": B1,
            "value_bytes": int(info["value_bytes"]),
            "dict_size": int(info["dict_size"]),
            "lanes_per_row": int(info["lanes"]),
            "sector_build_s": round(t_sector, 3),
            "csr_build_s": round(t_csr, 3),
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_from": traffic_from,
        },
        "multi_vector": multi,
        "gather_lane": gather,
        "lanczos_to_E0": lanczos,
        "cpu_baseline": cpu,
    }
    extra = [w.strip() for w in args.workload.split(",") if w.strip() and w.strip() != "matvec"]
    for w in extra:
        if w == "ftlm":
            out["ftlm"] = run_ftlm(args)
        elif w == "config1":
            out["config1"] = run_config1(args)
        elif w == "dynamics":
            out["dynamics"] = run_dynamics(args)
        elif w == "dynT":
            out["dynamics_finite_T"] = run_dynamics_finite_T(args)
        elif w == "serve":
            out["serve"] = run_serve(args)
        elif w == "spectrum":
            out["spectrum"] = run_spectrum(args)
        elif w == "refdyn":
            out["reference_dynamics"] = run_reference_dynamics(args)
        elif w == "observe":
            out["observe"] = run_observe(args)
            log(f"observe: {out['observe']}")
        elif w == "refftlm":
            out["reference_ftlm"] = run_reference_ftlm(args)
        elif w == "refeigs":
            out["reference_eigs"] = run_reference_eigs(args)
        else:
            raise SystemExit(f"unknown workload {w!r}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

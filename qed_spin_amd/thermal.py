"""Thermodynamics of a spin Hamiltonian over its symmetry blocks (the reference's qed.thermal /
ed::sectors::thermal, src/engine/thermal.cpp:345-650).

Every non-empty block of the (Sz, k) walk gets its own seed from the reference's per-block stream
(seed <- seed * 6364136223846793005 + 1442695040888963407, in walk order, thermal.cpp:471). Blocks of
at most ``dense_max_dim`` states are solved exactly (their whole spectrum, thermal.cpp:197-200);
all other blocks run FTLM together in the batched engine: every block and every random vector of
the walk advance through one packed multi-vector matvec per Lanczos step, the step replayed as a
hipGraph (qedh_ftlm_batched). The blocks' curves are combined by log-sum-exp with the law of total
variance (thermal.cpp:597-640).
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import os
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .errors import DeviceUnavailable, DeviceUnsupported, InvalidRequest, ResourceLimit
from .operator import Operator
from .sector import AUTOTUNE_MIN_ROWS, DEFAULT_CSR_BUDGET, LANE_CSR, Sector, SectorMatvec, parallel_build
from .symmetry import Symmetry

LCG_MUL = 6364136223846793005
LCG_ADD = 1442695040888963407
_U64 = (1 << 64) - 1
KBREAKDOWN_REL = 64.0 * np.finfo(np.float64).eps


# method='exact': the dense lanes' largest block (16 GiB as a dense complex matrix on the device)
EXACT_MAX_DIM = 32768

def next_block_seed(seed: int) -> int:
    """The next block's seed of the thermal walk (thermal.cpp:471)."""
    return (int(seed) * LCG_MUL + LCG_ADD) & _U64


def resolve_device(device: str, verb: str) -> None:
    """The reference's device= keyword (python/qed/_verbs/_device.py:18-29): "gpu" and "auto" run
    on the HIP device and raise DeviceUnavailable without one; "cpu" names the reference's host path,
    which this backend does not provide."""
    if device not in ("cpu", "gpu", "auto"):
        raise InvalidRequest(f"{verb}: device must be 'cpu', 'gpu' or 'auto', got {device!r}")
    if device == "cpu":
        raise DeviceUnsupported(f"{verb}: this backend runs on the HIP device; device='cpu' is the host path")
    if _lib.device_count() < 1:
        raise DeviceUnavailable(f"{verb}: no HIP device is visible")


def _handles(mats: Sequence[SectorMatvec]):
    return (C.c_void_p * len(mats))(*[m.h.value for m in mats])


def batch_lanczos(mats: Sequence[SectorMatvec], R: int, krylov: int, start: np.ndarray,
                  breakdown_tol: Optional[Sequence[float]] = None, use_graph: bool = True, reorth: int = 2):
    """Lanczos of R vectors per block, all blocks at once (reorth 2: LocalDGKS3, 0: none). start:
    complex (rows, R) with the blocks' rows stacked in order. Returns alpha, beta (krylov, blocks, R),
    steps (blocks, R)."""
    nb = len(mats)
    rows = sum(m.dim for m in mats)
    st = np.ascontiguousarray(np.asarray(start, np.complex128).reshape(rows, R))
    tol = np.ascontiguousarray(
        [KBREAKDOWN_REL * (m.op.norm_bound() or 1.0) for m in mats] if breakdown_tol is None else breakdown_tol,
        np.float64)
    alpha = np.zeros((krylov, nb, R))
    beta = np.zeros((krylov, nb, R))
    steps = np.zeros((nb, R), np.int32)
    t = np.zeros(3)
    _lib.call("qedh_batch_lanczos", _handles(mats), nb, int(R), int(krylov),
              st.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), _lib.ptr(tol, C.c_double),
              int(reorth), int(bool(use_graph)), _lib.ptr(alpha, C.c_double), _lib.ptr(beta, C.c_double),
              _lib.ptr(steps, C.c_int), _lib.ptr(t, C.c_double))
    return alpha, beta, steps, {"pack_s": t[0], "steps_s": t[1], "total_s": t[2]}


def ftlm_batched(mats: Sequence[SectorMatvec], samples: int, krylov: int, seeds: Sequence[int], betas,
                 breakdown_tol: Optional[Sequence[float]] = None, use_graph: bool = True):
    """FTLM of every block at once; block b draws its samples from seeds[b]. Returns lnZ, E, V
    (blocks, temperatures), e_min (blocks) and the engine's timings."""
    nb = len(mats)
    betas = np.ascontiguousarray(betas, np.float64)
    nT = betas.size
    sd = np.ascontiguousarray([int(s) & _U64 for s in seeds], np.uint64)
    tol = None if breakdown_tol is None else np.ascontiguousarray(breakdown_tol, np.float64)
    lnZ, E, V = np.zeros((nb, nT)), np.zeros((nb, nT)), np.zeros((nb, nT))
    emin = np.zeros(nb)
    stats = np.zeros(10)
    _lib.call("qedh_ftlm_batched", _handles(mats), nb, int(samples), int(krylov), _lib.ptr(sd, C.c_uint64),
              _lib.ptr(tol, C.c_double), int(nT), _lib.ptr(betas, C.c_double), int(bool(use_graph)),
              _lib.ptr(lnZ, C.c_double), _lib.ptr(E, C.c_double), _lib.ptr(V, C.c_double),
              _lib.ptr(emin, C.c_double), _lib.ptr(stats, C.c_double))
    keys = ["pack_s", "seed_s", "steps_s", "post_s", "total_s", "applies", "rows", "nnz", "schedule", "tune_s"]
    return lnZ, E, V, emin, dict(zip(keys, stats.tolist()))


def exact_thermal(mats: Sequence[SectorMatvec], betas):
    """Exact curves of small blocks (whole spectra of the dense blocks)."""
    nb = len(mats)
    betas = np.ascontiguousarray(betas, np.float64)
    nT = betas.size
    lnZ, E, V = np.zeros((nb, nT)), np.zeros((nb, nT)), np.zeros((nb, nT))
    emin = np.zeros(nb)
    _lib.call("qedh_exact_thermal", _handles(mats), nb, int(nT), _lib.ptr(betas, C.c_double),
              _lib.ptr(lnZ, C.c_double), _lib.ptr(E, C.c_double), _lib.ptr(V, C.c_double), _lib.ptr(emin, C.c_double))
    return lnZ, E, V, emin


def combine(betas, weights, lnZ, E, V) -> Dict[str, np.ndarray]:
    betas = np.ascontiguousarray(betas, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    lnZ, E, V = (np.ascontiguousarray(a, np.float64) for a in (lnZ, E, V))
    nb, nT = lnZ.shape
    out = {k: np.zeros(nT) for k in ("lnZ", "E", "C", "S", "F")}
    _lib.call("qedh_thermal_combine", int(nb), int(nT), _lib.ptr(betas, C.c_double), _lib.ptr(w, C.c_double),
              _lib.ptr(lnZ, C.c_double), _lib.ptr(E, C.c_double), _lib.ptr(V, C.c_double),
              *[_lib.ptr(out[k], C.c_double) for k in ("lnZ", "E", "C", "S", "F")])
    return out


@dataclass
class ThermalResult:
    """The reference's ThermalCurves (include/ed/sectors/thermal.h): per temperature ln Z, E, C, S,
    F; e0 the lowest level any block resolved; per-block records and timings."""

    T: np.ndarray
    lnZ: np.ndarray
    E: np.ndarray
    C: np.ndarray
    S: np.ndarray
    F: np.ndarray
    e0: float
    total_dim: int
    blocks: List[dict] = field(default_factory=list)
    timings: Dict[str, float] = field(default_factory=dict)
    engine: Dict[str, float] = field(default_factory=dict)


# the last walk's split of its build time (sectors, then operators), for the benchmarks
WALK_TIMES: Dict[str, float] = {}

# Device objects of finished walks being released on a thread of their own: a walk's hundreds of
# sectors and operators take seconds to free, which need not delay the verb's return. The next walk
# waits for them first, so their memory is back in the pool before it allocates.
_RELEASING: List[threading.Thread] = []


def release_later(objs: list) -> None:
    """Drop the list's references (and so free the device objects only it holds) on another thread."""
    t = threading.Thread(target=objs.clear, name="qsa-release")
    t.start()
    _RELEASING.append(t)


def wait_released() -> None:
    while _RELEASING:
        _RELEASING.pop().join()


class LazyMatvec:
    """A block's sector operator built when first used (any attribute): a walk whose blocks are not
    all applied -- a T = 0 dynamics run needs the ground search's blocks and the probes' targets --
    builds only those. Safe to first use from several threads."""

    def __init__(self, make):
        self._make = make
        self._m = None
        self._lock = threading.Lock()

    def _get(self):
        if self._m is None:
            with self._lock:
                if self._m is None:
                    self._m = self._make()
        return self._m

    def __getattr__(self, name):
        return getattr(self._get(), name)


def time_reversal_partners(H: Operator, blks) -> Dict[int, int]:
    """{index of a momentum-m block with 2m > N: index of its momentum-(N - m) partner} for a real H:
    time reversal maps the one block onto the complex conjugate of the other (same representatives
    and norms, conjugate characters), so the two have one spectrum and one trace of any function of H."""
    partner: Dict[int, int] = {}
    if not blks or not H.is_real():
        return partner
    N = H.num_sites
    where = {(b.n_up, b.momentum, b.parity, b.flip): i for i, b in enumerate(blks)}
    for i, b in enumerate(blks):
        if b.momentum is not None and 2 * b.momentum > N:
            p = where.get((b.n_up, N - b.momentum, b.parity, b.flip))
            if p is not None and np.allclose(np.asarray(blks[p].group.chars).conj(), b.group.chars):
                partner[i] = p
    return partner


def walk_blocks(H: Operator, sym: Symmetry, csr_budget: int = DEFAULT_CSR_BUDGET, tune: bool = True,
                fold: bool = False, lazy: bool = False):
    """The non-empty blocks of the walk in the reference's order (n_up ascending, then momentum),
    each with its sector operator on the device CSR lane. Sectors and the operators of small
    sectors are built on several host threads at once; operators large enough to time their SpMV
    variants are built one after another, so the timings see an otherwise idle device. tune=False
    (walks whose applies all run through a batched engine, which packs the blocks' CSRs and has
    kernels of its own): no operator times its variants and all are built on the threads. fold (a
    real H): the second block of each time-reversal pair is left out and the first stands for both
    (Block.weight 2), as the reference's walk folds them. lazy: the operators are LazyMatvec, built
    on first use."""
    blks = sym.blocks(H.num_sites)
    if fold:
        partner = time_reversal_partners(H, blks)
        for p in partner.values():
            blks[p] = dataclasses.replace(blks[p], weight=2 * blks[p].weight)
        blks = [b for i, b in enumerate(blks) if i not in partner]
    H._handle()
    wait_released()
    t0 = time.perf_counter()
    secs = parallel_build(lambda b: Sector(b.group, b.n_up), blks)
    WALK_TIMES["sectors_s"] = time.perf_counter() - t0
    keep = [(b, s) for b, s in zip(blks, secs) if s.dim > 0]
    mats = [None] * len(keep)
    small = [i for i, (_, s) in enumerate(keep) if s.dim < AUTOTUNE_MIN_ROWS or not tune]
    make = lambda i: SectorMatvec(H, keep[i][1], lane=LANE_CSR, csr_budget=csr_budget, tune=tune)
    if lazy:
        mats = [LazyMatvec(lambda i=i: make(i)) for i in range(len(keep))]
        small = []
    for i, m in zip(small, parallel_build(make, small)):
        mats[i] = m
    for i in range(len(keep)):
        if mats[i] is None:
            mats[i] = make(i)
    WALK_TIMES["operators_s"] = time.perf_counter() - t0 - WALK_TIMES["sectors_s"]
    release_later(secs)    # the empty sectors (the kept ones live on in the walk)
    return [(b, s, m) for (b, s), m in zip(keep, mats)]


def thermal(H: Operator, T, *, sym: Optional[Symmetry] = None, method: str = "ftlm", samples: int = 40,
            krylov: int = 100, seed: int = 0, device: str = "gpu", dense_max_dim: Optional[int] = None,
            use_graph: bool = True, csr_budget: int = DEFAULT_CSR_BUDGET) -> ThermalResult:
    """FTLM thermodynamics of H at temperatures T over the blocks of `sym` (default: every (Sz, k)
    block of a ring; with a real H each time-reversal pair of momentum blocks is one block of weight
    2, as the reference's walk folds them). seed 0 draws one. dense_max_dim (default 512, the reference's ThermalSpec
    default): blocks up to this dimension are solved exactly; 0 samples every block. method="exact":
    every block's whole spectrum through the dense lanes (blocks of at most EXACT_MAX_DIM states)."""
    resolve_device(device, "thermal")
    if method not in ("ftlm", "exact"):
        raise DeviceUnsupported(f"thermal: method {method!r} is not on the device path (ftlm and exact only)")
    T = np.asarray(T, np.float64).ravel()
    if T.size == 0 or np.any(~(T > 0)):
        raise InvalidRequest("thermal: temperatures must be positive")
    if samples < 1 or krylov < 1:
        raise InvalidRequest("thermal: samples and krylov must be >= 1")
    dmax = 512 if dense_max_dim is None else int(dense_max_dim)
    if dmax < 0:
        raise InvalidRequest(f"dense_max_dim must be >= 0 or None, got {dense_max_dim}")
    if method == "exact":   # every block through the dense lanes
        dmax = EXACT_MAX_DIM
    betas = 1.0 / T
    sym = sym if sym is not None else Symmetry()
    s = int(seed) if seed else int.from_bytes(os.urandom(8), "little")
    t0 = time.perf_counter()
    # every apply runs in the batched engine; time-reversal pairs fold into one block of weight 2
    blocks = walk_blocks(H, sym, csr_budget, tune=False, fold=True)
    t_build = time.perf_counter() - t0
    if not blocks:
        raise InvalidRequest("thermal: the symmetry selection holds no state")
    if method == "exact" and max(sec.dim for _, sec, _ in blocks) > EXACT_MAX_DIM:
        raise ResourceLimit(f"thermal: method 'exact' solves every block densely, at most {EXACT_MAX_DIM} states")
    recs = []
    for blk, sec, m in blocks:
        s = next_block_seed(s)
        recs.append({"label": blk.label, "dim": sec.dim, "seed": s, "exact": sec.dim <= dmax,
                     "weight": float(blk.weight)})
    nb, nT = len(blocks), betas.size
    lnZ, E, V, emin = np.zeros((nb, nT)), np.zeros((nb, nT)), np.zeros((nb, nT)), np.zeros(nb)
    ex = [i for i, r in enumerate(recs) if r["exact"]]
    sm = [i for i, r in enumerate(recs) if not r["exact"]]
    t1 = time.perf_counter()
    if ex:
        a = exact_thermal([blocks[i][2] for i in ex], betas)
        for arr, src in zip((lnZ, E, V, emin), a):
            arr[ex] = src
    t_exact = time.perf_counter() - t1
    engine = {}
    t2 = time.perf_counter()
    if sm:
        mats = [blocks[i][2] for i in sm]
        a = ftlm_batched(mats, samples, krylov, [recs[i]["seed"] for i in sm], betas, use_graph=use_graph)
        for arr, src in zip((lnZ, E, V, emin), a[:4]):
            arr[sm] = src
        engine = a[4]
    t_ftlm = time.perf_counter() - t2
    c = combine(betas, [r["weight"] for r in recs], lnZ, E, V)
    for i, r in enumerate(recs):
        r.update(lnZ=lnZ[i].copy(), E=E[i].copy(), V=V[i].copy(), e_min=float(emin[i]))
    mats = None
    release_later(blocks)
    return ThermalResult(
        T=T, lnZ=c["lnZ"], E=c["E"], C=c["C"], S=c["S"], F=c["F"], e0=float(np.min(emin)),
        total_dim=int(sum(r["dim"] * r["weight"] for r in recs)), blocks=recs,
        timings={"build_s": t_build, "exact_s": t_exact, "ftlm_s": t_ftlm, "total_s": time.perf_counter() - t0},
        engine=engine)

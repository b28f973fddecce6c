"""The serving loop over the C-ABI (qedh_serve): a stream of eigs / thermal / dynamics solves on spin
chains, admitted against an HBM budget, run concurrently on slots with a HIP stream each, sectors
resident between requests; and the mixed stream the benchmark serves.

The stream: Poisson arrivals, each request an eigs(k=4) / thermal(FTLM, 20 samples, depth 80) /
dynamics(S^z(pi), krylov 150) drawn 50/30/20, on a periodic chain of N uniform in {16, ..., 30}
with J1 ~ U[0.5, 1.5], J2 ~ U[0, 0.6] and chirality chi ~ U[-0.3, 0.3] (complex sectors), in the
sector n_up = N / 2, momentum 0.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from .errors import InvalidRequest
from .thermal import resolve_device

KIND = {"eigs": 0, "thermal": 1, "dynamics": 2}


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("n_sites", C.c_int),
        ("j1", C.c_double),
        ("j2", C.c_double),
        ("chi", C.c_double),
        ("arrival_s", C.c_double),
        ("k", C.c_int),
        ("samples", C.c_int),
        ("depth", C.c_int),
        ("krylov", C.c_int),
        ("seed", C.c_uint64),
    ]


class ServeOptions(C.Structure):
    _fields_ = [
        ("slots", C.c_int),
        ("batched", C.c_int),
        ("cache_sectors", C.c_int),
        ("realtime", C.c_int),
        ("hbm_budget", C.c_uint64),
        ("cache_budget", C.c_uint64),
        ("n_beta", C.c_int),
        ("betas", C.POINTER(C.c_double)),
        ("n_omega", C.c_int),
        ("omega", C.POINTER(C.c_double)),
        ("eta", C.c_double),
        ("progress_s", C.c_double),
    ]


DEFAULT_T = np.linspace(0.1, 4.0, 40)
DEFAULT_OMEGA = np.linspace(0.0, 4.0, 400)


def mixed_stream(n: int = 2000, rate: float = 50.0, seed: int = 3, n_min: int = 16, n_max: int = 30,
                 mix=(0.5, 0.3, 0.2)) -> List[dict]:
    """The benchmark's request stream (numpy default_rng(seed)): Poisson arrivals at `rate` per
    second, kinds drawn by `mix`, N uniform in {n_min..n_max}, couplings uniform."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    kinds = rng.choice(3, size=n, p=list(mix))
    Ns = rng.integers(n_min, n_max + 1, size=n)
    j1 = rng.uniform(0.5, 1.5, size=n)
    j2 = rng.uniform(0.0, 0.6, size=n)
    chi = rng.uniform(-0.3, 0.3, size=n)
    out = []
    for i in range(n):
        out.append(dict(kind=int(kinds[i]), n_sites=int(Ns[i]), j1=float(j1[i]), j2=float(j2[i]), chi=float(chi[i]),
                        arrival_s=float(t[i]), k=4, samples=20, depth=80, krylov=150, seed=1000 + i))
    return out


@dataclass
class ServeResult:
    values: List[np.ndarray]
    e0: np.ndarray
    timing: np.ndarray            # per request: arrival, start, finish, latency
    status: np.ndarray
    dims: np.ndarray
    stats: Dict[str, float] = field(default_factory=dict)


def serve(requests: List[dict], *, slots: int = 8, batched: bool = True, cache_sectors: bool = True,
          realtime: bool = True, hbm_budget: int = 0, cache_budget: int = 0, T=None, omega=None, eta: float = 0.05,
          device: str = "gpu", progress_s: float = 0.0) -> ServeResult:
    resolve_device(device, "serve")
    if not requests:
        raise InvalidRequest("serve: no request")
    T = DEFAULT_T if T is None else np.asarray(T, np.float64)
    betas = np.ascontiguousarray(1.0 / T, np.float64)
    omega = np.ascontiguousarray(DEFAULT_OMEGA if omega is None else omega, np.float64)
    n = len(requests)
    arr = (Request * n)()
    for i, r in enumerate(requests):
        q = arr[i]
        q.kind = int(KIND.get(r["kind"], r["kind"])) if isinstance(r["kind"], str) else int(r["kind"])
        q.n_sites = int(r["n_sites"])
        q.j1, q.j2, q.chi = float(r.get("j1", 1.0)), float(r.get("j2", 0.0)), float(r.get("chi", 0.0))
        q.arrival_s = float(r.get("arrival_s", 0.0))
        q.k = int(r.get("k", 4))
        q.samples = int(r.get("samples", 20))
        q.depth = int(r.get("depth", 80))
        q.krylov = int(r.get("krylov", 150))
        q.seed = int(r.get("seed", 1)) & ((1 << 64) - 1)
    o = ServeOptions()
    o.slots = int(slots)
    o.batched = int(bool(batched))
    o.cache_sectors = int(bool(cache_sectors))
    o.realtime = int(bool(realtime))
    o.hbm_budget = int(hbm_budget)
    o.cache_budget = int(cache_budget)
    o.n_beta = betas.size
    o.betas = _lib.ptr(betas, C.c_double)
    o.n_omega = omega.size
    o.omega = _lib.ptr(omega, C.c_double)
    o.eta = float(eta)
    o.progress_s = float(progress_s)
    stride = max(max(int(r.get("k", 4)) for r in requests), 3 * betas.size, omega.size)
    values = np.zeros((n, stride))
    e0 = np.zeros(n)
    timing = np.zeros((n, 4))
    status = np.zeros(n, np.int32)
    dims = np.zeros(n, np.int64)
    stats = np.zeros(16)
    _lib.call("qedh_serve", arr, n, C.byref(o), int(stride), _lib.ptr(values, C.c_double), _lib.ptr(e0, C.c_double),
              _lib.ptr(timing, C.c_double), _lib.ptr(status, C.c_int), _lib.ptr(dims, C.c_int64),
              _lib.ptr(stats, C.c_double))
    out_vals = []
    for i, r in enumerate(requests):
        kind = arr[i].kind
        width = arr[i].k if kind == 0 else (3 * betas.size if kind == 1 else omega.size)
        if kind == 0:
            width = min(width, int(dims[i])) if dims[i] > 0 else width
        out_vals.append(values[i, :width].copy())
    keys = ["wall_s", "solves_per_s", "p50_s", "p99_s", "mean_s", "max_s", "completed", "failed", "peak_reserved",
            "peak_in_flight", "sector_builds", "sector_hits", "build_s", "service_eigs_s", "service_thermal_s",
            "service_dynamics_s"]
    return ServeResult(values=out_vals, e0=e0, timing=timing, status=status, dims=dims,
                       stats=dict(zip(keys, stats.tolist())))

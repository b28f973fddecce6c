"""Dynamical correlation functions S_O(omega) at T = 0 (the reference's qed.dynamics /
ed::sectors::dynamics, src/engine/dynamics.cpp:282-690).

The ground manifold is every level within degeneracy_tol * s_H of the lowest level E0 over the
blocks of the walk (include/ed/sectors/dynamics.h:17-42). For every ground state |g> (block b) and
probe O, every block t the probe reaches gets phi = O|g> through the cross-sector operator
(CrossSectorMatVec, src/engine/internal.h:552-609, on the device), and each phi one continued
fraction on H of block t with the shift E0 and weight ||phi||^2 (cf.h:103-161). All (state, probe,
target) continued fractions run together in the batched engine (qedh_cf_batched): the targets'
CSRs packed into one block-diagonal matrix, the probe images as its right-hand sides, one
multi-vector matvec per Lanczos step replayed as a hipGraph. S_O is the sum over targets, averaged
over the ground manifold (dynamics.cpp:680-684).
"""

from __future__ import annotations

import cmath
import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .errors import DeviceUnsupported, InvalidRequest
from .operator import Operator, Program
from .sector import parallel_build, projected_program, DEFAULT_CSR_BUDGET, LANE_CSR, LANE_GATHER, Sector, SectorMatvec
from .symmetry import Symmetry
from .thermal import KBREAKDOWN_REL, _handles, release_later, resolve_device, time_reversal_partners, walk_blocks


def sz_q(N: int, q: float, sign: int = -1) -> Operator:
    """S^z(q) = N^-1/2 sum_j e^{sign i q j} S^z_j."""
    op = Operator(N)
    for j in range(N):
        op.add_product("z", (j,), cmath.exp(sign * 1j * q * j) / math.sqrt(N))
    return op


def _delta_up(O: Operator) -> int:
    """The change of up spins every product of O makes (one value, or the request is refused)."""
    ds = set()
    for ops, _, _ in O.products:
        d = 0
        for o in ops:
            d += {"+": 1, "-": -1}.get(o, 0)
            if o in "xy":
                raise DeviceUnsupported("dynamics: a probe must change Sz by one definite amount (use S+, S-, Sz)")
        ds.add(d)
    if len(ds) > 1:
        raise DeviceUnsupported("dynamics: a probe must change Sz by one definite amount")
    return ds.pop() if ds else 0


class _DevView:
    """A device pointer into a larger allocation (no ownership)."""

    def __init__(self, base: int, offset_bytes: int):
        self.ptr = C.c_void_p(base + offset_bytes)


@dataclass
class DynamicsResult:
    omega: np.ndarray
    S: List[np.ndarray]            # one spectral function per probe
    e0: float
    ground_manifold: int
    target_sectors: int
    timings: Dict[str, float] = field(default_factory=dict)
    engine: Dict[str, float] = field(default_factory=dict)
    norms: Optional[np.ndarray] = None   # T = 0: per probe, ||O|g>||^2 averaged over the ground manifold


def ground_manifold(blocks, window: float, k_max: int = 16, only=None, prune: bool = True, norm_bound: float = 1.0,
                    partners=None):
    """(E0, [(block index, device buffer, vector index, energy)]): every level within `window` of the
    lowest level of the walk (src/engine/dynamics.cpp:147-193). Each block's lowest level comes from a
    k = 1 eigs walk with the window (blocks whose 40-step estimate lies far above are pruned, as the
    reference prunes by default); the blocks within reach of the lowest then give their eigenpairs
    (Krylov-Schur, k doubled until a level outside the window shows or k_max), and E0 is the lowest
    eigenvalue among them. partners {mirror block: block} (time-reversal pairs of a real H): a mirror
    block is not scanned; it takes its partner's lowest level (its eigenpairs, when within reach, are
    its own)."""
    from .eigs import dense_floor, walk_levels

    partners = {} if partners is None else partners
    chosen = [i for i in range(len(blocks)) if only is None or i in only]
    idx = [i for i in chosen if not (i in partners and partners[i] in chosen)]
    lv, _ = walk_levels([blocks[i] for i in idx], 1, floor=dense_floor(1, None), prune=prune, window=window,
                        norm_bound=norm_bound)
    lows = [math.inf] * len(blocks)
    for q, (v, _, _) in lv.items():
        lows[idx[q]] = float(v[0]) if len(v) else math.inf
    for i in chosen:
        if i not in idx:
            lows[i] = lows[partners[i]]
    e_scan = min(lows)
    slack = window + 1e-9 * max(1.0, abs(e_scan))
    found = []
    for i in range(len(blocks)):
        if lows[i] > e_scan + slack:
            continue
        blk, sec, m = blocks[i]
        k = 2
        while True:
            k = min(k, sec.dim)
            buf = _lib_alloc(k * sec.dim * 16)
            vals = np.zeros(k)
            n, conv = C.c_int(), C.c_int()
            _lib.call("qedh_block_eigenpairs", m.h, int(k), C.c_int64(0), _lib.ptr(vals, C.c_double), buf.ptr,
                      C.byref(n), C.byref(conv))
            got = vals[: n.value]
            if len(got) == 0:
                raise RuntimeError(f"dynamics: no eigenpair of block {blk.label}")
            if got.max() > got.min() + slack or k >= sec.dim or k >= k_max:
                break
            k *= 2
        found.append((i, buf, got))
    e0 = min(float(g.min()) for _, _, g in found)
    states = []
    for i, buf, got in found:
        for j, v in enumerate(got):
            if v <= e0 + window:
                states.append((i, buf, j, float(v)))
    return e0, states


class _lib_alloc:
    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _lib.call("qedh_malloc", C.byref(p), C.c_size_t(max(int(nbytes), 16)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) and _lib._lib is not None:
            _lib.lib().qedh_free(self.ptr)
            self.ptr = None


def block_seed(base: int, key: int) -> int:
    """splitmix64 of base + phi (key + 1): a source sector's seed (include/ed/thermal/sample_seed.h:48-53)."""
    m = (1 << 64) - 1
    z = (int(base) + 0x9E3779B97F4A7C15 * (int(key) + 1)) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _reach(blocks, b, O, delta):
    """(target index, probe operator on the device CSR) for every block the probe reaches from block b."""
    blk_b, sec_b, _ = blocks[b]
    out = []
    for t, (blk_t, sec_t, _) in enumerate(blocks):
        if sec_b.n_up >= 0 and sec_t.n_up != sec_b.n_up + delta:
            continue
        prog = projected_program(O, sec_t, sec_b)   # no part between the two: nothing to build
        if prog.counts()[2] == 0:
            continue
        tgt = sec_t if t != b else Sector(blk_t.group, blk_t.n_up)
        out.append((t, SectorMatvec(O, tgt, sec_b, lane=LANE_CSR, tune=False, program=prog)))
    return out


def _dynamics_finite_T(H, probes, omega, T, eta, blocks, krylov, samples, seed, sources):
    """FTLM over every source sector (dynamics.cpp:693-1257) in one batched run
    (qedh_ftlm_dynamics_walk): per source its samples' Z and per (source, probe, target) channel its
    S, then S_O(omega, T) = sum_src w_src S_src / sum_src w_src Z_src with
    w_src = e^{-beta (emin_src - ref)}."""
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    nT, nW, P = T.size, omega.size, len(probes)
    deltas = [_delta_up(O) for O in probes]
    s0 = int(seed) if seed else int.from_bytes(__import__("os").urandom(8), "little")
    tol = KBREAKDOWN_REL * (H.norm_bound() or 1.0)
    for O in probes:
        O._handle()
    src = list(range(len(blocks))) if sources is None else list(sources)
    t_probes = time.perf_counter()
    seeds, chan_src, chan_h, chan_a, chan_probe, keep = [], [], [], [], [], []
    # every source's probe operators, built on the host threads (results in source order)
    reaches = parallel_build(
        lambda b: [(p, t, X) for p, O in enumerate(probes) for t, X in _reach(blocks, b, O, deltas[p])], src)
    for i, b in enumerate(src):
        blk, sec, m = blocks[b]
        irr = blk.group.G if blk.momentum is not None else 1
        key = ((sec.n_up + 1) * 3 + 0) * irr + (blk.momentum or 0)
        seeds.append(block_seed(s0, key))
        for p, t, X in reaches[i]:
            chan_src.append(i)
            chan_h.append(blocks[t][2].h.value)
            chan_a.append(X.h.value)
            chan_probe.append(p)
            keep.append(X)
    nS, nC = len(src), len(chan_src)
    t_probes = time.perf_counter() - t_probes
    sh = (C.c_void_p * nS)(*[blocks[b][2].h.value for b in src])
    seed_a = np.array(seeds, np.uint64)
    tol_a = np.full(nS, tol)
    cs = np.array(chan_src or [0], np.int32)
    ch = (C.c_void_p * max(nC, 1))(*(chan_h or [None]))
    ca = (C.c_void_p * max(nC, 1))(*(chan_a or [None]))
    Z = np.zeros((nS, nT))
    S = np.zeros((max(nC, 1), nT, nW, 2))
    emin = np.zeros(nS)
    done = np.zeros(nS, np.int32)
    stats = np.zeros(6)
    _lib.call("qedh_ftlm_dynamics_walk", nS, sh, _lib.ptr(seed_a, C.c_uint64), _lib.ptr(tol_a, C.c_double), nC,
              cs.ctypes.data_as(C.POINTER(C.c_int)), ch, ca, int(samples), int(krylov), nT, _lib.ptr(T, C.c_double), nW,
              _lib.ptr(omega, C.c_double), C.c_double(eta), C.c_uint64(0), _lib.ptr(Z, C.c_double),
              _lib.ptr(S, C.c_double), _lib.ptr(emin, C.c_double), done.ctypes.data_as(C.POINTER(C.c_int)),
              _lib.ptr(stats, C.c_double))
    keep = None
    ref = float(np.min(emin))
    out = []
    for p in range(P):
        Sres = np.zeros((nT, nW), np.complex128)
        for it in range(nT):
            beta = 1.0 / T[it]
            w = np.exp(-beta * (emin - ref))
            Zp = float(np.sum(w * Z[:, it]))
            for c in range(nC):
                if chan_probe[c] == p:
                    Sres[it] += w[chan_src[c]] * (S[c, it, :, 0] + 1j * S[c, it, :, 1])
            Sres[it] /= Zp
        out.append(Sres)
    engine = dict(zip(["lanczos_s", "overlap_s", "host_s", "total_s", "waves", "setup_s"], stats.tolist()))
    engine.update(probes_s=t_probes, channels=nC)
    return out, ref, engine


def dynamics(H: Operator, probes, omega, *, eta: float = 0.1, T=None, sym: Optional[Symmetry] = None,
             krylov: int = 200, device: str = "gpu", degeneracy_tol: float = 1e-8, use_graph: bool = True,
             csr_budget: int = DEFAULT_CSR_BUDGET, blocks=None, ground_blocks=None, samples: int = 30,
             seed: int = 0, sources=None) -> DynamicsResult:
    """S_O(omega) = sum_n |<n|O|g>|^2 L_eta(omega - E_n + E0), averaged over the ground manifold,
    for every probe O (an Operator or a list of them). blocks: the walk's (block, sector, matvec)
    triples when the caller already built them (thermal.walk_blocks). ground_blocks: indices of the
    blocks known to hold the ground manifold (default: every block is searched).

    T (a list of temperatures > 0): finite-temperature FTLM dynamics instead, `samples` random
    vectors per source sector from the seed (0 draws one); S[p] is then (len(T), len(omega)),
    complex (the autocorrelation's imaginary part is roundoff). sources: the source blocks
    (default: every block of the walk)."""
    resolve_device(device, "dynamics")
    probes = [probes] if isinstance(probes, Operator) else list(probes)
    if not probes:
        raise InvalidRequest("dynamics: no probe")
    omega = np.ascontiguousarray(omega, np.float64)
    if omega.size == 0 or not eta > 0:
        raise InvalidRequest("dynamics: an omega grid and eta > 0 are required")
    sym = sym if sym is not None else Symmetry()
    t0 = time.perf_counter()
    finite_T = T is not None and len(np.atleast_1d(T)) > 0
    own = blocks is None
    if blocks is None:
        # every apply of a walk runs in the batched engine (the continued fractions at T = 0, every
        # Lanczos run at T > 0) or is one of the ground search's few, so no block's SpMV is tuned;
        # T = 0 builds a block's operator only when the ground search or a probe image needs it
        blocks = walk_blocks(H, sym, csr_budget, tune=False, lazy=not finite_T)
    if not blocks:
        raise InvalidRequest("dynamics: the symmetry selection holds no state")
    t_walk = time.perf_counter() - t0
    if T is not None and len(np.atleast_1d(T)) > 0:
        if np.any(~(np.atleast_1d(T) > 0)):
            raise InvalidRequest("dynamics: temperatures must be positive")
        S, ref, engine = _dynamics_finite_T(H, probes, omega, T, eta, blocks, krylov, samples, seed, sources)
        if own:
            release_later(blocks)
        return DynamicsResult(omega=omega, S=S, e0=ref, ground_manifold=0, target_sectors=0,
                              timings={"walk_s": t_walk, "total_s": time.perf_counter() - t0}, engine=engine)
    sH = H.norm_bound() or 1.0
    window = degeneracy_tol * sH
    t1 = time.perf_counter()
    partners = time_reversal_partners(H, [b for b, _, _ in blocks])
    e0, states = ground_manifold(blocks, window, only=None if ground_blocks is None else set(ground_blocks),
                                 norm_bound=sH, partners=partners)
    t_gm = time.perf_counter() - t1
    # probe images: (state, probe) -> every block it reaches
    t2 = time.perf_counter()
    deltas = [_delta_up(O) for O in probes]
    images = []          # (target block index, probe index, device buffer of phi)
    keep = []            # the cross operators and buffers stay alive until the batched run ends
    for (b, buf, j, _) in states:
        blk_b, sec_b, _ = blocks[b]
        g = _DevView(buf.ptr.value, j * sec_b.dim * 16)
        for p, O in enumerate(probes):
            for t, (blk_t, sec_t, _) in enumerate(blocks):
                if sec_b.n_up >= 0 and sec_t.n_up != sec_b.n_up + deltas[p]:
                    continue
                # The probe is not invariant under the group, so even within the source's own block it
                # needs the lambda-projected program of two sectors: the target then gets a handle of
                # its own (the same representatives, enumerated again).
                prog = projected_program(O, sec_t, sec_b)
                if prog.counts()[2] == 0:
                    continue
                tgt = sec_t if t != b else Sector(blk_t.group, blk_t.n_up)
                X = SectorMatvec(O, tgt, sec_b, lane=LANE_GATHER, program=prog)
                phi = _lib_alloc(sec_t.dim * 16)
                X.apply(g, phi, complex_vectors=True)
                images.append((t, p, phi))
                keep.append(X)
    t_probe = time.perf_counter() - t2
    S = [np.zeros(omega.size) for _ in probes]
    norms = np.zeros(len(probes))
    engine = {}
    targets = sorted({t for t, _, _ in images})
    if images:
        slot = {t: [] for t in targets}
        for t, p, phi in images:
            slot[t].append((p, phi))
        R = max(len(v) for v in slot.values())
        if R > 64:
            raise InvalidRequest("dynamics: more than 64 probe images reach one block")
        mats = [blocks[t][2] for t in targets]
        starts = (C.c_void_p * (len(targets) * R))()
        owner = []
        for ti, t in enumerate(targets):
            for s, (p, phi) in enumerate(slot[t]):
                starts[ti * R + s] = phi.ptr.value
                owner.append((ti * R + s, p))
        out = np.zeros((len(targets) * R, omega.size))
        nsq = np.zeros(len(targets) * R)
        stats = np.zeros(9)
        _lib.call("qedh_cf_batched", _handles(mats), len(targets), int(R), starts, int(max(krylov, 2)), 1,
                  C.c_double(e0), C.c_double(KBREAKDOWN_REL * sH), int(omega.size), _lib.ptr(omega, C.c_double),
                  C.c_double(eta), int(bool(use_graph)), _lib.ptr(out, C.c_double), _lib.ptr(nsq, C.c_double),
                  None, None, _lib.ptr(stats, C.c_double))
        for i, p in owner:
            S[p] += out[i]
            norms[p] += nsq[i]
        engine = dict(zip(["pack_s", "steps_s", "host_s", "total_s", "applies", "rows", "nnz", "schedule", "tune_s"],
                          stats.tolist()))
    for p in range(len(probes)):
        S[p] /= float(len(states))
        norms[p] /= float(len(states))
    n_states = len(states)
    states = images = keep = None
    if own:
        release_later(blocks)
    return DynamicsResult(
        omega=omega, S=S, e0=e0, ground_manifold=n_states, target_sectors=len(targets),
        timings={"walk_s": t_walk, "ground_s": t_gm, "probe_s": t_probe, "total_s": time.perf_counter() - t0},
        engine=engine, norms=norms)

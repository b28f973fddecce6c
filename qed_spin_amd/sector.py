"""Symmetry sectors resident in HBM and the operators on them (the reference's RepSectorData +
RepSectorMatVec / CrossSectorMatVec device lanes, behind include/qedh.h)."""

from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _lib
from .errors import InvalidRequest
from .operator import Operator, Program

LANE_AUTO, LANE_GATHER, LANE_CSR = 0, 1, 2
LANE_NO_TUNE = 0x100             # OR-ed into the lane: skip the SpMV autotuning
DEFAULT_RANK_BUDGET = 16 << 30   # rank directory (C(N, n_up) / 4 bytes) up to 16 GiB of HBM
AUTOTUNE_MIN_ROWS = 1 << 18      # CSR operators from this many rows time their SpMV variants when built
BUILD_THREADS = int(os.environ.get("QED_SPIN_AMD_BUILD_THREADS", "8"))   # host threads building sectors and operators


def _own_stream() -> None:
    _lib.call("qedh_thread_stream", 1)


def parallel_build(fn: Callable, items: Sequence, threads: Optional[int] = None) -> List:
    """fn(item) for every item on host threads, each with a HIP stream of its own (builds of small
    sectors and operators are launch- and synchronisation-bound, so several in flight overlap);
    results in item order. Every build call synchronises its stream before it returns."""
    items = list(items)
    threads = BUILD_THREADS if threads is None else threads
    if len(items) <= 1 or threads <= 1:
        return [fn(x) for x in items]
    with ThreadPoolExecutor(max_workers=min(threads, len(items)), initializer=_own_stream) as ex:
        return list(ex.map(fn, items))
DEFAULT_CSR_BUDGET = 64 << 30


class DeviceArray:
    """A device allocation holding a 1-D array (double or complex128)."""

    def __init__(self, n: int, dtype=np.complex128):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _lib.call("qedh_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if d.nbytes:
            _lib.call("qedh_memcpy_htod", d.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(d.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            _lib.call("qedh_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes))
        return out

    def zero(self) -> None:
        if self.nbytes:
            _lib.call("qedh_memset", self.ptr, 0, C.c_size_t(self.nbytes))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib._lib is not None:
            _lib.lib().qedh_free(self.ptr)
            self.ptr = None


@dataclass
class GroupSpec:
    """|G| site permutations (bit i of g s = bit perms[g][i] of s), XOR flips, characters."""

    perms: np.ndarray            # (G, N) int32
    chars: np.ndarray            # (G,) complex128
    flips: Optional[np.ndarray] = None   # (G,) uint64

    @property
    def G(self) -> int:
        return int(self.perms.shape[0])

    @property
    def N(self) -> int:
        return int(self.perms.shape[1])


class Sector:
    """One symmetry sector, enumerated on the device: ascending representatives, 1/norms and the
    reverse lookup (n_up < 0: the full 2^N space)."""

    def __init__(self, group: GroupSpec, n_up: int, sublattice: int = -1, relaxed: bool = True,
                 rank_budget: int = DEFAULT_RANK_BUDGET, reps=None, inv_norms=None, irrep=None):
        """reps / inv_norms given: adopt them (the reference's RepSectorData) instead of enumerating;
        each representative is checked to find itself through this group's lookup. irrep given: the
        sector of that irrep, D(g) for every element of the group as an array (G, d, d) (a unitary
        representation; the group's characters are then not used): its basis states are the
        partner-0 states of the representatives whose stabiliser admits them, `dim` counts them."""
        perms = np.ascontiguousarray(group.perms, dtype=np.int32)
        chars = np.ascontiguousarray(np.asarray(group.chars, np.complex128).view(np.float64))
        flips = None if group.flips is None else np.ascontiguousarray(group.flips, dtype=np.uint64)
        if perms.ndim != 2 or chars.size != 2 * perms.shape[0]:
            raise InvalidRequest("sector: perms must be (G, N) and chars (G,)")
        self.group = group
        self.n_up = int(n_up)
        self.irrep_dim = 1
        h = C.c_void_p()
        head = (int(perms.shape[1]), int(n_up), int(perms.shape[0]), perms.ctypes.data_as(C.POINTER(C.c_int)),
                None if flips is None else flips.ctypes.data_as(C.POINTER(C.c_uint64)),
                chars.ctypes.data_as(C.POINTER(C.c_double)), int(sublattice), int(bool(relaxed)))
        if irrep is not None:
            D = np.ascontiguousarray(np.asarray(irrep, np.complex128))
            if D.ndim != 3 or D.shape[0] != perms.shape[0] or D.shape[1] != D.shape[2]:
                raise InvalidRequest("sector: irrep must be (G, d, d)")
            if reps is not None:
                raise InvalidRequest("sector: an irrep sector is enumerated, not imported")
            Dv = np.ascontiguousarray(D.reshape(-1).view(np.float64))
            _lib.call("qedh_sector_create_irrep", head[0], head[1], head[2], head[3], head[4], int(D.shape[1]),
                      _lib.ptr(Dv, C.c_double), int(sublattice), int(bool(relaxed)), C.c_uint64(int(rank_budget)),
                      C.byref(h))
            self.irrep_dim = int(D.shape[1])
            self.irrep = D
        elif reps is None:
            _lib.call("qedh_sector_create", *head, C.c_uint64(int(rank_budget)), C.byref(h))
        else:
            r = np.ascontiguousarray(reps, np.uint64)
            w = np.ascontiguousarray(inv_norms, np.float64)
            if r.shape != w.shape or r.ndim != 1:
                raise InvalidRequest("sector: reps and inv_norms must be 1-D of one length")
            _lib.call("qedh_sector_import", *head, C.c_int64(r.size), _lib.ptr(r, C.c_uint64), _lib.ptr(w, C.c_double),
                      C.c_uint64(int(rank_budget)), C.byref(h))
        self.h = h
        d, o, b, s = C.c_int64(), C.c_int64(), C.c_int(), C.c_double()
        _lib.call("qedh_sector_info", self.h, C.byref(d), C.byref(o), C.byref(b), C.byref(s))
        self.dim = int(d.value)
        self.n_orbits = int(o.value)
        self.sublattice_block = int(b.value)
        self.build_seconds = float(s.value)
        di, nr, nc = C.c_int(), C.c_int64(), C.c_int64()
        _lib.call("qedh_sector_irrep_info", self.h, C.byref(di), C.byref(nr), C.byref(nc))
        self.n_reps = int(nr.value)
        self.n_classes = int(nc.value)

    @property
    def real(self) -> bool:
        if self.irrep_dim > 1:
            return False
        return bool(np.all(np.asarray(self.group.chars).imag == 0.0))

    def export(self):
        reps = np.zeros(self.n_reps, np.uint64)
        inv = np.zeros(self.n_reps, np.float64)
        _lib.call("qedh_sector_export", self.h, _lib.ptr(reps, C.c_uint64), _lib.ptr(inv, C.c_double))
        return reps, inv

    def export_irrep(self):
        """A d > 1 sector's fields: per representative its stabiliser class and first state, per
        class its rank and C (d x d; columns from the rank on zero)."""
        if self.irrep_dim <= 1:
            raise InvalidRequest("export_irrep: the sector's irrep is one-dimensional")
        d = self.irrep_dim
        rep_class = np.zeros(self.n_reps, np.uint16)
        offset = np.zeros(self.n_reps + 1, np.uint64)
        rank = np.zeros(self.n_classes, np.uint8)
        Cm = np.zeros((self.n_classes, d, d), np.complex128)
        _lib.call("qedh_sector_irrep_export", self.h, rep_class.ctypes.data_as(C.c_void_p), _lib.ptr(offset, C.c_uint64),
                  rank.ctypes.data_as(C.c_void_p), _lib.ptr(Cm.reshape(-1).view(np.float64), C.c_double))
        return rep_class, offset, rank, Cm

    def key_bits(self):
        kb = np.zeros(self.group.N, np.int32)
        L = C.c_int()
        _lib.call("qedh_sector_key_bits", self.h, kb.ctypes.data_as(C.POINTER(C.c_int)), C.byref(L))
        return (kb if L.value else None), int(L.value)

    def index_and_projection(self, states):
        st = np.ascontiguousarray(states, dtype=np.uint64)
        idx = np.zeros(st.size, np.int64)
        proj = np.zeros(2 * st.size, np.float64)
        _lib.call("qedh_sector_index_and_projection", self.h, _lib.ptr(st, C.c_uint64), C.c_int64(st.size),
                  _lib.ptr(idx, C.c_int64), _lib.ptr(proj, C.c_double))
        return idx, proj.view(np.complex128)

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().qedh_sector_destroy(self.h)
            self.h = None


def projected_program(op: Operator, rows: Sector, cols: Sector) -> Program:
    """The program of O from sector cols to sector rows (lambda-projected, rows as ket); it depends
    on the sectors' groups and characters only, so any handles of them serve. counts()[2] == 0:
    O has no part between the two."""
    return Program(op, adjoint=True, src=rows, tgt=cols)


def matrix_elements(ops, bra: "Sector", ket: "Sector", bras, kets, programs=None) -> np.ndarray:
    """out[p, o] = <bras[p]| O_o |kets[p]> for device vectors (complex DeviceArrays) of the bra and ket
    sectors: every observable and pair in one walk of the bra sector (qedh_matrix_elements; the
    reference's rep_matrix_elements). programs: the observables' programs from the ket sector to the
    bra sector when the caller has them (projected_program(O, bra, ket))."""
    if programs is None:
        programs = [projected_program(O, bra, ket) for O in ops]
    n_obs, n_pairs = len(programs), len(bras)
    if len(kets) != n_pairs:
        raise InvalidRequest("matrix_elements: one ket per bra")
    out = np.zeros((n_pairs, n_obs), np.complex128)
    if n_obs == 0 or n_pairs == 0:
        return out
    ph = (C.c_void_p * n_obs)(*[p.h.value for p in programs])
    bh = (C.c_void_p * n_pairs)(*[b.ptr.value for b in bras])
    kh = (C.c_void_p * n_pairs)(*[k.ptr.value for k in kets])
    _lib.call("qedh_matrix_elements", bra.h, ket.h, n_obs, ph, n_pairs, bh, kh,
              _lib.ptr(out.reshape(-1).view(np.float64), C.c_double))
    return out


class SectorMatvec:
    """O restricted to a sector (or from sector `cols` to sector `rows`, one group) on device vectors.

    The program is the walk of O^dagger (for two sectors, lambda-projected with rows as ket)."""

    def __init__(self, op: Operator, rows: Sector, cols: Optional[Sector] = None, lane: int = LANE_AUTO,
                 csr_budget: int = DEFAULT_CSR_BUDGET, tune: bool = True, program: Optional[Program] = None):
        """tune=False: a large CSR operator keeps the default SpMV shape instead of timing the
        variants (for operators applied only through a batched engine). program: the compiled
        program when the caller has it already (Program(op, adjoint=True) for one sector,
        projected_program(op, rows, cols) for two)."""
        self.op = op
        self.rows = rows
        self.cols = cols if cols is not None else rows
        blocks = getattr(rows, "irrep_dim", 1) > 1 or getattr(self.cols, "irrep_dim", 1) > 1
        if blocks and self.cols is not rows:
            # the rows of an irrep sector are blocks written for an operator that commutes with the
            # group, between two sectors of one irrep (total S+-, S^2; the reference's
            # CrossSectorMatVec, src/engine/internal.h:557-563): its own program, not a projection
            if (rows.irrep_dim != self.cols.irrep_dim or rows.irrep.shape != self.cols.irrep.shape
                    or not np.array_equal(rows.irrep, self.cols.irrep)):
                raise InvalidRequest("matvec: two sectors of different irreps of dimension > 1")
        if program is not None:
            self.program = program
        elif self.cols is rows or blocks:
            self.program = Program(op, adjoint=True)
        else:
            self.program = projected_program(op, rows, self.cols)
        h = C.c_void_p()
        _lib.call("qedh_matvec_create", rows.h, None if self.cols is rows else self.cols.h, self.program.h,
                  C.c_double(op.norm_bound()), int(lane) | (0 if tune else LANE_NO_TUNE), C.c_uint64(int(csr_budget)),
                  C.byref(h))
        self.h = h
        self.info = self._info()

    def _info(self) -> dict:
        i = _lib.MatvecInfo()
        _lib.call("qedh_matvec_info", self.h, C.byref(i))
        return {k: getattr(i, k) for k, _ in _lib.MatvecInfo._fields_}

    @property
    def dim(self) -> int:
        return int(self.info["rows"])

    @property
    def is_real(self) -> bool:
        return bool(self.info["is_real"])

    @property
    def lane(self) -> str:
        return "device-csr" if self.info["lane"] == LANE_CSR else "device-gather"

    def dtype(self, complex_vectors=False):
        return np.float64 if (self.is_real and not complex_vectors) else np.complex128

    def apply(self, x: DeviceArray, y: DeviceArray, complex_vectors: bool = False) -> None:
        _lib.call("qedh_matvec_apply", self.h, x.ptr, y.ptr, int(bool(complex_vectors)))

    def apply_multi(self, xs: Sequence[DeviceArray], ys: Sequence[DeviceArray], complex_vectors=False) -> None:
        k = len(xs)
        ins = (C.c_void_p * k)(*[x.ptr.value for x in xs])
        outs = (C.c_void_p * k)(*[y.ptr.value for y in ys])
        _lib.call("qedh_matvec_apply_multi", self.h, ins, outs, k, int(bool(complex_vectors)))

    def apply_block(self, nv: int, X: DeviceArray, Y: DeviceArray, complex_vectors=False) -> None:
        _lib.call("qedh_matvec_apply_block", self.h, int(nv), X.ptr, Y.ptr, int(bool(complex_vectors)))

    def set_lanes(self, lanes: int) -> None:
        _lib.call("qedh_matvec_set_lanes", self.h, int(lanes))
        self.info = self._info()

    def apply_host(self, x: np.ndarray, complex_vectors: bool = False) -> np.ndarray:
        """y = O x for a host vector (staged through HBM)."""
        dt = self.dtype(complex_vectors)
        X = DeviceArray.from_host(np.asarray(x, dt))
        Y = DeviceArray(self.dim, dt)
        self.apply(X, Y, complex_vectors)
        return Y.to_host()

    def export_csr(self):
        nnz = int(self.info["nnz"])
        row_ptr = np.zeros(self.dim + 1, np.uint64)
        col = np.zeros(nnz, np.uint32)
        val = np.zeros(2 * nnz, np.float64)
        _lib.call("qedh_matvec_export_csr", self.h, _lib.ptr(row_ptr, C.c_uint64), _lib.ptr(col, C.c_uint32),
                  _lib.ptr(val, C.c_double))
        return row_ptr, col, val.view(np.complex128)

    def lanczos(self, v0: np.ndarray, max_iter: int, reorth: int = 1, breakdown_tol: float = 0.0,
                complex_vectors: bool = False):
        dt = self.dtype(complex_vectors)
        v0 = np.ascontiguousarray(v0, dtype=dt)
        alpha = np.zeros(max_iter, np.float64)
        beta = np.zeros(max_iter + 1, np.float64)
        iters = C.c_int()
        t = np.zeros(4, np.float64)
        _lib.call("qedh_lanczos", self.h, v0.ctypes.data_as(C.c_void_p), int(bool(complex_vectors)), int(max_iter),
                  int(reorth), C.c_double(breakdown_tol), _lib.ptr(alpha, C.c_double), _lib.ptr(beta, C.c_double),
                  C.byref(iters), _lib.ptr(t, C.c_double))
        m = int(iters.value)
        return alpha[:m], beta[: m + 1], {"apply_s": t[0], "reorth_s": t[1], "vector_s": t[2], "total_s": t[3]}

    def lowest(self, k: int = 1, max_iter: int = 0):
        vals = np.zeros(max(1, k), np.float64)
        n, conv, applies = C.c_int(), C.c_int(), C.c_int64()
        _lib.call("qedh_block_lowest", self.h, int(k), C.c_int64(max_iter), _lib.ptr(vals, C.c_double),
                  C.byref(n), C.byref(conv), C.byref(applies))
        return vals[: n.value], bool(conv.value), int(applies.value)

    def estimate(self) -> float:
        """The pruning estimate theta_1 - |r_1| of the block's lowest level (40 Lanczos steps)."""
        th, res, ap = C.c_double(), C.c_double(), C.c_int64()
        _lib.call("qedh_block_estimate", self.h, C.byref(th), C.byref(res), C.byref(ap))
        return th.value - res.value

    def ftlm(self, samples: int, krylov: int, seed: int, betas, breakdown_tol: float = 0.0):
        betas = np.ascontiguousarray(betas, np.float64)
        nT = betas.size
        lnZ, E, V = np.zeros(nT), np.zeros(nT), np.zeros(nT)
        emin = C.c_double()
        _lib.call("qedh_ftlm", self.h, int(samples), int(krylov), C.c_uint64(int(seed)), int(nT),
                  _lib.ptr(betas, C.c_double), C.c_double(breakdown_tol), _lib.ptr(lnZ, C.c_double),
                  _lib.ptr(E, C.c_double), _lib.ptr(V, C.c_double), C.byref(emin))
        return lnZ, E, V, float(emin.value)

    def continued_fraction(self, phi: np.ndarray, krylov: int, omega, eta: float, energy_shift=None,
                           breakdown_tol: float = 0.0):
        phi = np.ascontiguousarray(phi, np.complex128)
        omega = np.ascontiguousarray(omega, np.float64)
        S = np.zeros(omega.size)
        sh, nsq, ch = C.c_double(), C.c_double(), C.c_double()
        _lib.call("qedh_cf", self.h, phi.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), int(krylov),
                  int(omega.size), _lib.ptr(omega, C.c_double), C.c_double(eta), int(energy_shift is not None),
                  C.c_double(0.0 if energy_shift is None else energy_shift), C.c_double(breakdown_tol),
                  _lib.ptr(S, C.c_double), C.byref(sh), C.byref(nsq), C.byref(ch))
        return S, {"energy_shift": sh.value, "norm_sq": nsq.value, "convergence_change": ch.value}

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().qedh_matvec_destroy(self.h)
            self.h = None

"""Spin-1/2 operators as term records, the reference's qed.Operator / HamiltonianBuilder surface for the
hot path (python/qed/_builder.py; records -> canonical terms as src/ops/invariance.cpp:43-60).

A record is a product of single-site operators with a coefficient; the engine expands it into
canonical masked terms (set bit = spin up). Two-body records act as op1(site1) op2(site2) with op2
first; three-body records (op1 s1, op2 s2, op3 s3) with op1 first.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib
from .errors import InvalidRequest

OP_SPLUS, OP_SMINUS, OP_SZ = 0, 1, 2
_CHAR = {OP_SPLUS: "+", OP_SMINUS: "-", OP_SZ: "z"}


@dataclass
class Operator:
    """num_sites spins; products in insertion order (ops string, sites, coefficient), the first factor
    of the string leftmost (the last acts first)."""

    num_sites: int
    products: List[Tuple[str, Tuple[int, ...], complex]] = field(default_factory=list)

    def __post_init__(self):
        if not 1 <= int(self.num_sites) <= 63:
            raise InvalidRequest(f"num_sites must be in [1, 63], got {self.num_sites}")

    def add_product(self, ops: str, sites: Sequence[int], coeff: complex) -> "Operator":
        if len(ops) != len(sites):
            raise InvalidRequest("ops and sites differ in length")
        for s in sites:
            if not 0 <= int(s) < self.num_sites:
                raise InvalidRequest(f"site {s} out of range")
        for o in ops:
            if o not in "+-zxyudI":
                raise InvalidRequest(f"unknown single-site operator {o!r}")
        self.products.append((str(ops), tuple(int(s) for s in sites), complex(coeff)))
        return self

    def add_one_body(self, op: int, site: int, coeff: complex) -> "Operator":
        return self.add_product(_CHAR[op], (site,), coeff)

    def add_two_body(self, op1: int, site1: int, op2: int, site2: int, coeff: complex) -> "Operator":
        return self.add_product(_CHAR[op1] + _CHAR[op2], (site1, site2), coeff)

    def add_three_body(self, op1, site1, op2, site2, op3, site3, coeff) -> "Operator":
        # the reference folds (op1 s1, op2 s2, op3 s3) as the product op3 op2 op1 (op1 first)
        return self.add_product(_CHAR[op3] + _CHAR[op2] + _CHAR[op1], (site3, site2, site1), coeff)

    def add_terms(self, terms) -> "Operator":
        """Terms (coeff, ((op, site), ...)), the rightmost acting first."""
        for c, ops in terms:
            self.add_product("".join(o for o, _ in ops), [s for _, s in ops], c)
        return self

    # -- engine handles ---------------------------------------------------------------------------
    def __getstate__(self):
        d = dict(self.__dict__)
        d.pop("_engine", None)   # the engine handle stays with this object (copies rebuild their own)
        return d

    def _handle(self) -> "EngineOperator":
        """The engine's copy of the terms, rebuilt only when the products changed since the last call."""
        key = (int(self.num_sites), tuple(self.products))
        cached = self.__dict__.get("_engine")
        if cached is None or cached[0] != key:
            cached = (key, EngineOperator(self))
            self.__dict__["_engine"] = cached
        return cached[1]

    def norm_bound(self) -> float:
        return self._handle().info()[1]

    def is_real(self) -> bool:
        return bool(self._handle().info()[2])


class EngineOperator:
    """The canonical terms of an Operator inside the engine (qedh_operator)."""

    def __init__(self, op: Operator):
        h = C.c_void_p()
        _lib.call("qedh_operator_create", int(op.num_sites), C.byref(h))
        self.h = h
        self.n = int(op.num_sites)
        for ops, sites, c in op.products:
            arr = (C.c_int * max(1, len(sites)))(*sites)
            _lib.call("qedh_operator_add_product", self.h, ops.encode(), arr, float(c.real), float(c.imag))

    def info(self):
        n = C.c_int64()
        b = C.c_double()
        r = C.c_int()
        _lib.call("qedh_operator_info", self.h, C.byref(n), C.byref(b), C.byref(r))
        return int(n.value), float(b.value), int(r.value)

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().qedh_operator_destroy(self.h)
            self.h = None


class Program:
    """A compiled flip-grouped program (of O^dagger by default: the row program of a sector walk)."""

    def __init__(self, op: Operator, adjoint: bool = True, src=None, tgt=None):
        eop = op._handle()
        h = C.c_void_p()
        if src is None:
            _lib.call("qedh_program_compile", eop.h, int(adjoint), C.byref(h))
        else:
            _lib.call("qedh_program_compile_projected", eop.h, int(adjoint), src.h, tgt.h, C.byref(h))
        self.h = h

    @classmethod
    def from_arrays(cls, arrays: dict) -> "Program":
        """A program from the reference's MaskedProgram arrays (group_flip, group_setbits,
        group_vbegin, vsub_val, vsub_tbegin, term_sign, term_coeff as re/im pairs)."""
        a = {
            "group_flip": np.ascontiguousarray(arrays["group_flip"], np.uint64),
            "group_setbits": np.ascontiguousarray(arrays["group_setbits"], np.int32),
            "group_vbegin": np.ascontiguousarray(arrays["group_vbegin"], np.uint32),
            "vsub_val": np.ascontiguousarray(arrays["vsub_val"], np.uint64),
            "vsub_tbegin": np.ascontiguousarray(arrays["vsub_tbegin"], np.uint32),
            "term_sign": np.ascontiguousarray(arrays["term_sign"], np.uint64),
            "term_coeff": np.ascontiguousarray(np.asarray(arrays["term_coeff"]).view(np.float64)
                                               if np.iscomplexobj(arrays["term_coeff"]) else arrays["term_coeff"],
                                               np.float64),
        }
        self = cls.__new__(cls)
        h = C.c_void_p()
        P = _lib.ptr
        _lib.call("qedh_program_import", C.c_int64(a["group_flip"].size), P(a["group_flip"], C.c_uint64),
                  P(a["group_setbits"], C.c_int32), P(a["group_vbegin"], C.c_uint32), C.c_int64(a["vsub_val"].size),
                  P(a["vsub_val"], C.c_uint64), P(a["vsub_tbegin"], C.c_uint32), C.c_int64(a["term_sign"].size),
                  P(a["term_sign"], C.c_uint64), P(a["term_coeff"], C.c_double), C.byref(h))
        self.h = h
        return self

    def counts(self):
        g, v, t = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.call("qedh_program_counts", self.h, C.byref(g), C.byref(v), C.byref(t))
        return int(g.value), int(v.value), int(t.value)

    def export(self) -> dict:
        g, v, t = self.counts()
        out = {
            "group_flip": np.zeros(g, np.uint64),
            "group_setbits": np.zeros(g, np.int32),
            "group_vbegin": np.zeros(g + 1, np.uint32),
            "vsub_val": np.zeros(v, np.uint64),
            "vsub_tbegin": np.zeros(v + 1, np.uint32),
            "term_sign": np.zeros(t, np.uint64),
            "term_coeff": np.zeros(2 * t, np.float64),
        }
        P = _lib.ptr
        _lib.call(
            "qedh_program_export",
            self.h,
            P(out["group_flip"], C.c_uint64),
            P(out["group_setbits"], C.c_int32),
            P(out["group_vbegin"], C.c_uint32),
            P(out["vsub_val"], C.c_uint64),
            P(out["vsub_tbegin"], C.c_uint32),
            P(out["term_sign"], C.c_uint64),
            P(out["term_coeff"], C.c_double),
        )
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib.lib().qedh_program_destroy(self.h)
            self.h = None


class HamiltonianBuilder:
    """Bonds and couplings of a spin-1/2 model, written as records (python/qed/_builder.py)."""

    def __init__(self, num_sites: int):
        self.op = Operator(num_sites)

    def heisenberg(self, bonds, J: float = 1.0, Jz=None) -> "HamiltonianBuilder":
        """J (S+S- + S-S+)/2 + Jz SzSz per bond (XXZ when Jz differs)."""
        Jz = J if Jz is None else Jz
        for i, j in bonds:
            self.op.add_two_body(OP_SPLUS, i, OP_SMINUS, j, 0.5 * J)
            self.op.add_two_body(OP_SMINUS, i, OP_SPLUS, j, 0.5 * J)
            self.op.add_two_body(OP_SZ, i, OP_SZ, j, Jz)
        return self

    def xxz(self, bonds, Jxy: float, Delta: float) -> "HamiltonianBuilder":
        return self.heisenberg(bonds, J=Jxy, Jz=Jxy * Delta)

    def chirality(self, triangles, chi: float) -> "HamiltonianBuilder":
        """chi S_i.(S_j x S_k) per triangle: the six Levi-Civita products of S^x, S^y, S^z."""
        for i, j, k in triangles:
            for a, b, c, s in (
                ("x", "y", "z", 1),
                ("y", "z", "x", 1),
                ("z", "x", "y", 1),
                ("x", "z", "y", -1),
                ("y", "x", "z", -1),
                ("z", "y", "x", -1),
            ):
                self.op.add_product(a + b + c, (i, j, k), s * chi)
        return self

    def to_operator(self) -> Operator:
        return self.op

"""CPU checks of the C-ABI library: it loads, exports every entry point include/qedh.h declares, and
its host-side program compiler reproduces the reference's programs bit for bit (no device needed)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fixture_names, fixture_products, load_fixture

from qed_spin_amd import _lib, errors
from qed_spin_amd.operator import Operator, Program

HEADER = os.path.join(ROOT, "include", "qedh.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qedh_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_errors_map_to_classes():
    from qed_spin_amd import errors
    with pytest.raises(errors.InvalidRequest):
        Operator(70)
    h = C.c_void_p()
    with pytest.raises(errors.InvalidRequest):
        _lib.call("qedh_operator_create", 99, C.byref(h))


@pytest.mark.parametrize("name", fixture_names())
def test_program_bit_exact(name):
    fx = load_fixture(name)
    op = Operator(int(fx["N"]))
    for ops, sites, c in fixture_products(fx):
        op.add_product(ops, sites, c)
    P = Program(op, adjoint=True).export()
    assert np.array_equal(P["group_flip"], fx["program_group_flip"])
    assert np.array_equal(P["group_setbits"], fx["program_group_setbits"])
    assert np.array_equal(P["group_vbegin"], fx["program_group_vbegin"])
    assert np.array_equal(P["vsub_val"], fx["program_vsub_val"])
    assert np.array_equal(P["vsub_tbegin"], fx["program_vsub_tbegin"])
    assert np.array_equal(P["term_sign"], fx["program_term_sign"])
    assert np.array_equal(P["term_coeff"].view(np.uint64), fx["program_term_coeff"].view(np.uint64))


def test_norm_bound_and_realness():
    op = Operator(4)
    op.add_two_body(0, 0, 1, 1, 0.5).add_two_body(1, 0, 0, 1, 0.5).add_two_body(2, 0, 2, 1, 1.0)
    assert op.is_real()
    # S.S on one bond: 0.5 S+S- + 0.5 S-S+ + 0.25 ZZ -> s_H = 1.25
    assert abs(op.norm_bound() - 1.25) < 1e-15
    op.add_product("xyz", (0, 1, 2), 0.3)
    assert not op.is_real()


@pytest.mark.parametrize("name", ["j1j2_n12_k3", "chiral_n10_k2"])
def test_program_import_round_trip(name):
    """A program handed over as the reference's MaskedProgram arrays comes back unchanged."""
    fx = load_fixture(name)
    arrays = {k[len("program_"):]: fx[k] for k in fx if k.startswith("program_")}
    P = Program.from_arrays(arrays).export()
    for k, v in arrays.items():
        assert np.array_equal(np.asarray(P[k]).view(np.uint8), np.ascontiguousarray(v).view(np.uint8)), k


def test_program_import_rejects_bad_offsets():
    fx = load_fixture("j1j2_n12_k3")
    arrays = {k[len("program_"):]: fx[k].copy() for k in fx if k.startswith("program_")}
    arrays["group_vbegin"][-1] += 1
    with pytest.raises(errors.InvalidRequest):
        Program.from_arrays(arrays)

import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURES = os.path.join(ROOT, "tests", "golden", "fixtures")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libqedspin.so")


def fixture_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(FIXTURES, "*.npz")))


def load_fixture(name):
    with np.load(os.path.join(FIXTURES, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def fixture_products(fx):
    """The (ops, sites, coeff) records of a fixture."""
    out = []
    for ops, sites, c in zip(fx["rec_ops"], fx["rec_sites"], fx["rec_coeff"]):
        ops = str(ops)
        out.append((ops, tuple(int(s) for s in sites[: len(ops)]), complex(c)))
    return out


def fixture_program(fx):
    return {
        "group_flip": fx["program_group_flip"],
        "group_setbits": fx["program_group_setbits"],
        "group_vbegin": fx["program_group_vbegin"],
        "vsub_val": fx["program_vsub_val"],
        "vsub_tbegin": fx["program_vsub_tbegin"],
        "term_sign": fx["program_term_sign"],
        "term_coeff": fx["program_term_coeff"],
    }

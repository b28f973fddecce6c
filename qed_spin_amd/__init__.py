"""qed_spin_amd: an MI355X-native backend for QED_Spin's symmetry-sector matvec and Krylov drivers.

The sector enumeration, the on-the-fly and CSR sector operators and the Krylov vector work run in
hand-written HIP kernels for gfx950 (libqedspin.so, include/qedh.h); the drivers (Lanczos,
Krylov-Schur, FTLM, continued fraction) are C++ host code over that library. This package is the
Python surface: operators as term records, sectors, and the verbs with the reference's
``device=`` keyword.
"""

from . import errors
from .operator import OP_SMINUS, OP_SPLUS, OP_SZ, HamiltonianBuilder, Operator, Program
from .sector import LANE_AUTO, LANE_CSR, LANE_GATHER, DeviceArray, GroupSpec, Sector, SectorMatvec, matrix_elements
from .symmetry import Symmetry

__all__ = [
    "errors",
    "Operator",
    "HamiltonianBuilder",
    "Program",
    "OP_SPLUS",
    "OP_SMINUS",
    "OP_SZ",
    "Sector",
    "SectorMatvec",
    "matrix_elements",
    "GroupSpec",
    "DeviceArray",
    "Symmetry",
    "LANE_AUTO",
    "LANE_GATHER",
    "LANE_CSR",
]

"""Host helpers of the Krylov drivers' outputs: the tridiagonal eigensolve (the reference keeps it on
the host too, include/ed/krylov/tridiag.h)."""

from __future__ import annotations

import numpy as np
from scipy.linalg import eigh_tridiagonal


def tridiag_eigvals(alpha, beta):
    """Eigenvalues (ascending) of the tridiagonal with diagonal alpha[0..m) and off-diagonal beta[1..m)."""
    a = np.asarray(alpha, np.float64)
    b = np.asarray(beta, np.float64)[1 : len(a)]
    if len(a) == 1:
        return a.copy()
    return eigh_tridiagonal(a, b, eigvals_only=True)


def tridiag_lowest(alpha, beta) -> float:
    return float(tridiag_eigvals(alpha, beta)[0])

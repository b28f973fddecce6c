"""Groups with irreps of dimension 3 and 4 for the tests of d > 1 sectors: the symmetric group S_n
permuting n sites of each of several layers at once, and its standard irrep (the permutation
matrices on the vectors of coordinate sum zero)."""

from itertools import permutations

import numpy as np

from qed_spin_amd.symmetry import GroupSpec


def layered_symmetric_group(n, layers):
    """S_n on layers x n sites (site l n + i): every element permutes the sites of each layer alike.
    Element p: bit l n + i of g s = bit l n + p[i] of s."""
    perms = []
    for p in permutations(range(n)):
        perms.append([l * n + p[i] for l in range(layers) for i in range(n)])
    perms = np.array(perms, np.int32)
    return perms, [list(p) for p in permutations(range(n))]


def standard_irrep(elements, n):
    """D(g) = B^T P_g B, (P_g)_{ij} = [p[i] = j], B an orthonormal basis of the sum-zero vectors: a
    real orthogonal representation of dimension n - 1 in the composition p_{g1 g2} = p2[p1]."""
    J = np.ones((n, n)) / n
    w, V = np.linalg.eigh(np.eye(n) - J)
    B = V[:, w > 0.5]
    D = []
    for p in elements:
        P = np.zeros((n, n))
        for i in range(n):
            P[i, p[i]] = 1.0
        D.append(B.T @ P @ B)
    return np.array(D, np.complex128)


def layered_heisenberg(qs, n, layers, rung=0.7):
    """All-to-all Heisenberg within each layer (J = 1) and rungs l n + i -- (l + 1) n + i (J = rung)."""
    b = qs.HamiltonianBuilder(n * layers)
    b.heisenberg([(l * n + i, l * n + j) for l in range(layers) for i in range(n) for j in range(i + 1, n)], 1.0)
    if layers > 1:
        b.heisenberg([(l * n + i, (l + 1) * n + i) for l in range(layers - 1) for i in range(n)], rung)
    return b.to_operator()

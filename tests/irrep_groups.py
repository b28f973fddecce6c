"""Groups with irreps of dimension 2 to 8 for the tests of d > 1 sectors, each permuting n sites
of each of several layers at once, and its standard irrep (the permutation matrices on the vectors
of coordinate sum zero, dimension n - 1), irreducible because every group here is 2-transitive on
its n points (the permutation character then has norm 2):
- the symmetric group S_n (n <= 5);
- the affine group x -> a x + b of the field of q elements (q = 5, 7, 8, 9; sharply 2-transitive,
  q (q - 1) elements);
- PSL(2, 5) acting on the projective line over F_5 (6 points, 60 elements)."""

from itertools import permutations

import numpy as np

from qed_spin_amd.symmetry import GroupSpec


def layered(elements, n, layers):
    """Site permutations of point permutations on layers x n sites (site l n + i): every element
    permutes the sites of each layer alike. Element p: bit l n + i of g s = bit l n + p[i] of s."""
    return np.array([[l * n + p[i] for l in range(layers) for i in range(n)] for p in elements], np.int32)


def layered_symmetric_group(n, layers):
    """S_n on layers x n sites."""
    els = [list(p) for p in permutations(range(n))]
    return layered(els, n, layers), els


def _field(q):
    """(add, mul) tables of the field of q elements, q in {5, 7, 8, 9}: integers mod a prime; F_8 as
    polynomials over F_2 mod x^3 + x + 1 (bit k = the coefficient of x^k); F_9 as u + v i, i^2 = -1
    over F_3 (element 3 u + v)."""
    if q in (5, 7):
        e = np.arange(q)
        return (e[:, None] + e[None, :]) % q, (e[:, None] * e[None, :]) % q
    if q == 8:
        add = np.arange(8)[:, None] ^ np.arange(8)[None, :]
        mul = np.zeros((8, 8), int)
        for a in range(8):
            for b in range(8):
                r = 0
                for k in range(3):
                    if (b >> k) & 1:
                        r ^= a << k
                for k in (4, 3):   # reduce by x^3 + x + 1 = 0b1011
                    if (r >> k) & 1:
                        r ^= 0b1011 << (k - 3)
                mul[a, b] = r
        return add, mul
    if q == 9:
        add = np.zeros((9, 9), int)
        mul = np.zeros((9, 9), int)
        for a in range(9):
            for b in range(9):
                (u, v), (x, y) = divmod(a, 3), divmod(b, 3)
                add[a, b] = 3 * ((u + x) % 3) + (v + y) % 3
                mul[a, b] = 3 * ((u * x - v * y) % 3) + (u * y + v * x) % 3
        return add, mul
    raise ValueError(q)


def affine_group(q, layers):
    """x -> a x + b over the field of q elements (a != 0) on layers x q sites; the identity first."""
    add, mul = _field(q)
    one = next(a for a in range(1, q) if all(mul[a, x] == x for x in range(q)))
    units = [one] + [a for a in range(1, q) if a != one]
    els = [[int(add[mul[a, x], b]) for x in range(q)] for a in units for b in range(q)]
    assert len({tuple(p) for p in els}) == q * (q - 1)
    return layered(els, q, layers), els


def psl2_5(layers):
    """PSL(2, 5) on the projective line over F_5 (points 0..4 and infinity = 5) on layers x 6 sites:
    the closure of x -> x + 1 and x -> -1/x; the identity first."""
    inv = {1: 1, 2: 3, 3: 2, 4: 4}
    t = [1, 2, 3, 4, 0, 5]
    s = [5] + [(-inv[x]) % 5 for x in range(1, 5)] + [0]
    ident = list(range(6))
    els, seen, frontier = [ident], {tuple(ident)}, [ident]
    while frontier:
        nxt = []
        for p in frontier:
            for g in (t, s):
                c = [g[p[i]] for i in range(6)]
                if tuple(c) not in seen:
                    seen.add(tuple(c))
                    els.append(c)
                    nxt.append(c)
        frontier = nxt
    assert len(els) == 60
    return layered(els, 6, layers), els


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

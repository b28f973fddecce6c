"""The reference's small audit models (tests/python/support/models.py:505-646: make_audit_models,
the square and triangular tori, the NLCE clusters),
restated as this backend's Operators, with the symmetry selections of its golden cases
(tests/python/golden/cases.py:60-110) mapped onto this backend's Symmetry.

Each model is a list of (ops, sites, coeff) products: "+-" on (i, j) is S^+_i S^-_j. The random
models draw their couplings from numpy's default_rng exactly as the reference does (one uniform per
pair for the density test, then the normals), so the same numpy streams give the same terms; the CPU
tests pin every restatement against the reference's recorded whole spectrum."""

from __future__ import annotations

import numpy as np

from qed_spin_amd.operator import Operator
from qed_spin_amd.symmetry import Symmetry


def heisenberg(bonds, J=1.0, Jz=None):
    Jz = J if Jz is None else Jz
    out = []
    for i, j in bonds:
        out += [("+-", (i, j), 0.5 * J), ("-+", (i, j), 0.5 * J), ("zz", (i, j), Jz)]
    return out


def ring(N):
    return [(i, (i + 1) % N) for i in range(N)]


def triangular_torus(Lx, Ly):
    """Nearest and next-nearest bonds (each once) and the counter-clockwise up and down triangles of
    the Lx x Ly triangular torus, site (x, y) -> x + Lx y (models.py:439-464)."""
    def idx(x, y):
        return (x % Lx) + Lx * (y % Ly)

    nn, nnn, tri = [], [], []
    for y in range(Ly):
        for x in range(Lx):
            i = idx(x, y)
            nn += [(i, idx(x + 1, y)), (i, idx(x, y + 1)), (i, idx(x - 1, y + 1))]
            nnn += [(i, idx(x + 1, y + 1)), (i, idx(x - 2, y + 1)), (i, idx(x + 1, y - 2))]
            tri.append((i, idx(x + 1, y), idx(x, y + 1)))
            tri.append((idx(x + 1, y), idx(x + 1, y + 1), idx(x, y + 1)))

    def once(bs):
        seen, out = set(), []
        for i, j in bs:
            key = (min(i, j), max(i, j))
            if i != j and key not in seen:
                seen.add(key)
                out.append((i, j))
        return out

    return once(nn), once(nnn), tri


def chiral(triangles, Jchi):
    """Jchi S_a . (S_b x S_c) = Jchi (i/2) sum over the cyclic (a, b, c) of S^z_a (S^+_b S^-_c - S^-_b S^+_c)."""
    out = []
    for i, j, k in triangles:
        for a, b, c in ((i, j, k), (j, k, i), (k, i, j)):
            out += [("z+-", (a, b, c), 0.5j * Jchi), ("z-+", (a, b, c), -0.5j * Jchi)]
    return out


def random_real(N, seed, density=1.0):
    r = np.random.default_rng(seed)
    out = []
    for i in range(N):
        for j in range(i + 1, N):
            if r.random() > density:
                continue
            J = float(r.normal())
            Jz = float(r.normal())
            out += [("+-", (i, j), 0.5 * J), ("-+", (i, j), 0.5 * J), ("zz", (i, j), Jz)]
    return out


def random_complex(N, seed):
    r = np.random.default_rng(seed)
    out = []
    for i in range(N):
        for j in range(i + 1, N):
            c = complex(r.normal(), r.normal()) * 0.5
            out += [("+-", (i, j), c), ("-+", (i, j), np.conj(c)), ("zz", (i, j), float(r.normal()))]
    return out


def _tfim(N=10):
    out = [("zz", (i, (i + 1) % N), -1.0) for i in range(N)]
    out += [p for i in range(N) for p in (("+", (i,), -0.35), ("-", (i,), -0.35))]
    return out


def _tri_chiral():
    nn, nnn, tri = triangular_torus(4, 3)
    return heisenberg(nn) + heisenberg(nnn, 0.2) + chiral(tri, 0.5)


def square_torus(Lx, Ly):
    """Nearest-neighbour bonds of the periodic Lx x Ly square lattice, each once."""
    out = set()
    for y in range(Ly):
        for x in range(Lx):
            i = x + Lx * y
            for j in (((x + 1) % Lx) + Lx * y, x + Lx * ((y + 1) % Ly)):
                if i != j:
                    out.add((min(i, j), max(i, j)))
    return sorted(out)


def square_j1j2(L, J2):
    """models.py:604-617: J1 on the nearest, J2 on the diagonal bonds of the L x L torus."""
    def idx(x, y):
        return (x % L) + L * (y % L)

    nn, nnn = set(), set()
    for y in range(L):
        for x in range(L):
            i = idx(x, y)
            for j in (idx(x + 1, y), idx(x, y + 1)):
                nn.add((min(i, j), max(i, j)))
            for j in (idx(x + 1, y + 1), idx(x - 1, y + 1)):
                nnn.add((min(i, j), max(i, j)))
    return heisenberg(sorted(nn)) + heisenberg(sorted(nnn), J2)


_HEX = [(0, i) for i in range(1, 7)] + [(i, i % 6 + 1) for i in range(1, 7)]
_TRI6 = [(0, 1), (1, 2), (0, 3), (1, 3), (1, 4), (2, 4), (3, 4), (3, 5), (4, 5)]
_STRIP = [(i, i + 1) for i in range(0, 9)] + [(i, i + 2) for i in range(0, 8)]


def _tri_chiral_3x3():
    nn, _, tri = triangular_torus(3, 3)
    return heisenberg(nn) + chiral(tri, 0.5)


# name: (N, products, u1, su2, translation-invariant ring)
MODELS = {
    "dimer": (2, lambda: heisenberg([(0, 1)]), True, True, True),
    "triangle3": (3, lambda: heisenberg(ring(3)), True, True, True),
    "chain4": (4, lambda: heisenberg(ring(4)), True, True, True),
    "chain5_odd": (5, lambda: heisenberg(ring(5)), True, True, True),
    "open_chain10": (10, lambda: heisenberg([(i, i + 1) for i in range(9)]), True, True, False),
    "xy_chain10": (10, lambda: heisenberg(ring(10), 1.0, 0.0), True, False, True),
    "staggered_field10": (10, lambda: heisenberg(ring(10)) + [("z", (i,), 0.3 * (-1) ** i) for i in range(10)],
                          True, False, False),
    "tfim10": (10, _tfim, False, False, True),
    "random_real8": (8, lambda: random_real(8, 1), True, False, False),
    "random_complex8": (8, lambda: random_complex(8, 2), True, False, False),
    "tri_chiral4x3": (12, _tri_chiral, True, True, False),
    "square4x3": (12, lambda: heisenberg(square_torus(4, 3)), True, True, False),
    "square4x4_j2=1": (16, lambda: square_j1j2(4, 1.0), True, True, False),
    "tri_chiral_3x3": (9, _tri_chiral_3x3, True, True, False),
    # the open clusters of the NLCE engine (models.py:630-646)
    "nlce_triangle3": (3, lambda: heisenberg([(0, 1), (1, 2), (2, 0)]), True, True, False),
    "nlce_rhombus4": (4, lambda: heisenberg([(0, 1), (1, 2), (2, 3), (3, 0), (1, 3)]), True, True, False),
    "nlce_bowtie5": (5, lambda: heisenberg([(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 2)]), True, True, False),
    "nlce_hexagon7": (7, lambda: heisenberg(_HEX), True, True, False),
    "nlce_hexagon7_xxz": (7, lambda: heisenberg(_HEX, 1.0, 0.75), True, False, False),
    "nlce_triangle6": (6, lambda: heisenberg(_TRI6), True, True, False),
    "nlce_strip10": (10, lambda: heisenberg(_STRIP), True, True, False),
}


def operator(name: str) -> Operator:
    N, make, *_ = MODELS[name]
    op = Operator(N)
    for ops, sites, c in make():
        op.add_product(ops, sites, c)
    return op


def symmetry(name: str, label: str):
    """This backend's Symmetry for a golden case label (and the total_spin request, or None)."""
    N, _, u1, su2, ring_ok = MODELS[name]
    if label == "none":
        return Symmetry.none(), None
    if label in ("spatial=none", "auto/point_group=off"):
        return (Symmetry(translations=False) if u1 else Symmetry.none()), None
    if label == "auto":
        if ring_ok:
            return Symmetry(sz="all" if u1 else "off"), None
        return (Symmetry(translations=False) if u1 else Symmetry.none()), None
    if label.startswith("sz="):
        return Symmetry(sz=[int(label[3:])], translations=False), None
    if label.startswith("auto/sz="):
        n = int(label[8:])
        return (Symmetry(sz=[n]) if ring_ok else Symmetry(sz=[n], translations=False)), None
    if label == "total_spin=min":
        return Symmetry(translations=False), "min"
    raise KeyError(label)


LABELS = ["none", "spatial=none", "auto", "sz", "auto/sz", "total_spin=min"]


def record_keys(name: str, records) -> list:
    """(verb, label, key) of every eigs / spectrum record of the model this backend can run."""
    out = []
    for k in records:
        if not k.startswith(name + "/api/"):
            continue
        rest = k[len(name) + 5:]
        if rest.startswith("spectrum/"):
            verb, label = "spectrum", rest[len("spectrum/"):]
        elif rest.startswith("eigs/") and rest.endswith("/k=4"):
            verb, label = "eigs", rest[len("eigs/"):-len("/k=4")]
        else:
            continue
        try:
            symmetry(name, label)
        except KeyError:
            continue
        out.append((verb, label, k))
    return out

"""Group data of the sectors the verbs solve: translations (and reflections) of a ring, their
characters, and the (Sz, k) walk.

Conventions follow the reference: an element acts as bit i of g s = bit perm[i] of s; the translation
generator is T[i] = (i + 1) mod N (tests/python/support/models.py:97); the abelian characters are
chi_m(T^a) = e^{2 pi i m a / N}, exact at quarter turns (src/basis/irreps.cpp:161-175 root_of_unity,
:181-242 abelian_characters with the group in the order T^0, T^1, ...).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .sector import GroupSpec


def root_of_unity(p: int, L: int) -> complex:
    """e^{2 pi i p / L}, exact at the quarter turns; cos/sin of (2 pi p) / L otherwise."""
    p %= L
    if (4 * p) % L == 0:
        return [1 + 0j, 1j, -1 + 0j, -1j][(4 * p) // L]
    th = 6.283185307179586476925286766559 * float(p) / float(L)
    return complex(math.cos(th), math.sin(th))


def translations(N: int) -> np.ndarray:
    """T^a, a = 0..N-1: T^a[i] = (i + a) mod N."""
    return np.array([[(i + a) % N for i in range(N)] for a in range(N)], dtype=np.int32)


def reflections(N: int) -> np.ndarray:
    """T^a R with R[i] = -i mod N (R first): i -> (a - i) mod N."""
    return np.array([[(a - i) % N for i in range(N)] for a in range(N)], dtype=np.int32)


def momentum_chars(N: int, m: int) -> np.ndarray:
    return np.array([root_of_unity(m * a, N) for a in range(N)], dtype=np.complex128)


def chain_group(N: int, m: int) -> GroupSpec:
    """The momentum-m sector of the N translations."""
    return GroupSpec(perms=translations(N), chars=momentum_chars(N, m))


def dihedral_group(N: int, m: int, parity: int) -> GroupSpec:
    """Translations and reflections at a reflection-invariant momentum (m = 0 or N/2): characters
    chi(T^a) = e^{2 pi i m a/N}, chi(T^a R) = parity chi(T^a)."""
    if (2 * m) % N != 0:
        raise ValueError("a reflection-parity sector needs m = 0 or N/2")
    ck = momentum_chars(N, m)
    perms = np.concatenate([translations(N), reflections(N)])
    chars = np.concatenate([ck, parity * ck])
    return GroupSpec(perms=perms, chars=chars)


def compose(perms: np.ndarray, flips=None) -> np.ndarray:
    """prod[g1, g2] = the index of the element g1 g2 (U(g1) U(g2) = U(g1 g2); with (g s)_i = s_{p[i]},
    g1 g2 has p[i] = p2[p1[i]] and flip f1 ^ p1(f2)), or -1 when the product is not in the list."""
    perms = np.asarray(perms)
    G, N = perms.shape
    fl = np.zeros(G, np.uint64) if flips is None else np.asarray(flips, np.uint64)

    def permute(p, mask):   # (g m)_i = m_{p[i]}
        out = 0
        for i in range(N):
            out |= ((int(mask) >> int(p[i])) & 1) << i
        return out

    index = {(tuple(perms[g]), int(fl[g])): g for g in range(G)}
    prod = np.full((G, G), -1, np.int64)
    for a in range(G):
        for b in range(G):
            p = perms[b][perms[a]]
            f = int(fl[a]) ^ permute(perms[a], fl[b])
            prod[a, b] = index.get((tuple(p), f), -1)
    return prod


def is_representation(perms: np.ndarray, D, flips=None, tol: float = 1e-12) -> bool:
    """Whether D (G, d, d) is a unitary representation of the group in the composition above."""
    D = np.asarray(D, np.complex128)
    prod = compose(perms, flips)
    if np.any(prod < 0):
        return False
    d = D.shape[1]
    for g in range(D.shape[0]):
        if np.abs(D[g] @ D[g].conj().T - np.eye(d)).max() > tol:
            return False
    return bool(np.abs(D[:, None] @ D[None, :] - D[prod]).max() <= tol)


def dihedral_irrep(N: int, m: int):
    """The two-dimensional irrep E_m (0 < m < N/2) of the dihedral group of the N-ring, for the
    elements of dihedral_group's order: translations T^a (i -> i + a), then the reflections
    i -> a - i, which are R T^a with R: i -> -i. On the momentum pair (m, -m): D(T^a) = diag(w^a,
    w^-a), w = e^{2 pi i m / N}, D(R T^a) = [[0, w^-a], [w^a, 0]]. Returns (the group with the traces
    as characters, D of shape (2N, 2, 2))."""
    if not 0 < 2 * m < N:
        raise ValueError("E_m needs 0 < m < N/2")
    D = np.zeros((2 * N, 2, 2), np.complex128)
    for a in range(N):
        w = root_of_unity(m * a, N)
        wc = root_of_unity(-m * a, N)
        D[a] = [[w, 0.0], [0.0, wc]]
        D[N + a] = [[0.0, wc], [w, 0.0]]
    perms = np.concatenate([translations(N), reflections(N)])
    return GroupSpec(perms=perms, chars=np.trace(D, axis1=1, axis2=2)), D


def with_spin_flip(g: GroupSpec, z: int) -> GroupSpec:
    """The group g times the global spin flip (every site flipped), the flipped elements' characters
    z chi(g): the sector of flip parity z (only for n_up = N / 2, where the flip keeps Sz)."""
    N = g.N
    full = np.uint64((1 << N) - 1)
    base = np.zeros(g.G, np.uint64) if g.flips is None else np.asarray(g.flips, np.uint64)
    return GroupSpec(perms=np.concatenate([g.perms, g.perms]), chars=np.concatenate([g.chars, z * g.chars]),
                     flips=np.concatenate([base, base ^ full]))


def trivial_group(N: int) -> GroupSpec:
    return GroupSpec(perms=np.arange(N, dtype=np.int32)[None, :], chars=np.ones(1, np.complex128))


@dataclass
class Block:
    """One symmetry block of the walk: its group, magnetisation and labels."""

    group: GroupSpec
    n_up: int
    momentum: Optional[int] = None
    parity: Optional[int] = None
    flip: Optional[int] = None
    weight: int = 1      # blocks it stands for (2: it and its time-reversal partner)

    @property
    def label(self) -> str:
        parts = [f"n_up={self.n_up}"]
        if self.momentum is not None:
            parts.append(f"k={self.momentum}")
        if self.parity is not None:
            parts.append(f"R={self.parity:+d}")
        if self.flip is not None:
            parts.append(f"Z={self.flip:+d}")
        return ",".join(parts)


@dataclass
class Symmetry:
    """A chain's symmetry request: translations (momentum sectors), Sz, reflection parity.

    sz: an n_up, a list of them, or "all" (every n_up) / "off" (the full space, no U(1));
    momenta: "all" or a list of m; reflection: None, +1 or -1 (only with momenta 0 or N/2).
    translations=False: no spatial symmetry (one block per n_up)."""

    sz: object = "all"
    momenta: object = "all"
    translations: bool = True
    reflection: Optional[int] = None
    spin_flip: Optional[object] = None   # None, +1, -1 or "all" (both parities); applies at n_up = N / 2

    @classmethod
    def none(cls) -> "Symmetry":
        return cls(sz="off", momenta=None, translations=False)

    def blocks(self, N: int) -> List[Block]:
        if self.sz == "off":
            nups = [-1]
        elif self.sz == "all":
            nups = list(range(N + 1))
        elif isinstance(self.sz, (list, tuple)):
            nups = [int(x) for x in self.sz]
        else:
            nups = [int(self.sz)]
        if self.spin_flip not in (None, 1, -1, "all"):
            raise ValueError(f"spin_flip must be None, +1, -1 or 'all', got {self.spin_flip!r}")
        out = []
        for n_up in nups:
            if not self.translations:
                blocks = [Block(trivial_group(N), n_up)]
            else:
                blocks = []
                ms = range(N) if self.momenta in ("all", None) else [int(m) % N for m in self.momenta]
                for m in ms:
                    if self.reflection is not None:
                        blocks.append(Block(dihedral_group(N, m, int(self.reflection)), n_up, m, int(self.reflection)))
                    else:
                        blocks.append(Block(chain_group(N, m), n_up, m))
            if self.spin_flip is not None and 2 * n_up == N:
                zs = (1, -1) if self.spin_flip == "all" else (int(self.spin_flip),)
                blocks = [Block(with_spin_flip(b.group, z), b.n_up, b.momentum, b.parity, z) for b in blocks for z in zs]
            out.extend(blocks)
        return out

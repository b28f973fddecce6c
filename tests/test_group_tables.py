"""The sublattice key-order tables built on the host (qedh_sublattice_tables, no device): the key
bits against the oracle's restatement of the reference's block-system rule
(include/ed/basis/sublattice_code.h:252-300), and the private accelerator tables (the least leading
block reachable from each block pattern and the elements reaching it) against brute force over
every element applied to every pattern."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from qed_spin_amd import _lib
from qed_spin_amd import symmetry as sym


def _tables(perms, flips, N, mode, relaxed=True):
    perms = np.ascontiguousarray(perms, np.int32)
    G = perms.shape[0]
    flips = np.ascontiguousarray(np.zeros(G, np.uint64) if flips is None else flips, np.uint64)
    L = C.c_int()
    n_cand = C.c_int64()
    kb = np.zeros(N, np.int32)
    args = (N, G, _lib.ptr(perms, C.c_int), _lib.ptr(flips, C.c_uint64), mode, int(relaxed), C.byref(L))
    _lib.call("qedh_sublattice_tables", *args, _lib.ptr(kb, C.c_int), None, None, None, 0, C.byref(n_cand))
    if L.value == 0:
        return None
    M = (N // L.value) << L.value
    lead = np.zeros(M, np.uint16)
    off = np.zeros(M + 1, np.uint32)
    cand = np.zeros(n_cand.value, np.uint16)
    _lib.call("qedh_sublattice_tables", *args, _lib.ptr(kb, C.c_int), lead.ctypes.data_as(C.c_void_p),
              _lib.ptr(off, C.c_uint32), cand.ctypes.data_as(C.c_void_p), n_cand.value, C.byref(n_cand))
    return L.value, kb, lead, off, cand


def _square_group(Lx, Ly):
    """Translations of an Lx x Ly torus (site x + Lx y) and the 180-degree rotation."""
    N = Lx * Ly
    out = []
    for rot in (False, True):
        for a in range(Lx):
            for b in range(Ly):
                p = []
                for i in range(N):
                    x, y = i % Lx, i // Lx
                    if rot:
                        x, y = -x, -y
                    p.append((x + a) % Lx + Lx * ((y + b) % Ly))
                out.append(p)
    return np.array(out, np.int32)


def _brute(perms, flips, N, L, kb):
    """lead / candidate lists from the definition: element g maps the sites of one block (its feed)
    onto block 0; the leading pattern of g s has bit t = bit perm[site_0t] of s XOR g's flip there."""
    m = N // L
    members = [sorted(range(N), key=lambda s: kb[s]) [N - (j + 1) * L: N - j * L] for j in range(m)]
    block = np.zeros(N, np.int64)
    for j, ms in enumerate(members):
        block[ms] = j
    P = 1 << L
    xs = np.arange(P, dtype=np.int64)
    lead = np.full(m * P, 0xFFFF, np.int64)
    imgs = []
    for g, p in enumerate(perms):
        feed = int(block[p[members[0][0]]])
        y = np.zeros(P, np.int64)
        for t in range(L):
            src = p[members[0][t]]
            u = members[feed].index(src)
            y |= ((xs >> u) & 1) << t
            if (int(flips[g]) >> members[0][t]) & 1:
                y ^= 1 << t
        imgs.append((feed, y))
        sl = slice(feed * P, (feed + 1) * P)
        lead[sl] = np.minimum(lead[sl], y)
    cands = [[] for _ in range(m * P)]
    for g, (feed, y) in enumerate(imgs):
        hit = np.nonzero(y == lead[feed * P:(feed + 1) * P])[0]
        for x in hit:
            cands[feed * P + int(x)].append(g)
    return lead, cands


CASES = [
    ("dihedral16", lambda: (sym.dihedral_group(16, 0, 1).perms, None, 16, 1)),
    ("dihedral16_flip", lambda: (sym.with_spin_flip(sym.dihedral_group(16, 0, 1), 1).perms,
                                  sym.with_spin_flip(sym.dihedral_group(16, 0, 1), 1).flips, 16, 1)),
    ("ring24_size_rule", lambda: (sym.translations(24), None, 24, -1)),
    ("dihedral24", lambda: (sym.dihedral_group(24, 0, 1).perms, None, 24, -1)),
    ("torus4x4", lambda: (_square_group(4, 4), None, 16, 1)),
    ("torus6x4_flip", lambda: (np.concatenate([_square_group(6, 4)] * 2),
                               np.concatenate([np.zeros(48, np.uint64), np.full(48, (1 << 24) - 1, np.uint64)]), 24, 1)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_key_order_tables_against_brute_force(name, make):
    perms, flips, N, mode = make()
    perms = np.asarray(perms, np.int32)
    flips = np.zeros(len(perms), np.uint64) if flips is None else np.asarray(flips, np.uint64)
    t = _tables(perms, flips, N, mode)
    kb_ref, L_ref = po.sublattice_key_bits(perms, flips, N, mode)
    if t is None:
        assert L_ref == 0
        return
    L, kb, lead, off, cand = t
    assert L == L_ref and np.array_equal(kb, kb_ref)
    want_lead, want_cand = _brute(perms, flips, N, L, kb)
    assert np.array_equal(lead.astype(np.int64), want_lead)
    got = [list(cand[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
    assert got == want_cand


def test_no_block_system_leaves_the_tables_off():
    # a prime ring has no nontrivial block system; the size rule also refuses small groups
    assert _tables(sym.translations(13), None, 13, 1) is None
    assert _tables(sym.translations(16), None, 16, -1) is None

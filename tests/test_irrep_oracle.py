"""The oracle of sectors of irreps of dimension d > 1 (oracle/pyoracle.py irrep_fields and
irrep_basis_matrix), pinned by what the mathematics fixes, on the CPU:
- the E_m matrices of the dihedral group are a unitary representation in this package's composition
  of elements (and the other order of composition is not);
- the partner-0 states of E_m are orthonormal, and H restricted to them has the spectrum of the
  momentum-m block (the E_m copies and the momentum-m states are in one-to-one correspondence);
- the irreps tile the Sz sector: sum over irreps of d x dim = C(N, n_up)."""

from math import comb

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import symmetry as sym


def ring_h(N, J2=0.5):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    b.heisenberg([(i, (i + 2) % N) for i in range(N)], J2)
    return b.to_operator()


def sz_hamiltonian(H, N, n_up):
    triv = po.OracleSector(N, n_up, np.arange(N, dtype=np.int32)[None, :], np.ones(1))
    rp, col, val = po.sector_csr(qs.Program(H).export(), triv)
    A = np.zeros((triv.dim, triv.dim), complex)
    for r in range(triv.dim):
        A[r, col[rp[r]:rp[r + 1]]] += val[rp[r]:rp[r + 1]]
    return A, triv.reps


@pytest.mark.parametrize("N", [6, 8, 10])
def test_dihedral_irreps_are_representations(N):
    for m in range(1, (N + 1) // 2):
        if 2 * m == N:
            continue
        g, D = sym.dihedral_irrep(N, m)
        assert sym.is_representation(g.perms, D), m
        # composing the other way round (T^a R instead of R T^a for the reflections) is not one
        D2 = np.concatenate([D[:N], np.array([D[a] @ D[N] for a in range(N)])])
        assert not sym.is_representation(g.perms, D2), m


@pytest.mark.parametrize("N,n_up", [(8, 4), (8, 3), (10, 5), (10, 4)])
def test_irrep_basis_orthonormal_with_momentum_spectrum_and_tiling(N, n_up):
    H = ring_h(N)
    Hs, states = sz_hamiltonian(H, N, n_up)
    total = 0
    for m in range(N):
        g = sym.chain_group(N, m)
        o = po.OracleSector(N, n_up, g.perms, g.chars)
        Ek = po.sector_basis_matrix(o, states)
        ek = np.linalg.eigvalsh(Ek.conj().T @ Hs @ Ek)
        if not 0 < 2 * m < N:
            continue
        gD, D = sym.dihedral_irrep(N, m)
        reps, Cs = po.irrep_fields(N, n_up, gD.perms, D)
        E = po.irrep_basis_matrix(N, gD.perms, D, reps, Cs, states)
        assert E.shape[1] == o.dim, m
        assert np.abs(E.conj().T @ E - np.eye(E.shape[1])).max() <= 1e-12, m
        ev = np.linalg.eigvalsh(E.conj().T @ Hs @ E)
        assert np.abs(np.sort(ev) - np.sort(ek)).max() <= 1e-12 * max(1.0, np.abs(ek).max()), m
        total += 2 * E.shape[1]
    # the one-dimensional irreps of D_N: momenta 0 (and N/2) with both reflection parities
    for m in ([0, N // 2] if N % 2 == 0 else [0]):
        for par in (1, -1):
            g = sym.dihedral_group(N, m, par)
            total += po.OracleSector(N, n_up, g.perms, g.chars).dim
    assert total == comb(N, n_up)


def test_rank_follows_the_stabiliser():
    # N = 4, n_up = 2, E_1: |0011> is fixed by the identity and one reflection, whose D has
    # eigenvalues +1 and -1: rank 1. |0101> is fixed by T^2, and D(T^2) = -I: rank 0 (the Neel
    # orbit carries momenta 0 and 2 only), so it holds no state.
    g, D = sym.dihedral_irrep(4, 1)
    reps, Cs = po.irrep_fields(4, 2, g.perms, D)
    ranks = {int(r): C.shape[1] for r, C in zip(reps, Cs)}
    assert ranks == {0b0011: 1}
    assert sum(ranks.values()) == po.OracleSector(4, 2, sym.translations(4), sym.momentum_chars(4, 1)).dim


def isotypic_spectrum(Hs, states, N, perms, D):
    """Eigenvalues of H on the isotypic subspace of the irrep D in the Sz basis `states` (the range
    of P = (d / |G|) sum_g conj(chi(g)) U(g)); every level there comes d times."""
    index = {int(s): i for i, s in enumerate(states)}
    G, d = D.shape[0], D.shape[1]
    chi = np.trace(D, axis1=1, axis2=2)
    P = np.zeros((len(states), len(states)), complex)
    for g in range(G):
        for j, s in enumerate(states):
            t = 0
            for i in range(N):
                t |= ((int(s) >> int(perms[g][i])) & 1) << i
            P[index[t], j] += np.conj(chi[g])
    P *= d / G
    w, V = np.linalg.eigh(0.5 * (P + P.conj().T))
    Q = V[:, w > 0.5]
    return np.linalg.eigvalsh(Q.conj().T @ Hs @ Q), Q.shape[1]


@pytest.mark.parametrize("n,layers,n_up", [(3, 2, 3), (4, 2, 4), (4, 2, 3), (5, 2, 5), (5, 2, 4)])
def test_standard_irreps_of_symmetric_groups(n, layers, n_up):
    from irrep_groups import layered_heisenberg, layered_symmetric_group, standard_irrep

    N = n * layers
    perms, els = layered_symmetric_group(n, layers)
    D = standard_irrep(els, n)
    d = D.shape[1]
    assert sym.is_representation(perms, D)
    H = layered_heisenberg(qs, n, layers)
    Hs, states = sz_hamiltonian(H, N, n_up)
    reps, Cs = po.irrep_fields(N, n_up, perms, D)
    E = po.irrep_basis_matrix(N, perms, D, reps, Cs, states)
    assert np.abs(E.conj().T @ E - np.eye(E.shape[1])).max() <= 1e-12
    iso, rank = isotypic_spectrum(Hs, states, N, perms, D)
    assert rank == d * E.shape[1]
    ev = np.linalg.eigvalsh(E.conj().T @ Hs @ E)
    assert np.abs(np.sort(np.repeat(ev, d)) - np.sort(iso)).max() <= 1e-11


TWO_TRANSITIVE = {"agl5": (5, 4), "psl2_5": (6, 5), "agl7": (7, 6), "agl8": (8, 7), "agl9": (9, 8)}


def two_transitive(name, layers):
    """(site permutations, point permutations, n) of the 2-transitive groups of irrep_groups."""
    from irrep_groups import affine_group, psl2_5

    n = TWO_TRANSITIVE[name][0]
    perms, els = psl2_5(layers) if name == "psl2_5" else affine_group(n, layers)
    return perms, els, n


@pytest.mark.parametrize("name,layers,n_up", [("agl5", 2, 4), ("psl2_5", 2, 5), ("agl7", 2, 4), ("agl8", 2, 3),
                                              ("agl9", 2, 3)])
def test_standard_irreps_of_two_transitive_groups_d4_to_d8(name, layers, n_up):
    """d = n - 1 = 4..8: D is a unitary representation (and the reversed composition is not), its
    character has norm 1 (irreducible), and the oracle's sector has an orthonormal basis on which H
    has the isotypic spectrum (each level d times)."""
    from irrep_groups import layered_heisenberg, standard_irrep

    perms, els, n = two_transitive(name, layers)
    N = n * layers
    D = standard_irrep(els, n)
    d = D.shape[1]
    assert d == TWO_TRANSITIVE[name][1]
    G = len(els)
    assert sym.is_representation(perms, D)
    # g -> D(g)^dag reverses products: a representation only of an abelian group
    assert not sym.is_representation(perms, np.conj(np.transpose(D, (0, 2, 1))))
    chi = np.trace(D, axis1=1, axis2=2)
    assert abs(np.vdot(chi, chi).real / G - 1.0) <= 1e-12
    H = layered_heisenberg(qs, n, layers)
    Hs, states = sz_hamiltonian(H, N, n_up)
    reps, Cs = po.irrep_fields(N, n_up, perms, D)
    E = po.irrep_basis_matrix(N, perms, D, reps, Cs, states)
    assert E.shape[1] > 0
    assert np.abs(E.conj().T @ E - np.eye(E.shape[1])).max() <= 1e-12
    iso, rank = isotypic_spectrum(Hs, states, N, perms, D)
    assert rank == d * E.shape[1]
    ev = np.linalg.eigvalsh(E.conj().T @ Hs @ E)
    assert np.abs(np.sort(np.repeat(ev, d)) - np.sort(iso)).max() <= 1e-11

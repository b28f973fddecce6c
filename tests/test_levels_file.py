"""The EigResult .npz file and the host pieces of the level-vector verbs, without a device: the
array layout written and read back, every refusal of a damaged file, the sublattice fingerprint
(64-bit FNV-1a, pinned by its published test vectors) and the split of operators by their change
of magnetisation (pinned by the textbook S^x, S^y matrices)."""

import numpy as np
import pytest

from oracle import pyoracle as po

from qed_spin_amd import eigs as ev
from qed_spin_amd import errors
from qed_spin_amd import levels as lv
from qed_spin_amd import symmetry as sym
from qed_spin_amd.operator import Operator


def test_fingerprint_is_fnv1a_64():
    """FNV-1a of the bytes b: feeding key_bit = b - 1 hashes the byte b. Published vectors:
    FNV-1a-64("a") = 0xaf63dc4c8601ec8c, ("foobar") = 0x85944171f73967e8, ("") = the offset basis."""
    assert lv.sublattice_fingerprint([]) == 0xCBF29CE484222325
    assert lv.sublattice_fingerprint([ord("a") - 1]) == 0xAF63DC4C8601EC8C
    assert lv.sublattice_fingerprint([c - 1 for c in b"foobar"]) == 0x85944171F73967E8


def test_split_by_shift_against_textbook_matrices():
    """One site (state 0 = down, 1 = up): S^x = [[0, 1/2], [1/2, 0]], S^y = [[0, i/2], [-i/2, 0]] in
    the (down, up) order; the parts raise (d = +1) and lower (d = -1) by one spin."""
    x = lv.split_by_shift(Operator(1).add_product("x", (0,), 1.0))
    y = lv.split_by_shift(Operator(1).add_product("y", (0,), 1.0))
    assert sorted(x) == [-1, 1] and sorted(y) == [-1, 1]
    Mx = {d: po.dense_matrix(P.products, 1) for d, P in x.items()}
    My = {d: po.dense_matrix(P.products, 1) for d, P in y.items()}
    assert np.allclose(Mx[1], [[0, 0], [0.5, 0]]) and np.allclose(Mx[-1], [[0, 0.5], [0, 0]])
    assert np.allclose(My[1], [[0, 0], [-0.5j, 0]]) and np.allclose(My[-1], [[0, 0.5j], [0, 0]])


def test_split_by_shift_parts_change_n_up_by_their_shift():
    N = 5
    O = (Operator(N).add_product("xz", (0, 1), 0.7).add_product("y+", (2, 3), 1.3j)
         .add_product("+-", (4, 0), -0.4).add_product("xy", (1, 2), 0.25))
    parts = lv.split_by_shift(O)
    assert sorted(parts) == [-2, -1, 0, 1, 2]
    total = np.zeros((1 << N, 1 << N), complex)
    pc = np.array([bin(s).count("1") for s in range(1 << N)])
    for d, P in parts.items():
        M = po.dense_matrix(P.products, N)
        t, s = np.nonzero(np.abs(M) > 0)
        assert np.all(pc[t] - pc[s] == d)
        total += M
    assert np.abs(total - po.dense_matrix(O.products, N)).max() < 1e-15


def _result(N=6, n_up=3):
    """A result built by hand: two levels of the k = 0 and k = 1 sectors with made-up vectors."""
    bases, vecs, levels = {}, [], []
    rng = np.random.default_rng(3)
    for i, m in enumerate((0, 1)):
        g = sym.chain_group(N, m)
        os_ = po.OracleSector(N, n_up, g.perms, g.chars)
        bases[i] = lv.LevelBasis(perms=g.perms.astype(np.int32), chars=g.chars, flips=np.zeros(N, np.uint64),
                                 n_up=n_up, reps=os_.reps, inv_norms=os_.inv_norms)
        v = rng.normal(size=os_.dim) + 1j * rng.normal(size=os_.dim)
        vecs.append((i, v / np.linalg.norm(v)))
        levels.append({"energy": -2.0 + i, "n_up": n_up, "momentum": m, "parity": None, "dim": os_.dim,
                       "label": f"n_up={n_up},k={m}", "vector": i})
    return ev.EigsResult(energies=np.array([-2.0, -1.0]), blocks=[{"dim": b.dim} for b in bases.values()],
                         k=2, n_sites=N, levels=levels, _bases=bases, _vectors=vecs)


def test_round_trip_of_the_arrays(tmp_path):
    r = _result()
    p = tmp_path / "r.npz"
    r.save(p)
    with np.load(p) as f:
        keys = set(f.files)
        assert int(f["format_version"]) == 2 and int(f["k"]) == 2
        tag = f["level_tag"].reshape(2, 11)
        assert list(tag[:, 0]) == [3, 3] and list(tag[:, 3]) == [0, 1] and list(tag[:, 9]) == [r.levels[0]["dim"],
                                                                                            r.levels[1]["dim"]]
        assert list(f["basis1_shape"]) == [6, 6, 3, 1]
        assert f["vector_amplitudes"].dtype == np.complex128
    for k in ("level_energy", "level_mirror", "level_fold", "level_multiplicity", "level_vector", "level_momentum",
              "level_momentum_offset", "level_irrep_elems", "level_irrep_chars", "level_irrep_offset", "n_bases",
              "vector_basis", "vector_offset", "spec_scalars", "result_scalars", "basis0_reps", "basis0_inv_norms",
              "basis0_characters", "basis0_perms", "basis0_flip_masks", "basis0_sublattice"):
        assert k in keys, k
    s = ev.load_eigs(p)
    assert np.array_equal(s.energies, r.energies) and s.k == 2 and s.n_sites == 6
    assert [L["momentum"] for L in s.levels] == [0, 1]
    assert [L["vector"] for L in s.levels] == [0, 1]
    for (ka, va), (kb, vb) in zip(r._vectors, s._vectors):
        assert ka == kb and np.array_equal(va, vb)
    for b in r._bases:
        assert np.array_equal(r._bases[b].reps, s._bases[b].reps)
        assert np.array_equal(r._bases[b].inv_norms, s._bases[b].inv_norms)
        assert np.array_equal(r._bases[b].perms, s._bases[b].perms)
        assert np.array_equal(r._bases[b].chars, s._bases[b].chars)


def _damaged(tmp_path, edit):
    r = _result()
    p = tmp_path / "good.npz"
    r.save(p)
    with np.load(p) as f:
        d = {k: f[k] for k in f.files}
    edit(d)
    q = tmp_path / "bad.npz"
    np.savez(q, **d)
    return q


@pytest.mark.parametrize("what,edit", [
    ("format 1", lambda d: d.__setitem__("format_version", np.int64(1))),
    ("no version", lambda d: d.pop("format_version")),
    ("level_tag shape", lambda d: d.__setitem__("level_tag", d["level_tag"][:-1])),
    ("level lengths", lambda d: d.__setitem__("level_vector", d["level_vector"][:1])),
    ("vector offsets", lambda d: d.__setitem__("vector_offset", d["vector_offset"][::-1].copy())),
    ("vector basis", lambda d: d.__setitem__("vector_basis", np.array([0, 5]))),
    ("level vector", lambda d: d.__setitem__("level_vector", np.array([0, 2]))),
    ("perms", lambda d: d.__setitem__("basis0_perms", np.zeros_like(d["basis0_perms"]))),
    ("shape", lambda d: d.__setitem__("basis0_shape", np.array([6, 70, 3, 1]))),
    ("irrep dim", lambda d: d.__setitem__("basis0_shape", np.array([6, 6, 3, 2]))),
    ("characters", lambda d: d.__setitem__("basis0_characters", d["basis0_characters"][:2])),
    ("vector length", lambda d: d.__setitem__("vector_offset", np.array([0, 3, d["vector_offset"][-1]]))),
    ("momentum offsets", lambda d: d.__setitem__("level_momentum_offset", np.array([0, 5, 2]))),
])
def test_damaged_files_are_refused(tmp_path, what, edit):
    with pytest.raises(errors.InvalidRequest):
        ev.load_eigs(_damaged(tmp_path, edit))


def test_a_result_without_vectors_says_so():
    r = ev.EigsResult(energies=np.array([-1.0]), k=1, n_sites=4,
                      levels=[{"energy": -1.0, "n_up": 2, "momentum": 0, "parity": None, "dim": 2, "vector": -1}])
    for call in (lambda: r.vectors(), lambda: r.expect(Operator(4).add_product("z", (0,), 1.0)),
                 lambda: r.correlations(Operator(4).add_product("z", (0,), 1.0))):
        with pytest.raises(errors.InvalidRequest):
            call()


def test_spin_squared_is_the_casimir():
    """S^2 of three spins 1/2: eigenvalues 3/4 (two doublets) and 15/4 (the quartet), four each;
    of two: 0 (singlet) once and 2 (triplet) three times."""
    w3 = np.linalg.eigvalsh(po.dense_matrix(ev.spin_squared(3).products, 3))
    assert np.allclose(w3, [0.75] * 4 + [3.75] * 4)
    w2 = np.linalg.eigvalsh(po.dense_matrix(ev.spin_squared(2).products, 2))
    assert np.allclose(w2, [0.0, 2.0, 2.0, 2.0])


def test_total_spin_requests_are_checked():
    assert ev.two_spin("min", 12) == 0 and ev.two_spin("min", 9) == 1
    assert ev.two_spin(1.5, 9) == 3 and ev.two_spin(2, 12) == 4
    for bad, N in ((1.0, 9), (0.5, 12), (0.3, 12), (7, 12), (-1, 12), ("max", 12)):
        with pytest.raises(errors.InvalidRequest):
            ev.two_spin(bad, N)


def test_total_spin_record_is_the_lowest_singlets():
    """The reference's j1j2_chain12 total_spin=min record (Sz = 0 sector, no spatial symmetry) is the
    4 lowest eigenvalues of H on the S^2 = 0 eigenspace of the dense Sz = 0 block, each singlet once."""
    import json
    import os

    from conftest import FIXTURES

    with open(os.path.join(FIXTURES, "reference_records.json")) as f:
        want = json.load(f)["records"]["j1j2_chain12/api/eigs/total_spin=min/k=4"]["values"]["energies"]
    N = 12
    H = Operator(N)
    for d, J in ((1, 1.0), (2, 0.5)):
        for i in range(N):
            j = (i + d) % N
            H.add_product("zz", (i, j), J).add_product("+-", (i, j), 0.5 * J).add_product("-+", (i, j), 0.5 * J)
    states = [x for x in range(1 << N) if bin(x).count("1") == N // 2]
    Hd = po.dense_matrix(H.products, N, states)
    w, Q = np.linalg.eigh(po.dense_matrix(ev.spin_squared(N).products, N, states))
    Q = Q[:, np.abs(w) < 1e-8]
    got = np.linalg.eigvalsh(Q.conj().T @ Hd @ Q)[:4]
    assert np.allclose(got, want, atol=1e-10)

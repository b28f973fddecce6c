"""GPU parity of the HIP path against the oracle and the reference's fixtures.

Bit-exact: representatives, 1/norms, key order, index_and_projection, the merged CSR (columns and
values). Matvec within 1e-13 relative norm (both lanes, single / multi / block vectors; multi equals
single bit for bit). Eigenvalues within 1e-10 of dense diagonalisation. Every case prints its name
before it touches the device."""

import math
import sys

import numpy as np
import pytest

from conftest import fixture_names, fixture_products, load_fixture
from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import symmetry as sym

pytestmark = pytest.mark.gpu

NAMES = fixture_names()


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def fixture_sector(fx):
    g = qs.GroupSpec(perms=fx["perms"], chars=fx["chars"], flips=fx["flips"])
    return qs.Sector(g, int(fx["n_up"]), sublattice=int(fx["sublattice_mode"]))


def fixture_operator(fx):
    op = qs.Operator(int(fx["N"]))
    for ops, sites, c in fixture_products(fx):
        op.add_product(ops, sites, c)
    return op


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("name", NAMES)
def test_sector_bit_exact(name):
    say("sector", name)
    fx = load_fixture(name)
    s = fixture_sector(fx)
    reps, inv = s.export()
    assert s.n_orbits == int(fx["n_orbits"][0])
    assert np.array_equal(reps, fx["reps"])
    assert np.array_equal(bits(inv), bits(fx["inv_norms"]))
    kb, L = s.key_bits()
    if "key_bit" in fx:
        assert L > 0 and np.array_equal(kb, fx["key_bit"])
    else:
        assert L == 0


@pytest.mark.parametrize("name", NAMES)
def test_index_and_projection_bit_exact(name):
    say("index_and_projection", name)
    fx = load_fixture(name)
    s = fixture_sector(fx)
    idx, proj = s.index_and_projection(fx["probes"])
    assert np.array_equal(idx, fx["probe_index"])
    ref = fx["probe_proj"].view(np.complex128)
    hit = idx >= 0
    assert np.array_equal(bits(proj[hit].view(np.float64)), bits(ref[hit].view(np.float64)))


@pytest.mark.parametrize("name", NAMES)
def test_csr_bit_exact(name):
    say("csr", name)
    fx = load_fixture(name)
    s = fixture_sector(fx)
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=qs.LANE_CSR)
    row_ptr, col, val = m.export_csr()
    assert np.array_equal(row_ptr, fx["csr_row_ptr"])
    assert np.array_equal(col, fx["csr_col"])
    ref = fx["csr_val"].view(np.complex128)
    if m.is_real:   # a real operator keeps the real parts (its imaginary parts are zeros)
        assert np.array_equal(bits(val.real), bits(ref.real))
        assert np.all(ref.imag == 0.0)
    else:
        assert np.array_equal(bits(val.view(np.float64)), bits(ref.view(np.float64)))
    assert m.info["nnz"] == int(fx["csr_meta"][0])
    if m.info["value_bytes"] in (1, 2):
        assert m.info["dict_size"] == int(fx["csr_meta"][1])
    assert m.info["id_width_ref"] == int(fx["csr_meta"][2])


@pytest.mark.parametrize("lane", [qs.LANE_GATHER, qs.LANE_CSR])
@pytest.mark.parametrize("name", NAMES)
def test_matvec_matches_reference(name, lane):
    say("matvec", name, lane)
    fx = load_fixture(name)
    s = fixture_sector(fx)
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=lane)
    x = fx["gather_in"].view(np.complex128)
    ref = fx["gather_out"].view(np.complex128)
    y = m.apply_host(x, complex_vectors=True)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    if m.is_real:   # real vectors: the real and imaginary parts separately
        yr = m.apply_host(x.real) + 1j * m.apply_host(x.imag)
        assert np.linalg.norm(yr - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("lane", [qs.LANE_GATHER, qs.LANE_CSR])
def test_multi_equals_single_bitwise(lane):
    say("multi", lane)
    fx = load_fixture("j1j2_n16_k3_slc")
    s = fixture_sector(fx)
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=lane)
    rng = np.random.default_rng(7)
    xs = [rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim) for _ in range(11)]
    X = [qs.DeviceArray.from_host(x) for x in xs]
    Y = [qs.DeviceArray(s.dim) for _ in xs]
    m.apply_multi(X, Y)
    for x, y in zip(xs, Y):
        single = m.apply_host(x)
        assert np.array_equal(bits(y.to_host().view(np.float64)), bits(single.view(np.float64)))
    if lane == qs.LANE_CSR:
        for nv in (1, 2, 4, 8):
            B = np.stack(xs[:nv], axis=1)   # row-major rows x nv
            XB = qs.DeviceArray.from_host(B.ravel())
            YB = qs.DeviceArray(s.dim * nv)
            m.apply_block(nv, XB, YB)
            got = YB.to_host().reshape(s.dim, nv)
            for i in range(nv):
                assert np.array_equal(bits(got[:, i].copy().view(np.float64)),
                                      bits(m.apply_host(xs[i]).view(np.float64)))


def dense_of(m):
    row_ptr, col, val = m.export_csr()
    A = np.zeros((m.dim, m.dim), np.complex128)
    for r in range(m.dim):
        A[r, col[row_ptr[r]:row_ptr[r + 1]]] += val[row_ptr[r]:row_ptr[r + 1]]
    return A


@pytest.mark.parametrize("name", ["j1j2_n16_k3_plain", "j1j2_n16_d_k0_slc", "chiral_n10_k2"])
def test_lanczos_and_lowest_levels_against_dense(name):
    say("lanczos", name)
    fx = load_fixture(name)
    s = fixture_sector(fx)
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=qs.LANE_CSR)
    w = np.linalg.eigvalsh(dense_of(m))
    rng = np.random.default_rng(2)
    v0 = rng.normal(size=s.dim) if m.is_real else rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim)
    for reorth in (0, 1, 2):
        a, b, _ = m.lanczos(v0, min(150, s.dim), reorth)
        ritz = po.tridiag_eig(a, b)[0]
        assert abs(ritz[0] - w[0]) <= 1e-10 * max(1.0, abs(w[0]))
    vals, conv, _ = m.lowest(1)
    assert conv and abs(vals[0] - w[0]) <= 1e-10 * max(1.0, abs(w[0]))
    vals, conv, _ = m.lowest(4)
    assert conv
    assert np.allclose(vals, w[:4], rtol=1e-10, atol=1e-10)


def j1j2_operator(N, J2):
    b = qs.HamiltonianBuilder(N)
    b.heisenberg([(i, (i + 1) % N) for i in range(N)], 1.0)
    b.heisenberg([(i, (i + 2) % N) for i in range(N)], J2)
    return b.to_operator()


def test_heisenberg_n20_k0_lowest4_against_oracle():
    """The N = 20 ring's k = 0 sector (dim 9252): four lowest levels against the oracle's own sector
    and a dense diagonalisation of its merged CSR."""
    say("heisenberg n20 k0")
    N = 20
    op = j1j2_operator(N, 0.0)
    g = sym.chain_group(N, 0)
    s = qs.Sector(g, N // 2)
    assert s.dim == 9252
    o = po.OracleSector(N, N // 2, g.perms, g.chars)
    reps, inv = s.export()
    assert np.array_equal(reps, o.reps) and np.array_equal(bits(inv), bits(o.inv_norms))
    m = qs.SectorMatvec(op, s, lane=qs.LANE_CSR)
    row_ptr, col, val = po.sector_csr(qs.Program(op).export(), o)
    A = np.zeros((o.dim, o.dim))
    for r in range(o.dim):
        A[r, col[row_ptr[r]:row_ptr[r + 1]]] += val[row_ptr[r]:row_ptr[r + 1]].real
    w = np.linalg.eigvalsh(A)
    vals, conv, _ = m.lowest(4)
    assert conv and np.allclose(vals, w[:4], rtol=1e-10, atol=0)


def test_majumdar_ghosh_n32_full_size():
    """The N = 32 J1-J2 ring at J2 = 1/2, k = 0, reflection-even: dim 9,398,520 (Burnside), the
    Hermitian form <x, H y> = <H x, y>, and E0 = -3N/8 = -12 exactly (the Majumdar-Ghosh point)."""
    say("majumdar-ghosh n32")
    N = 32
    op = j1j2_operator(N, 0.5)
    s = qs.Sector(sym.dihedral_group(N, 0, +1), N // 2)
    assert s.dim == 9398520
    assert s.sublattice_block == 16
    m = qs.SectorMatvec(op, s, lane=qs.LANE_CSR)
    assert m.is_real
    rng = np.random.default_rng(4)
    x = rng.normal(size=s.dim)
    y = rng.normal(size=s.dim)
    hx, hy = m.apply_host(x), m.apply_host(y)
    assert abs(x @ hy - hx @ y) <= 1e-12 * abs(x @ hy)
    vals, conv, _ = m.lowest(1)
    assert conv and abs(vals[0] + 12.0) <= 1e-10 * 12.0


def test_device_memory_probe():
    say("device memory probe")
    from qed_spin_amd import _lib

    free, total = _lib.device_memory()
    assert 0 < free <= total
    assert total >= 64 << 30   # an MI355X holds 288 GB


def test_sectors_built_on_threads_with_own_streams_equal_sequential_builds():
    """Sectors and CSR operators built on host threads, each with a stream of its own, are the
    same bit for bit as the ones built one after another on the calling thread."""
    say("threaded builds")
    from qed_spin_amd.sector import parallel_build

    names = NAMES[:6]
    seq = [fixture_sector(load_fixture(n)).export() for n in names]
    par = parallel_build(lambda n: fixture_sector(load_fixture(n)).export(), names, threads=3)
    for (r1, i1), (r2, i2) in zip(seq, par):
        assert np.array_equal(r1, r2) and np.array_equal(bits(i1), bits(i2))
    n = 14
    H = qs.HamiltonianBuilder(n).heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0).to_operator()
    blks = qs.Symmetry(sz=[n // 2]).blocks(n)

    def csr(b):
        s = qs.Sector(b.group, b.n_up)
        m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
        return m.export_csr()

    one = [csr(b) for b in blks]
    many = parallel_build(csr, blks, threads=4)
    for a, b in zip(one, many):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_dictionary_keeps_values_and_their_negatives_apart():
    """A momentum sector whose merged values hold pairs v, -v (J1-J2 ring N = 22, k = 1: 345 distinct
    values) keeps its value dictionary (ids, not full values), and its entries are the oracle's bit
    for bit (as a multiset: the sector may list its representatives in key order)."""
    say("dictionary of v and -v")
    n = 22
    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    b.heisenberg([(i, (i + 2) % n) for i in range(n)], 0.3)
    H = b.to_operator()
    g = sym.chain_group(n, 1)
    s = qs.Sector(g, n // 2)
    m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    assert m.info["value_bytes"] <= 2 and m.info["dict_size"] == 345
    os_ = po.OracleSector(n, n // 2, g.perms, g.chars)
    rp, col, val = po.sector_csr(qs.Program(H, adjoint=True).export(), os_)
    r2, c2, v2 = m.export_csr()
    assert len(c2) == len(col)
    w1 = np.sort(bits(np.asarray(val).view(np.float64)).reshape(-1, 2).view([("re", np.uint64), ("im", np.uint64)]), axis=0)
    w2 = np.sort(bits(np.asarray(v2).view(np.float64)).reshape(-1, 2).view([("re", np.uint64), ("im", np.uint64)]), axis=0)
    assert np.array_equal(w1, w2)


@pytest.mark.parametrize("name", NAMES)
def test_imported_reference_sector_gives_the_reference_csr(name):
    """The reference's own representatives and 1/norms (its RepSectorData, written by its code into the
    fixture) adopted as they are (qedh_sector_import): the CSR built on them is the reference's bit for
    bit, and a representative out of order is refused by name."""
    say("sector import", name)
    fx = load_fixture(name)
    g = qs.GroupSpec(perms=fx["perms"], chars=fx["chars"], flips=fx["flips"])
    s = qs.Sector(g, int(fx["n_up"]), sublattice=int(fx["sublattice_mode"]), reps=fx["reps"], inv_norms=fx["inv_norms"])
    assert s.dim == len(fx["reps"])
    reps, inv = s.export()
    assert np.array_equal(reps, fx["reps"]) and np.array_equal(bits(inv), bits(fx["inv_norms"]))
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=qs.LANE_CSR)
    row_ptr, col, val = m.export_csr()
    assert np.array_equal(row_ptr, fx["csr_row_ptr"]) and np.array_equal(col, fx["csr_col"])
    ref = fx["csr_val"].view(np.complex128)
    if m.is_real:
        assert np.array_equal(bits(val.real), bits(ref.real))
    else:
        assert np.array_equal(bits(val.view(np.float64)), bits(ref.view(np.float64)))
    if s.dim >= 3:
        bad = np.array(fx["reps"], np.uint64).copy()
        bad[[1, 2]] = bad[[2, 1]]
        with pytest.raises(qs.errors.InvalidRequest):
            qs.Sector(g, int(fx["n_up"]), sublattice=int(fx["sublattice_mode"]), reps=bad, inv_norms=fx["inv_norms"])


def test_apply_on_host_buffers_equals_device_apply():
    """qedh_matvec_apply_host (the reference's LinearOperator::apply on host buffers) gives the device
    apply's output bit for bit, real and complex vectors."""
    import ctypes as C

    from qed_spin_amd import _lib

    say("apply on host buffers")
    fx = load_fixture(NAMES[0])
    s = fixture_sector(fx)
    m = qs.SectorMatvec(fixture_operator(fx), s, lane=qs.LANE_CSR)
    rng = np.random.default_rng(5)
    for cv in (False, True):
        dt = m.dtype(cv)
        x = rng.normal(size=m.dim) + (1j * rng.normal(size=m.dim) if dt == np.complex128 else 0.0)
        x = np.ascontiguousarray(x, dt)
        y = np.zeros(m.dim, dt)
        _lib.call("qedh_matvec_apply_host", m.h, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), int(cv))
        want = m.apply_host(x, complex_vectors=cv)
        assert np.array_equal(bits(np.asarray(y).view(np.float64)), bits(np.asarray(want).view(np.float64)))

"""The Krylov drivers' vector backend through the C-ABI (BasicBackend<Scalar>,
include/ed/matvec/backend.h:56-179): every entry point against numpy on device vectors, real and
complex, including column arrays of separately allocated vectors and more columns than one launch
takes; the release fence of objects shared across threads; the eigs walk with a time-reversal pair
listed mirror first. Every case prints its name before it touches the device."""

import ctypes as C
import sys
import threading

import numpy as np
import pytest

import qed_spin_amd as qs
from qed_spin_amd import _lib, errors
from qed_spin_amd import eigs as ev
from qed_spin_amd import symmetry as sym
from qed_spin_amd.sector import DeviceArray, parallel_build

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _vec(rng, n, cplx):
    x = rng.normal(size=n)
    if cplx:
        x = x + 1j * rng.normal(size=n)
    return x


def _dev(x):
    return DeviceArray.from_host(np.ascontiguousarray(x))


def _scal(a, cplx):
    return np.array([a.real, a.imag] if cplx else [a.real], np.float64)


def _p(a):
    return _lib.ptr(a, C.c_double)


SIZES = [1, 1000, 1_000_003]


@pytest.mark.parametrize("cplx", [0, 1])
@pytest.mark.parametrize("n", SIZES)
def test_level1_against_numpy(cplx, n):
    say("backend level 1", cplx, n)
    rng = np.random.default_rng(n + cplx)
    x, y, z = (_vec(rng, n, cplx) for _ in range(3))
    a = complex(0.7, -0.3) if cplx else 0.7
    b = complex(-1.1, 0.4) if cplx else -1.1
    tol = 1e-13
    dx, dy, dz = _dev(x), _dev(y), _dev(z)
    out = np.zeros(2)
    _lib.call("qedh_dot", cplx, dx.ptr, dy.ptr, n, _p(out))
    ref = np.vdot(x, y)
    got = complex(out[0], out[1]) if cplx else out[0]
    assert abs(got - ref) <= tol * np.linalg.norm(x) * np.linalg.norm(y)
    _lib.call("qedh_nrm2", cplx, dx.ptr, n, _p(out))
    assert abs(out[0] - np.linalg.norm(x)) <= tol * np.linalg.norm(x)
    # y <- a x + y
    _lib.call("qedh_axpy", cplx, _p(_scal(a, cplx)), dx.ptr, dy.ptr, n)
    y1 = a * x + y
    assert np.allclose(dy.to_host(), y1, rtol=0, atol=1e-14 * np.abs(y1).max())
    # y <- y + b z, then z... axpy_dot returns x^H y after the update
    _lib.call("qedh_axpy_dot", cplx, _p(_scal(b, cplx)), dz.ptr, dy.ptr, dx.ptr, n, _p(out))
    y2 = y1 + b * z
    got = complex(out[0], out[1]) if cplx else out[0]
    assert abs(got - np.vdot(x, y2)) <= tol * np.linalg.norm(x) * np.linalg.norm(y2)
    assert np.allclose(dy.to_host(), y2, rtol=0, atol=1e-14 * np.abs(y2).max())
    _lib.call("qedh_axpy_nrm2", cplx, _p(_scal(a, cplx)), dx.ptr, dy.ptr, n, _p(out))
    y3 = y2 + a * x
    assert abs(out[0] - np.linalg.norm(y3)) <= tol * np.linalg.norm(y3)
    _lib.call("qedh_axpby", cplx, _p(_scal(b, cplx)), dz.ptr, _p(_scal(a, cplx)), dy.ptr, n)
    y4 = b * z + a * y3
    assert np.allclose(dy.to_host(), y4, rtol=0, atol=1e-14 * np.abs(y4).max())
    _lib.call("qedh_scale", cplx, _p(_scal(b, cplx)), dy.ptr, n)
    assert np.allclose(dy.to_host(), b * y4, rtol=0, atol=1e-14 * np.abs(b * y4).max())
    _lib.call("qedh_fill_zero", cplx, dy.ptr, n)
    assert not np.any(dy.to_host())


@pytest.mark.parametrize("cplx", [0, 1])
@pytest.mark.parametrize("cols", [1, 5, 300])
def test_dot_many_axpy_many_on_separate_vectors(cplx, cols):
    """basis[] of separately allocated device vectors (not one strided block), more columns than one
    launch of the tall-skinny kernels takes (256)."""
    say("backend dot_many / axpy_many", cplx, cols)
    n = 40_001
    rng = np.random.default_rng(cols)
    vs = [_vec(rng, n, cplx) for _ in range(cols)]
    w = _vec(rng, n, cplx)
    dvs = [_dev(v) for v in vs]
    dw = _dev(w)
    table = (C.c_void_p * cols)(*[d.ptr.value for d in dvs])
    out = np.zeros(cols * (2 if cplx else 1))
    _lib.call("qedh_dot_many", cplx, table, cols, dw.ptr, n, _p(out))
    got = out.view(np.complex128) if cplx else out
    ref = np.array([np.vdot(v, w) for v in vs])
    scale = np.linalg.norm(w) * max(np.linalg.norm(v) for v in vs)
    assert np.abs(got - ref).max() <= 1e-13 * scale
    alphas = _vec(rng, cols, cplx)
    a_in = np.ascontiguousarray(alphas).view(np.float64) if cplx else np.ascontiguousarray(alphas)
    _lib.call("qedh_axpy_many", cplx, _p(a_in), table, cols, dw.ptr, n)
    ref_w = w + sum(a * v for a, v in zip(alphas, vs))
    assert np.abs(dw.to_host() - ref_w).max() <= 1e-13 * np.abs(ref_w).max() * np.sqrt(cols)


@pytest.mark.parametrize("cplx", [0, 1])
@pytest.mark.parametrize("ops", ["NN", "CN", "NT", "TC"])
def test_gemm_against_numpy(cplx, ops):
    say("backend gemm", cplx, ops)
    rng = np.random.default_rng(3)
    m, n, k = 37, 11, 53
    op = {"N": lambda A: A, "T": lambda A: A.T, "C": lambda A: A.conj().T}
    shape_a = (m, k) if ops[0] == "N" else (k, m)
    shape_b = (k, n) if ops[1] == "N" else (n, k)
    mk = (lambda s: rng.normal(size=s) + 1j * rng.normal(size=s)) if cplx else (lambda s: rng.normal(size=s))
    A, B, Cm = mk(shape_a), mk(shape_b), mk((m, n))
    alpha = complex(0.5, 0.25) if cplx else 0.5
    beta = complex(-0.75, 0.5) if cplx else -0.75
    dA, dB, dC = (_dev(np.asfortranarray(M).ravel(order="F")) for M in (A, B, Cm))
    _lib.call("qedh_gemm", cplx, ops[0].encode(), ops[1].encode(), m, n, k, _p(_scal(alpha, cplx)), dA.ptr,
              shape_a[0], dB.ptr, shape_b[0], _p(_scal(beta, cplx)), dC.ptr, m)
    ref = alpha * op[ops[0]](A) @ op[ops[1]](B) + beta * Cm
    got = dC.to_host().reshape((n, m)).T
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * k


def test_gemm_refuses_bad_op_and_leading_dimension():
    say("backend gemm refusals")
    d = _dev(np.zeros(16))
    one = np.array([1.0])
    with pytest.raises(errors.InvalidRequest):
        _lib.call("qedh_gemm", 0, b"X", b"N", 2, 2, 2, _p(one), d.ptr, 2, d.ptr, 2, _p(one), d.ptr, 2)
    with pytest.raises(errors.InvalidRequest):
        _lib.call("qedh_gemm", 0, b"N", b"N", 4, 2, 2, _p(one), d.ptr, 2, d.ptr, 2, _p(one), d.ptr, 4)


def _stats():
    f, s = C.c_uint64(0), C.c_int64(0)
    _lib.call("qedh_stream_stats", C.byref(f), C.byref(s))
    return int(f.value), int(s.value)


def _ring(n, j2=0.0):
    b = qs.HamiltonianBuilder(n)
    b.heisenberg([(i, (i + 1) % n) for i in range(n)], 1.0)
    if j2:
        b.heisenberg([(i, (i + 2) % n) for i in range(n)], j2)
    return b.to_operator()


def test_shared_objects_fence_before_release_and_streams_are_reused():
    """Operators built and applied on worker threads (streams of their own), then dropped on
    another thread: each release fences every engine stream first. Repeated thread pools reuse the
    pooled streams instead of creating (and destroying) new ones."""
    say("release fence and stream pool")
    H = _ring(16, 0.5)

    def build(m):
        s = qs.Sector(sym.chain_group(16, m), 8)
        op = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
        x = np.random.default_rng(m).normal(size=s.dim) + 0j
        return op, op.apply_host(x, complex_vectors=True)

    built = parallel_build(build, list(range(8)))
    f0, s0 = _stats()
    assert s0 >= 1
    for _ in range(3):
        parallel_build(build, list(range(8)))
    f1, s1 = _stats()
    # a thread hands its stream back as it exits, which may land just after the pool's join returns:
    # at most one pool's worth of extra streams, not one per pool
    assert s1 <= s0 + 8, ("thread pools created new streams instead of reusing the pooled ones", s0, s1)
    assert f1 > f0
    ys = [y for _, y in built]
    ops = [op for op, _ in built]
    del built
    done = threading.Event()

    def drop():
        ops.clear()   # the last references, dropped on a thread of the null stream
        done.set()

    t = threading.Thread(target=drop)
    t.start()
    t.join()
    assert done.is_set()
    f2, _ = _stats()
    assert f2 >= f1 + 8, (f1, f2)
    # the same operators built again give the same applies (the freed memory was not in use)
    again = parallel_build(build, list(range(8)))
    for (_, y), y0 in zip(again, ys):
        assert np.array_equal(y, y0)


def test_eigs_mirror_listed_first_matches_all_momenta():
    """Symmetry(momenta=[N - 1, 1]) lists the time-reversal mirror before its partner: the walk still
    returns both copies of the degenerate levels, as momenta='all' restricted to those blocks does."""
    say("eigs mirror first")
    N = 12
    H = _ring(N, 0.3)
    a = ev.eigs(H, 4, sym=qs.Symmetry(sz=[N // 2], momenta=[N - 1, 1]), dense_max_dim=0)
    b = ev.eigs(H, 4, sym=qs.Symmetry(sz=[N // 2], momenta=[1, N - 1]), dense_max_dim=0)
    assert np.allclose(a.energies, b.energies, rtol=0, atol=1e-10)
    # the pair is degenerate: each of the lowest levels appears twice
    assert abs(a.energies[0] - a.energies[1]) < 1e-10 and abs(a.energies[2] - a.energies[3]) < 1e-10

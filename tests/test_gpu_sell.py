"""The SELL-64 copy of the merged CSR (64-row slices, one lane per row, each row summed in entry
order): its applies against the CSR layout's and the oracle's within 1e-13; bit for bit the same
under every gather depth and workgroup size the autotuner may pick; k-vector and block-interleaved
applies equal single applies bit for bit; a large tuned operator takes it by itself. Every case
prints its name before it touches the device."""

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

import qed_spin_amd as qs
from qed_spin_amd import _lib
from qed_spin_amd import symmetry as sym

pytestmark = pytest.mark.gpu

TUNE_XCD, TUNE_UNROLL, TUNE_UNROLL8, TUNE_WG256, TUNE_WG1024, TUNE_UNROLL2, TUNE_UNROLL16, TUNE_CSR, TUNE_SELL = (
    2, 4, 8, 32, 64, 128, 256, 512, 1024)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def tune(m, flags):
    _lib.call("qedh_matvec_set_tuning", m.h, 0, int(flags))


def ring(N, d=1):
    return [(i, (i + d) % N) for i in range(N)]


@pytest.mark.parametrize("N,chi", [(16, 0.0), (14, 0.25)])
def test_sell_matches_csr_and_is_independent_of_tuning(N, chi):
    say("sell vs csr", N, chi)
    b = qs.HamiltonianBuilder(N)
    b.heisenberg(ring(N), 1.0)
    b.heisenberg(ring(N, 2), 0.5)
    if chi:
        b.chirality([(j, (j + 1) % N, (j + 2) % N) for j in range(N)], chi)
    H = b.to_operator()
    g = sym.chain_group(N, 0 if chi == 0.0 else 1)   # a real block, and a complex one
    s = qs.Sector(g, N // 2)
    m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR, tune=False)
    rng = np.random.default_rng(N)
    x = rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim)
    tune(m, TUNE_XCD | TUNE_UNROLL)          # the CSR layout
    y_csr = m.apply_host(x, complex_vectors=True)
    tune(m, TUNE_XCD | TUNE_UNROLL8 | TUNE_SELL)
    assert m._info()["lanes"] == 1
    y = m.apply_host(x, complex_vectors=True)
    assert np.linalg.norm(y - y_csr) <= 1e-13 * np.linalg.norm(y_csr)
    # the oracle's CSR of the same sector
    o = po.OracleSector(N, N // 2, g.perms, g.chars)
    row_ptr, col, val = po.sector_csr(qs.Program(H).export(), o)
    ref = po.csr_matvec(row_ptr, col, val, x)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    # every gather depth and workgroup size: the same bits
    for flags in (0, TUNE_XCD, TUNE_UNROLL2, TUNE_UNROLL | TUNE_WG256, TUNE_UNROLL8 | TUNE_WG1024,
                  TUNE_XCD | TUNE_UNROLL16, TUNE_UNROLL16 | TUNE_WG256):
        tune(m, flags)
        assert np.array_equal(bits(m.apply_host(x, complex_vectors=True).view(np.float64)), bits(y.view(np.float64))), flags
    # k vectors at once and block-interleaved vectors: bit for bit the single applies
    tune(m, TUNE_XCD | TUNE_UNROLL8)
    xs = [rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim) for _ in range(11)]
    X = [qs.DeviceArray.from_host(v) for v in xs]
    Y = [qs.DeviceArray(s.dim) for _ in xs]
    m.apply_multi(X, Y, complex_vectors=True)
    singles = [m.apply_host(v, complex_vectors=True) for v in xs]
    for yv, single in zip(Y, singles):
        assert np.array_equal(bits(yv.to_host().view(np.float64)), bits(single.view(np.float64)))
    for nv in (1, 2, 4, 8):
        B = np.stack(xs[:nv], axis=1)
        XB = qs.DeviceArray.from_host(B.ravel())
        YB = qs.DeviceArray(s.dim * nv)
        m.apply_block(nv, XB, YB, complex_vectors=True)
        got = YB.to_host().reshape(s.dim, nv)
        for i in range(nv):
            assert np.array_equal(bits(got[:, i].copy().view(np.float64)), bits(singles[i].view(np.float64)))
    if m.is_real:   # real vectors through the real kernels
        xr = rng.normal(size=s.dim)
        yr = m.apply_host(xr)
        tune(m, TUNE_XCD | TUNE_UNROLL16)
        assert np.array_equal(bits(m.apply_host(xr)), bits(yr))
        ref_r = po.csr_matvec(row_ptr, col, val, xr.astype(complex)).real
        assert np.linalg.norm(yr - ref_r) <= 1e-13 * np.linalg.norm(ref_r)


def test_large_tuned_operator_takes_the_sell_layout():
    say("sell: large tuned operator")
    N = 28
    b = qs.HamiltonianBuilder(N)
    b.heisenberg(ring(N), 1.0)
    b.heisenberg(ring(N, 2), 0.5)
    H = b.to_operator()
    s = qs.Sector(sym.dihedral_group(N, 0, 1), N // 2)
    assert s.dim >= 1 << 18
    m = qs.SectorMatvec(H, s, lane=qs.LANE_CSR)
    assert m._info()["lanes"] == 1
    rng = np.random.default_rng(5)
    x = rng.normal(size=s.dim)
    y = m.apply_host(x)
    tune(m, TUNE_XCD | TUNE_UNROLL | TUNE_CSR)
    y_csr = m.apply_host(x)
    assert np.linalg.norm(y - y_csr) <= 1e-13 * np.linalg.norm(y_csr)

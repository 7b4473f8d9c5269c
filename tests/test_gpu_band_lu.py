"""BandGeneral factor/solve with the reference's row swaps, on the device,
against the oracle's restatement of the reference backend
(linear_solver_backend.mojo:631-679 band LU, :652-657 and :175-214 dense
partial-pivot fallback, :735-769 and :216-232 solves).

The device's pivoted band LU (the register kernel, the global-memory one
and the blocked one: sixteen-column panels and trailing updates over many
workgroups) makes the reference's pivot choices and divisions in the
reference's order, so its factors are the reference's; its back substitution sums
by columns where the reference sums by rows, so solutions agree to
rounding: within 1e-12 of the largest solution entry on these bands (the
measured figure is printed).  The bands are indefinite, nonsymmetric, with
zero and tiny diagonal entries -- the cases where the reference leaves its
band LU for the dense fallback and where elimination without swaps fails
or loses accuracy -- plus a singular one.
"""

import numpy as np
import pytest

from oracle import oracle
from strut_amd.linear_system import BandSystem
from strut_amd.model import MaterialTable
from strut_amd.sections import SectionRuntime
from strut_amd.transient3d import band_solve

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def rt():
    r = SectionRuntime(MaterialTable([{"id": 1, "type": "Elastic", "params": {"E": 1.0}}]), device=0)
    yield r
    r.close()


def _to_band(A, hbw):
    n = len(A)
    band = np.zeros((n, 2 * hbw + 1))
    for i in range(n):
        for j in range(max(0, i - hbw), min(n - 1, i + hbw) + 1):
            band[i, j - i + hbw] = A[i, j]
    return band


def _general_band(n, hbw, seed, kind):
    """kind: 'indefinite' symmetric with random signs and a weak diagonal;
    'nonsym' nonsymmetric; 'zero_diag' symmetric with zero diagonal entries
    at every third row; 'tiny_pivot' diagonally dominant except a 1e-13
    first pivot under a unit sub-diagonal entry."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for d in range(1, hbw + 1):
        lo = rng.normal(size=n - d)
        up = lo if kind in ("indefinite", "zero_diag") else rng.normal(size=n - d)
        A[np.arange(d, n), np.arange(n - d)] = lo
        A[np.arange(n - d), np.arange(d, n)] = up
    if kind == "tiny_pivot":
        A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=1) + 1.0
        A[0, 0] = 1e-13
        if n > 1:
            A[1, 0] = A[0, 1] = 1.0
    else:
        A[np.arange(n), np.arange(n)] = rng.normal(size=n) * 0.5
        if kind == "zero_diag":
            A[np.arange(0, n, 3), np.arange(0, n, 3)] = 0.0
    return A


CASES = [(1, 0), (5, 1), (9, 3), (40, 39), (130, 64), (300, 110), (2300, 110), (700, 119), (520, 127), (200, 150),
         (60, 59), (17, 2), (900, 300), (1200, 560)]


@pytest.mark.parametrize("n,hbw", CASES)
@pytest.mark.parametrize("kind", ["indefinite", "nonsym", "zero_diag", "tiny_pivot"])
def test_pivoted_band_solve_matches_the_reference_backend(rt, n, hbw, kind):
    if n == 1 and kind != "tiny_pivot":
        kind = "tiny_pivot"  # a 1 x 1 system: only a nonzero pivot makes sense
    A = _general_band(n, hbw, seed=7 * n + hbw, kind=kind)
    if n == 1:
        A[0, 0] = 2.5
    band = _to_band(A, hbw)
    rhs = np.random.default_rng(n + 1).normal(size=n)
    ok, want, fallback = oracle.band_solve(band, hbw, rhs)
    assert ok
    fits = hbw <= 127 and n <= 3072
    for variant in ((0, 5, 6, 8) if fits else (0, 6, 8)):
        x, info = band_solve(rt, band, rhs, variant=variant)
        err = np.abs(x - want).max() / np.abs(want).max()
        print(f"n {n} hbw {hbw} {kind} variant {variant}: info {info} reference fallback {fallback} err {err:.2e}",
              flush=True)
        assert info == (1 if fallback else 0)
        assert err <= TOL


def test_elimination_without_swaps_fails_where_the_reference_pivots(rt):
    """A zero leading pivot: the reference's scan swaps rows and solves; the
    elimination without swaps (variants 1-4, 7) stops on the pivot."""
    n, hbw = 40, 6
    A = _general_band(n, hbw, seed=3, kind="zero_diag")
    band = _to_band(A, hbw)
    rhs = np.arange(1.0, n + 1.0)
    ok, want, fallback = oracle.band_solve(band, hbw, rhs)
    assert ok and fallback
    x, info = band_solve(rt, band, rhs, variant=0)
    assert info == 1 and np.abs(x - want).max() <= TOL * np.abs(want).max()
    assert np.abs(A @ x - rhs).max() <= 1e-10 * np.abs(rhs).max()
    _, info_np = band_solve(rt, band, rhs, variant=7)
    assert info_np == 2


def test_singular_band_is_reported(rt):
    """A column whose largest magnitude is 0: the reference aborts
    (BandGeneral matrix is singular or near-singular); the device reports 2."""
    n, hbw = 30, 4
    A = _general_band(n, hbw, seed=5, kind="nonsym")
    A[:, 11] = 0.0
    band = _to_band(A, hbw)
    rhs = np.ones(n)
    ok, _, _ = oracle.band_solve(band, hbw, rhs)
    assert not ok
    for variant in (0, 5, 6, 8):
        _, info = band_solve(rt, band, rhs, variant=variant)
        assert info == 2


@pytest.mark.parametrize("n,hbw", [(12, 3), (300, 110), (400, 140)])
def test_band_system_factor_once_solve_many(rt, n, hbw):
    """create / load (dense, out-of-band entries dropped) / add scaled /
    factor / several solves, against the oracle on the in-band matrix; the
    split solve gives the one-shot solve's bits where both take the register
    kernel (wider bands: the blocked factorization, whose one-shot back
    substitution sums by rows where the split solve sums by columns)."""
    rng = np.random.default_rng(n)
    K = _general_band(n, hbw, seed=n, kind="indefinite")
    M = np.diag(1.0 + rng.random(n))
    K_dense = K.copy()
    K_dense[0, n - 1] = K_dense[n - 1, 0] = 123.0  # outside the band: dropped as the reference drops it
    a0 = 0.75
    eff = _to_band(K, hbw) + a0 * _to_band(M, hbw)
    sysm = BandSystem(rt, n, hbw)
    sysm.load(K_dense)
    sysm.load(M, scale=a0, add=True)
    info = sysm.factor()
    for s in range(3):
        rhs = rng.normal(size=n)
        ok, want, fallback = oracle.band_solve(eff, hbw, rhs)
        assert ok and info == (1 if fallback else 0)
        x = sysm.solve(rhs)
        assert np.abs(x - want).max() <= TOL * np.abs(want).max()
        one_shot, _ = band_solve(rt, eff, rhs, variant=0)
        if hbw <= 127:
            assert np.array_equal(x, one_shot)
        else:
            assert np.abs(x - one_shot).max() <= TOL * np.abs(want).max()
    sysm.close()


def test_band_system_refuses_out_of_order_calls(rt):
    from strut_amd._native import StrutGpuError

    sysm = BandSystem(rt, 10, 2)
    with pytest.raises(StrutGpuError, match="no factorization"):
        sysm.solve(np.ones(10))
    with pytest.raises(StrutGpuError, match="no loaded values"):
        sysm.factor()
    with pytest.raises(StrutGpuError, match="no loaded values"):
        sysm.load(np.eye(10), add=True)
    with pytest.raises(ValueError):
        sysm.load(np.eye(9))
    sysm.close()
    with pytest.raises(StrutGpuError):
        BandSystem(rt, 10, 10)


@pytest.mark.parametrize("n,hbw", [(5, 1), (9, 3), (40, 39), (130, 64), (300, 110), (2300, 110), (200, 150),
                                   (700, 300), (640, 200)])
@pytest.mark.parametrize("kind", ["indefinite", "nonsym", "zero_diag", "tiny_pivot"])
def test_factors_are_the_reference_factors(rt, n, hbw, kind):
    """The U rows of the device factorization equal, bit for bit, those of
    the oracle's reference backend: its band LU's where it needs no swap,
    its dense partial-pivot LU's rows where it falls back.  Bands wider than
    the register kernel holds go through the blocked kernels."""
    A = _general_band(n, hbw, seed=11 * n + hbw, kind=kind)
    band = _to_band(A, hbw)
    ok, want, fallback = oracle.band_factor_rows(band, hbw)
    assert ok
    sysm = BandSystem(rt, n, hbw)
    sysm.load(band, band_layout=True)
    info = sysm.factor()
    got, _, piv = sysm.factors()
    sysm.close()
    diff = np.abs(got - want)
    print(f"n {n} hbw {hbw} {kind}: info {info} fallback {fallback} rows swapped {int(np.sum(piv != np.arange(n)))} "
          f"max |U - U_ref| {diff.max():.2e}", flush=True)
    assert info == (1 if fallback else 0)
    np.testing.assert_array_equal(got, want)

"""GPU parity of the 3D frame transient (sgpu_frame3d_*, strut_amd/csrc/
frame3d.hip) with the oracle's driver (oracle/transient3d.c) on the same
seeded cases: RC frames with forceBeamColumn3d columns and dispBeamColumn3d
beams, rigid floor diaphragms under the Transformation handler and two
simultaneous horizontal records.  The bar: the same status, the same Newton
iteration count at every step, and every displacement within 1e-8 of the
history's peak.  The device's band LU makes the reference's row swaps
(where the reference hands a factorization to its dense partial-pivot LU,
the device pivots within the band and gets the same factors), and the count
of such factorizations must equal the oracle's; the elimination without
swaps, kept as an opt-in, is checked on one frame too.  Both band rules
run: the reference's (element free map, whose truncation of the collapsed
tangent costs extra iterations) and the full collapsed band."""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.frame3d import load_frame3d_case
from strut_amd.transient3d import Frame3dRun

pytestmark = pytest.mark.gpu


def _compare(fc, hbw_rule, steps, band_solver=0):
    ref = oracle.run_transient3d(fc, hbw_rule=hbw_rule, steps=steps)
    dev = Frame3dRun(fc, hbw_rule=hbw_rule, band_solver=band_solver)
    assert dev.hbw == ref["hbw"] and dev.free_count == ref["free_count"]
    got = dev.run(steps)
    dev.close()
    print(f"rule {hbw_rule}: hbw {dev.hbw} n {dev.free_count} status {got['status']}/{ref['status']} "
          f"iters {got['iters'].sum()}/{ref['iters'].sum()} dense fallbacks {got['dense_fallbacks']}/"
          f"{ref['dense_fallbacks']}", flush=True)
    assert got["status"] == ref["status"]
    assert got["steps_done"] == ref["steps_done"]
    assert got["dense_fallbacks"] == ref["dense_fallbacks"]
    k = ref["steps_done"]
    np.testing.assert_array_equal(got["iters"][:k], ref["iters"][:k])
    scale = np.abs(ref["u"][:k]).max()
    err = np.abs(got["u"][:k] - ref["u"][:k]).max()
    assert err <= 1e-8 * scale, err / scale
    return ref


@pytest.mark.parametrize("hbw_rule", [0, 1])
def test_gpu_effective_band_matches_oracle(hbw_rule):
    """Element state, collapse, band gather and K + a1 C + a0 M of the first
    Newton solve, against the band the oracle factors there."""
    case = cases.rc_frame3d_case(2, 2, 1, steps=2, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    fc = load_frame3d_case(case)
    ref = oracle.run_transient3d(fc, hbw_rule=hbw_rule, steps=1, capture_band=True)
    dev = Frame3dRun(fc, hbw_rule=hbw_rule)
    got = dev.effective_band()
    dev.close()
    want = ref["band_first"]
    diff = np.abs(got - want)
    worst = np.unravel_index(np.argmax(diff), diff.shape)
    print(f"band rule {hbw_rule}: max diff {diff.max():.3e} at {worst} got {got[worst]:.17g} want {want[worst]:.17g}; "
          f"scale {np.abs(want).max():.3e}", flush=True)
    assert diff.max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("hbw_rule", [0, 1])
def test_gpu_diaphragm_frame_bidirectional_matches_oracle(hbw_rule):
    case = cases.rc_frame3d_case(2, 2, 1, steps=100, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    ref = _compare(load_frame3d_case(case), hbw_rule, 100)
    assert ref["status"] == 0 and np.abs(ref["u"]).max() > 1e-2  # well into the nonlinear range


@pytest.mark.parametrize("hbw_rule", [0, 1])
def test_gpu_blocked_band_solver_matches_oracle(hbw_rule):
    """The blocked row-swap kernels (sixteen-column panels, trailing updates
    over many workgroups) in the 3D driver: the same bar as the register
    kernel."""
    case = cases.rc_frame3d_case(2, 2, 1, steps=60, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    _compare(load_frame3d_case(case), hbw_rule, 60, band_solver=2)


def test_gpu_frame_without_constraints_matches_oracle():
    case = cases.rc_frame3d_case(2, 1, 1, steps=60, dt=0.005, n_core=4, n_bars=4, diaphragms=False,
                                 bidirectional=False, pgas_g=(0.6, 0.0))
    _compare(load_frame3d_case(case), 0, 60)


def test_gpu_band_elimination_without_swaps_opt_in_matches_oracle():
    """The opt-in solver that eliminates without row swaps (it counts where
    the reference would swap): on this frame's symmetric, diagonally strong
    tangents it follows the oracle within the displacement bar."""
    case = cases.rc_frame3d_case(2, 2, 1, steps=60, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    _compare(load_frame3d_case(case), 0, 60, band_solver=1)


def test_gpu_config1_full_size_first_steps_match_oracle():
    """Configuration 1 at its own size (20 stories x 5 x 5 bays, 1920
    elements, 268-fiber sections, rigid diaphragms, both records, the
    reference's band rule: 2220 free DOFs, half bandwidth 110): the first
    three steps against the oracle's driver (dense K, collapse, band LU with
    the dense fallback), identical iteration counts, dense-fallback count
    and displacements within 1e-8 of the peak."""
    fc = load_frame3d_case(cases.config1_case(steps=3))
    ref = _compare(fc, 0, 3)
    assert ref["hbw"] == 110 and ref["free_count"] == 2220 and ref["status"] == 0


def _spd_band(n, hbw, seed, dominant):
    """A symmetric positive definite band: diagonally dominant, or with
    off-diagonal entries above the diagonal (where the reference's factor
    moves to its dense partial-pivot LU)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    if dominant:
        for d in range(1, hbw + 1):
            v = rng.normal(size=n - d)
            A[np.arange(n - d), np.arange(d, n)] = v
            A[np.arange(d, n), np.arange(n - d)] = v
        A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=1) + 1.0
    else:  # B B^T with B lower banded (half the band) and a small B[0, 0]
        B = np.diag(1.0 + np.abs(rng.normal(size=n)))
        h2 = hbw // 2
        for d in range(1, h2 + 1):
            B[np.arange(d, n), np.arange(n - d)] = rng.normal(size=n - d) * (0.5 / h2)
        B[0, 0] = 0.1
        if n > 1 and h2 > 0:
            B[1, 0] = 0.8  # A[1, 0] = 0.08 outweighs A[0, 0] = 0.01
        A = B @ B.T
    band = np.zeros((n, 2 * hbw + 1))
    for i in range(n):
        for j in range(max(0, i - hbw), min(n - 1, i + hbw) + 1):
            band[i, j - i + hbw] = A[i, j]
    return A, band


@pytest.mark.parametrize("n,hbw", [(1, 0), (7, 0), (9, 3), (40, 39), (300, 110), (2300, 110), (700, 119), (520, 127),
                                   (200, 150)])
@pytest.mark.parametrize("dominant", [True, False])
def test_gpu_band_solve_kernels_match_the_reference_backend(n, hbw, dominant):
    from strut_amd.model import MaterialTable
    from strut_amd.sections import SectionRuntime
    from strut_amd.transient3d import band_solve

    A, band = _spd_band(n, hbw, seed=n + hbw, dominant=dominant)
    rhs = np.random.default_rng(n).normal(size=n)
    ok, want, fallback = oracle.band_solve(band, hbw, rhs)
    assert ok
    rt = SectionRuntime(MaterialTable([{"id": 1, "type": "Elastic", "params": {"E": 1.0}}]), device=0)
    variants = (1, 2, 3, 4) if hbw <= 115 else (1, 2, 3) if hbw <= 119 else (1, 2) if hbw <= 127 else (1,)
    for variant in variants:
        x, info = band_solve(rt, band, rhs, variant=variant)
        print(f"n {n} hbw {hbw} dominant {dominant} variant {variant}: info {info} fallback {fallback} "
              f"err {np.abs(x - want).max() / np.abs(want).max():.2e}", flush=True)
        assert info == (1 if fallback else 0)
        assert np.abs(x - want).max() <= 1e-12 * np.abs(want).max()
    if hbw > 127:
        with pytest.raises(Exception, match="register-window"):
            band_solve(rt, band, rhs, variant=2)
    rt.close()


def test_gpu_frame3d_entry_points_refuse_bad_calls():
    """Host-side refusals of the C ABI (nothing reaches a kernel)."""
    from strut_amd._native import StrutGpuError
    from strut_amd.transient3d import Frame3dDesc

    case = cases.rc_frame3d_case(1, 1, 1, steps=4, dt=0.005, n_core=4, n_bars=4)
    fc = load_frame3d_case(case)
    run = Frame3dRun(fc)
    with pytest.raises(StrutGpuError, match="no recorder"):
        run.envelopes()
    with pytest.raises(StrutGpuError, match="out of range"):
        run.set_recorder([fc.total_dofs], 4)
    # a second frame in the same context
    import ctypes
    d = Frame3dDesc()
    d.n_nodes = 1
    with pytest.raises(StrutGpuError, match="already holds"):
        run.rt._check(run.rt.lib.sgpu_frame3d_create(run.rt.ctx, ctypes.byref(d), None, None))
    # a record naming elements record 0 owns, or ids that are no elements
    from ctypes import POINTER, c_double, c_int32
    from strut_amd._native import dptr, iptr
    ts = [np.zeros(10) for _ in fc.patterns]
    tv = (POINTER(c_double) * 3)(*[dptr(v) for v in ts])
    tn = (c_int32 * 3)(*[10] * len(ts))
    tdt = (c_double * 3)(*[0.005] * len(ts))
    tf = (c_double * 3)(*[1.0] * len(ts))
    out = c_int32()
    for ids, msg in ((np.array(run._keep["ids"], dtype=np.int32), "already used"),
                     (np.full(fc.n_elem, 10**6, dtype=np.int32), "not an sgpu_elem3d element")):
        with pytest.raises(StrutGpuError, match=msg):
            run.rt._check(run.rt.lib.sgpu_frame3d_add_record(run.rt.ctx, iptr(ids), tv, tn, tdt, tf,
                                                             ctypes.byref(out)))
    run.set_recorder([0, 6], 4)
    with pytest.raises(StrutGpuError, match="already set"):
        run.set_recorder([0], 4)
    with pytest.raises(StrutGpuError, match="out of range"):
        run.envelopes([0], [5], [3.0])
    with pytest.raises(StrutGpuError, match="zero denominator"):
        run.envelopes([0], [1], [0.0])
    assert run.run(4)["status"] == 0
    run.close()


def test_gpu_records_run_together_match_each_oracle_run():
    """Three records of one rigid-diaphragm frame in one context (their
    element and band-solve work in the same launches) against the oracle
    run of each record alone: identical iteration counts and statuses,
    displacements within 1e-8 of the peak."""
    from strut_amd.frame3d import Pattern, recorder_plan

    steps = 60
    base = cases.rc_frame3d_case(2, 1, 1, steps=steps, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    fc = load_frame3d_case(base)
    recs, fcs = [], []
    for seeds, pgas in (((5, 6), (0.6, 0.3)), ((7, 8), (0.2, 0.45))):
        c = cases.rc_frame3d_case(2, 1, 1, steps=steps, dt=0.005, n_core=4, n_bars=4, seeds=seeds, pgas_g=pgas)
        f = load_frame3d_case(c)
        fcs.append(f)
        recs.append([Pattern(p.direction, p.values, p.dt, p.factor) for p in f.patterns])
    cols, ci, cj, dx = recorder_plan(fc)
    full = np.arange(fc.total_dofs, dtype=np.int32)
    run = Frame3dRun(fc, records=recs)
    run.set_recorder(full, steps)
    got = run.run_records(steps)
    for r, f in enumerate([fc] + fcs):
        ref = oracle.run_transient3d(f)
        k = ref["steps_done"]
        print(f"record {r}: status {got['status'][r]}/{ref['status']} iters {got['iters'][r].sum()}/"
              f"{ref['iters'].sum()}", flush=True)
        assert got["status"][r] == ref["status"] and got["steps_done"][r] == k
        np.testing.assert_array_equal(got["iters"][r][:k], ref["iters"][:k])
        hist = run.recorded(r)
        scale = np.abs(ref["u"][:k]).max()
        assert np.abs(hist - ref["u"][:k]).max() <= 1e-8 * scale
        fi, fj = cols[ci], cols[cj]  # the drift channels over the full-DOF recording
        _, drift = run.envelopes(fi, fj, dx, record=r)
        _, want = oracle.envelopes(hist, k, fi, fj, dx)
        np.testing.assert_array_equal(drift, want)
    run.close()


def test_gpu_records_that_stop_early_leave_the_others_unchanged():
    """Four records together, two of them scaled until they stop (the
    reference aborts on an element's local failure, and on Newton
    non-convergence).  Every record gives the same bits as the same record
    run alone on the device (statuses, steps done, iteration counts and
    recorded displacements), so a record that stops leaves the live list
    without disturbing the others' lockstep steps, commits or recorder rows.
    Against the oracle's run of each record alone: the same status and stop
    step for all four; identical counts and displacements within 1e-8 for
    the three that do not end in stalled Newton steps.  The record that
    stops on Newton non-convergence (scaled 10x) ends in stalled steps
    whose iterates depend on the last ulp of libm's pow in the reference
    itself, so its stop step and status are its bar."""
    from strut_amd.frame3d import Pattern

    steps = 60
    base = cases.rc_frame3d_case(2, 1, 1, steps=steps, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    fc = load_frame3d_case(base)
    recs, fcs = [], []
    # scaled 30x: an element's local failure at step 42; 10x: Newton
    # non-convergence at step 51 (the oracle's runs, unchanged when the
    # records are scaled by a further 1 +- 1e-13)
    for seeds, pgas, scale in (((5, 6), (0.6, 0.3), 30.0), ((7, 8), (0.2, 0.45), 1.0), ((9, 10), (0.6, 0.3), 10.0)):
        c = cases.rc_frame3d_case(2, 1, 1, steps=steps, dt=0.005, n_core=4, n_bars=4, seeds=seeds, pgas_g=pgas)
        f = load_frame3d_case(c)
        f.patterns = [Pattern(p.direction, p.values, p.dt, p.factor * scale) for p in f.patterns]
        fcs.append(f)
        recs.append(list(f.patterns))
    full = np.arange(fc.total_dofs, dtype=np.int32)
    run = Frame3dRun(fc, records=recs)
    run.set_recorder(full, steps)
    got = run.run_records(steps)
    hists = [run.recorded(r) for r in range(4)]
    run.close()
    stopped = 0
    for r, f in enumerate([fc] + fcs):
        alone = Frame3dRun(f)
        alone.set_recorder(full, steps)
        one = alone.run(steps, record_history=False)
        one_hist = alone.recorded(0)
        alone.close()
        ref = oracle.run_transient3d(f)
        k = ref["steps_done"]
        print(f"record {r}: status {got['status'][r]}/{one['status']}/{ref['status']} steps "
              f"{got['steps_done'][r]}/{one['steps_done']}/{k} iters {got['iters'][r].sum()}/{one['iters'].sum()}/"
              f"{ref['iters'].sum()}", flush=True)
        assert got["status"][r] == one["status"] and got["steps_done"][r] == one["steps_done"]
        np.testing.assert_array_equal(got["iters"][r], one["iters"])
        np.testing.assert_array_equal(hists[r], one_hist)
        assert got["status"][r] == ref["status"] and got["steps_done"][r] == k
        stopped += ref["status"] != 0
        if ref["status"] > 0:
            continue  # ends in stalled Newton steps: the same stop is the bar
        n_it = k  # an element failure's last step has no oracle count
        np.testing.assert_array_equal(got["iters"][r][:n_it], ref["iters"][:n_it])
        if k:
            scale = np.abs(ref["u"][:k]).max()
            assert np.abs(hists[r] - ref["u"][:k]).max() <= 1e-8 * scale
    assert stopped == 2 and got["status"][1] < 0 and got["status"][3] > 0


def _tower(steps, **kw):
    args = dict(dt=0.005, core=(1, 1), wall_h_div=2, wall_v_div=2, n_core=4, n_bars=4, pgas_g=(1.2, 0.84),
                wall_E=3.0e6)
    args.update(kw)
    return cases.rc_tower_case(2, 2, steps=steps, **args)


def test_gpu_tower_effective_band_matches_oracle():
    """A tower with an elastic shell core wall: beam-column and shell
    stiffness, collapse and band of the first Newton solve."""
    fc = load_frame3d_case(_tower(2))
    ref = oracle.run_transient3d(fc, hbw_rule=0, steps=1, capture_band=True)
    dev = Frame3dRun(fc, hbw_rule=0)
    got = dev.effective_band()
    dev.close()
    want = ref["band_first"]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("hbw_rule", [0, 1])
def test_gpu_tower_with_a_shell_wall_matches_oracle(hbw_rule):
    """RC frame around a shell core wall (wall nodes on the floors slaved to
    the diaphragms, wall nodes between floors free), two records, well into
    the frame's nonlinear range."""
    ref = _compare(load_frame3d_case(_tower(100)), hbw_rule, 100)
    assert ref["status"] == 0 and np.abs(ref["u"]).max() > 5e-3


def test_gpu_tower_records_run_together_match_each_oracle_run():
    from strut_amd.frame3d import Pattern

    steps = 50
    fc = load_frame3d_case(_tower(steps))
    fcs, recs = [], []
    for seeds, pgas in (((5, 6), (0.9, 1.3)), ((7, 8), (0.4, 0.5))):
        f = load_frame3d_case(_tower(steps, seeds=seeds, pgas_g=pgas))
        fcs.append(f)
        recs.append([Pattern(p.direction, p.values, p.dt, p.factor) for p in f.patterns])
    run = Frame3dRun(fc, records=recs)
    run.set_recorder(np.arange(fc.total_dofs, dtype=np.int32), steps)
    got = run.run_records(steps)
    for r, f in enumerate([fc] + fcs):
        ref = oracle.run_transient3d(f)
        k = ref["steps_done"]
        assert got["status"][r] == ref["status"] and got["steps_done"][r] == k
        np.testing.assert_array_equal(got["iters"][r][:k], ref["iters"][:k])
        scale = np.abs(ref["u"][:k]).max()
        assert np.abs(run.recorded(r) - ref["u"][:k]).max() <= 1e-8 * scale
    run.close()

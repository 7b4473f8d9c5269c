"""GPU parity of the 3D frame transient (sgpu_frame3d_*, strut_amd/csrc/
frame3d.hip) with the oracle's driver (oracle/transient3d.c) on the same
seeded cases: RC frames with forceBeamColumn3d columns and dispBeamColumn3d
beams, rigid floor diaphragms under the Transformation handler and two
simultaneous horizontal records.  The bar: the same status, the same Newton
iteration count at every step, and every displacement within 1e-8 of the
history's peak.  Where the reference's band LU hands a factorization to its
dense partial-pivot LU the device keeps the band elimination (the frames'
effective tangents are symmetric with a dominant diagonal): the solutions
agree within rounding, which the displacement bar checks.  Both band rules run: the reference's (element free map,
whose truncation of the collapsed tangent costs extra iterations) and the
full collapsed band."""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.frame3d import load_frame3d_case
from strut_amd.transient3d import Frame3dRun

pytestmark = pytest.mark.gpu


def _compare(fc, hbw_rule, steps):
    ref = oracle.run_transient3d(fc, hbw_rule=hbw_rule, steps=steps)
    dev = Frame3dRun(fc, hbw_rule=hbw_rule)
    assert dev.hbw == ref["hbw"] and dev.free_count == ref["free_count"]
    got = dev.run(steps)
    dev.close()
    print(f"rule {hbw_rule}: hbw {dev.hbw} n {dev.free_count} status {got['status']}/{ref['status']} "
          f"iters {got['iters'].sum()}/{ref['iters'].sum()} dense fallbacks {got['dense_fallbacks']}/"
          f"{ref['dense_fallbacks']}", flush=True)
    assert got["status"] == ref["status"]
    assert got["steps_done"] == ref["steps_done"]
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


def test_gpu_frame_without_constraints_matches_oracle():
    case = cases.rc_frame3d_case(2, 1, 1, steps=60, dt=0.005, n_core=4, n_bars=4, diaphragms=False,
                                 bidirectional=False, pgas_g=(0.6, 0.0))
    _compare(load_frame3d_case(case), 0, 60)

"""GPU parity of the device-side recorders (strut_amd/csrc/recorders.hip):
envelopes and drift envelopes computed on the device from the recorded
history equal the oracle's restatement (oracle.envelopes, helpers.mojo:
3113-3186) applied to the same history, bit for bit (min / max / abs of
the same divisions), and agree within 1e-8 of the peak with the oracle's
own transient history.  2D: a batch of two frames x three records with
inter-story drift channels on every column line; 3D: a rigid-diaphragm
frame under two records with the generator's drift and floor envelope
recorders."""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.frame2d import load_frame2d_case
from strut_amd.frame3d import load_frame3d_case, recorder_plan
from strut_amd.transient import FrameBatch
from strut_amd.transient3d import Frame3dRun

pytestmark = pytest.mark.gpu


def _drift_channels(stories, bays, h=3.0):
    col_i, col_j = [], []
    for k in range(1, stories + 1):
        for c in range(bays + 1):
            col_i.append(3 * ((k - 1) * (bays + 1) + c))
            col_j.append(3 * (k * (bays + 1) + c))
    return np.array(col_i), np.array(col_j), np.full(len(col_i), h)


def test_gpu_2d_batch_envelopes():
    frames = [(2, 1), (3, 2)]
    fcs = [load_frame2d_case(cases.rc_frame2d_case(s, b, steps=200, dt=0.01, seed=0, pga_g=0.5, strong=1.5))
           for s, b in frames]
    fb = FrameBatch(fcs[0].materials)
    fids = [fb.add_frame(fc, record_all=True) for fc in fcs]
    inst = {}
    for fid, fc in zip(fids, fcs):
        inst[fid] = [fb.add_motion(fid, cases.ground_motion(seed, fc.dt, 200, 0.3 + 0.2 * seed), fc.dt, 200)
                     for seed in range(3)]
    fb.finalize()
    fb.run(rounds=64)
    for fid, (s, b), fc in zip(fids, frames, fcs):
        ci, cj, dx = _drift_channels(s, b)
        disp, drift = fb.envelopes(fid, ci, cj, dx)
        for k, i in enumerate(inst[fid]):
            r = fb.result(i)
            n = r["steps_done"]
            want_d, want_dr = oracle.envelopes(r["u"], n, ci, cj, dx)
            np.testing.assert_array_equal(disp[k], want_d)
            np.testing.assert_array_equal(drift[k], want_dr)
            print(f"frame {s}x{b} record {k}: {n} steps, peak drift {drift[k][:, 2].max():.3e}", flush=True)
    fb.close()
    # against the oracle's own run of one record
    fc = fcs[1]
    ref_case = cases.rc_frame2d_case(3, 2, steps=200, dt=0.01, seed=0, pga_g=0.3, strong=1.5)
    ref = oracle.run_transient2d(load_frame2d_case(ref_case))
    ci, cj, dx = _drift_channels(3, 2)
    _, want = oracle.envelopes(ref["u"], ref["steps_done"], ci, cj, dx)
    fb = FrameBatch(fc.materials)
    fid = fb.add_frame(load_frame2d_case(ref_case), record_all=True)
    fb.add_motion(fid, ref_case["time_series"][0]["values"], 0.01, 200)
    fb.finalize()
    fb.run(rounds=64)
    _, got = fb.envelopes(fid, ci, cj, dx)
    fb.close()
    assert np.abs(got[0] - want).max() <= 1e-8 * np.abs(want).max()


def test_gpu_3d_recorder_and_envelopes():
    case = cases.rc_frame3d_case(2, 2, 1, steps=80, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    fc = load_frame3d_case(case)
    cols, ci, cj, dx = recorder_plan(fc)
    run = Frame3dRun(fc)
    run.set_recorder(cols, 80)
    r = run.run(80)
    assert r["status"] == 0
    hist = run.recorded()
    np.testing.assert_array_equal(hist, r["u"][:, cols])
    disp, drift = run.envelopes(ci, cj, dx)
    run.close()
    want_d, want_dr = oracle.envelopes(hist, len(hist), ci, cj, dx)
    np.testing.assert_array_equal(disp, want_d)
    np.testing.assert_array_equal(drift, want_dr)
    ref = oracle.run_transient3d(fc)
    ref_d, ref_dr = oracle.envelopes(ref["u"][:, cols], ref["steps_done"], ci, cj, dx)
    assert np.abs(drift - ref_dr).max() <= 1e-8 * np.abs(ref_dr).max()
    assert np.abs(disp - ref_d).max() <= 1e-8 * np.abs(ref_d).max()
    print(f"3D: peak drifts {drift[:, 2]}", flush=True)

"""Pins the oracle's recorder restatement (oracle.envelopes:
_drift_value, helpers.mojo:3113-3142; _update_envelope, :3153-3186) by
closed forms: a sampled sine's extremes, the first step seeding the
envelope (an increasing positive series keeps its first value as the
minimum), a rigid translation's zero drift and a linear profile's exact
slope; and against the independent peak-drift formula of the case
generator on a 2D history."""

import numpy as np

from oracle import oracle
from strut_amd import cases
from strut_amd.frame3d import load_frame3d_case, recorder_plan


def test_envelope_closed_forms():
    t = np.arange(400) * 0.01
    hist = np.stack([2.5 * np.sin(2 * np.pi * t), 1.0 + t, -3.0 - t], axis=1)  # sine hits +-2.5 at t = 0.25, 0.75
    disp, drift = oracle.envelopes(hist, 400)
    np.testing.assert_allclose(disp[0], [-2.5, 2.5, 2.5], rtol=0, atol=1e-12)
    assert disp[1].tolist() == [1.0, hist[399, 1], hist[399, 1]]  # seeded by the first value, not zero
    assert disp[2].tolist() == [hist[399, 2], -3.0, abs(hist[399, 2])]
    assert drift.shape == (0, 3)
    one, _ = oracle.envelopes(hist, 1)
    assert one[2].tolist() == [-3.0, -3.0, 3.0]
    none, _ = oracle.envelopes(hist, 0)
    assert not none.any()


def test_drift_of_rigid_translation_and_linear_profile():
    rng = np.random.default_rng(0)
    base = rng.normal(size=200)
    hist = np.stack([base, base, base + 0.75 * 4.0], axis=1)  # columns 0, 1 move together; 2 leads by 0.75 * 4 m
    _, drift = oracle.envelopes(hist, 200, col_i=[0, 0], col_j=[1, 2], dx=[3.0, 4.0])
    assert drift[0].tolist() == [0.0, 0.0, 0.0]
    np.testing.assert_allclose(drift[1], [0.75, 0.75, 0.75], rtol=1e-15)


def test_drift_envelope_matches_the_peak_drift_formula():
    from strut_amd.frame2d import load_frame2d_case

    case = cases.rc_frame2d_case(3, 2, steps=150, dt=0.01, seed=4, pga_g=0.4, strong=1.0)
    fc2 = load_frame2d_case(case)
    r = oracle.run_transient2d(fc2)
    u = r["u"][:r["steps_done"]]
    cols = 3
    col_i, col_j = [], []
    for k in range(1, 4):
        for c in range(cols):
            col_i.append(3 * ((k - 1) * cols + c))
            col_j.append(3 * (k * cols + c))
    _, drift = oracle.envelopes(u, len(u), col_i, col_j, [3.0] * len(col_i))
    peak = drift[:, 2].reshape(3, cols).max(axis=1)
    np.testing.assert_allclose(peak, cases.peak_story_drift(u, 3, 2), rtol=1e-14)


def test_3d_loader_reads_the_drift_and_envelope_recorders():
    fc = load_frame3d_case(cases.rc_frame3d_case(2, 1, 1, steps=10, n_core=4, n_bars=4))
    assert len(fc.drift) == 4 and all(d == 3.0 for _, _, d in fc.drift)
    cols, ci, cj, dx = recorder_plan(fc)
    assert len(cols) == len(fc.envelope_dofs) + 6  # three corner nodes x (ux, uy); the base's are fixed DOFs
    for (i_dof, j_dof, _), a, b in zip(fc.drift, ci, cj):
        assert cols[a] == i_dof and cols[b] == j_dof

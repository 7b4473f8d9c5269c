"""The seeded ground motions: the Butterworth band-pass design against its
closed-form magnitude, the zero-phase filter's properties, and records of
any length (no minimum, no numerical library in the generator)."""

import math

import numpy as np
import pytest

from strut_amd import cases


def _response(sos, w):
    z = np.exp(1j * w)
    h = np.ones_like(w, dtype=complex)
    for b0, b1, b2, a1, a2 in sos:
        h *= (b0 + b1 / z + b2 / z ** 2) / (1.0 + a1 / z + a2 / z ** 2)
    return h


@pytest.mark.parametrize("f_lo,f_hi,fs", [(0.1, 25.0, 100.0), (0.1, 2.0, 100.0), (0.5, 15.0, 200.0)])
def test_band_pass_magnitude_is_the_butterworth_closed_form(f_lo, f_hi, fs):
    """The bilinear transform maps the digital frequency w to the analog
    W = 2 fs tan(w / 2), so |H(e^iw)|^2 = 1 / (1 + ((W^2 - W1 W2) /
    ((W2 - W1) W))^(2N)) with pre-warped edges W1, W2: -3 dB exactly at
    f_lo and f_hi, 1 at the geometric centre, zeros at DC and Nyquist."""
    n = 4
    sos = cases.butterworth_bandpass_sos(n, f_lo, f_hi, fs)
    assert len(sos) == n
    w = np.linspace(1e-4, math.pi - 1e-4, 4001)
    big_w = 2.0 * fs * np.tan(w / 2.0)
    w1, w2 = (2.0 * fs * math.tan(math.pi * f / fs) for f in (f_lo, f_hi))
    want = 1.0 / (1.0 + ((big_w ** 2 - w1 * w2) / ((w2 - w1) * big_w)) ** (2 * n))
    got = np.abs(_response(sos, w)) ** 2
    assert np.abs(got - want).max() < 1e-9
    edges = np.array([2.0 * math.pi * f_lo / fs, 2.0 * math.pi * f_hi / fs])
    np.testing.assert_allclose(np.abs(_response(sos, edges)) ** 2, 0.5, rtol=1e-9)
    assert abs(_response(sos, np.array([1e-9]))[0]) < 1e-8
    for b0, b1, b2, a1, a2 in sos:  # stable: every pole pair inside the unit circle
        assert 0.0 < a2 < 1.0 and abs(a1) < 1.0 + a2


def test_zero_phase_filter_passes_a_mid_band_sine_unchanged():
    fs, f = 100.0, 3.0
    sos = cases.butterworth_bandpass_sos(4, 0.1, 25.0, fs)
    t = np.arange(16000) / fs
    x = np.sin(2.0 * math.pi * f * t)
    y = np.array(cases.sos_filtfilt(sos, x))
    gain = np.abs(_response(sos, np.array([2.0 * math.pi * f / fs])))[0] ** 2
    # 70 s from either end the 0.1 Hz edge's start-up transient has decayed:
    # no phase shift, the squared gain
    mid = slice(7000, 9000)
    assert np.abs(y[mid] - gain * x[mid]).max() < 1e-6


def test_filter_matches_scipy_where_scipy_is_present():
    signal = pytest.importorskip("scipy.signal")
    x = cases.gaussian_noise(3, 3001)
    for f_lo, f_hi, fs in [(0.1, 25.0, 100.0), (0.1, 2.0, 100.0)]:
        ref = signal.sosfiltfilt(signal.butter(4, [f_lo, f_hi], btype="bandpass", fs=fs, output="sos"), x)
        got = np.array(cases.sos_filtfilt(cases.butterworth_bandpass_sos(4, f_lo, f_hi, fs), x))
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("steps", [1, 2, 5, 99, 1000])
def test_ground_motion_any_length(steps):
    a = cases.ground_motion(7, 0.01, steps, 0.4)
    assert len(a) == steps + 1 and np.isfinite(a).all()
    assert abs(np.abs(a).max() - 0.4 * 9.81) < 1e-12 or steps < 2
    np.testing.assert_array_equal(a, cases.ground_motion(7, 0.01, steps, 0.4))

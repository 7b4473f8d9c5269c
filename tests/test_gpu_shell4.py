"""GPU parity of the elastic MITC4 shell stiffness (csrc/shell4.hip,
sgpu_shell4_stiffness) against the oracle's restatement of
shell4_mindlin_stiffness (oracle/shell4.c), through the C ABI.

Every expression keeps the reference's order on both sides, so the entries
agree to the last bits of atan2 / sin / cos (device and host libraries may
round those differently): the bar is 1e-13 of the largest entry.  Shapes:
rectangles, trapezoids, general and warped quads in random orientations,
several sections in one launch."""

import numpy as np
import pytest

from oracle import oracle
from strut_amd.shells import shell4_stiffness

pytestmark = pytest.mark.gpu


def _shells(n, seed=5):
    rng = np.random.default_rng(seed)
    xyz, props = [], []
    base = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [2.0, 1.5, 0.0], [0.0, 1.5, 0.0]])
    for k in range(n):
        p = base + rng.uniform(-0.35, 0.35, size=(4, 3)) * np.array([1.0, 1.0, 0.3 * (k % 3)])
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        xyz.append(p @ q.T * rng.uniform(0.5, 3.0) + rng.normal(size=3) * 10.0)
        props.append([rng.uniform(2.0e7, 4.0e7), rng.uniform(0.0, 0.35), rng.uniform(0.1, 0.6)])
    return np.array(xyz), np.array(props)


@pytest.mark.parametrize("n", [1, 7, 300])
def test_gpu_shell4_stiffness_matches_oracle(n):
    xyz, props = _shells(n)
    got = shell4_stiffness(xyz, props)
    for e in range(n):
        want = oracle.shell4_stiffness(*props[e], xyz[e])
        err = np.abs(got[e] - want).max() / np.abs(want).max()
        assert err < 1e-13, f"shell {e}: {err:.2e}"


def test_gpu_shell4_axis_aligned_rectangle_is_oracle_bits():
    """An axis-aligned rectangle takes atan2 / sin / cos at exact angles, so
    the device entries are the oracle's bits."""
    xyz = np.array([[[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [2.0, 1.5, 0.0], [0.0, 1.5, 0.0]]])
    props = np.array([[3.0e7, 0.2, 0.4]])
    got = shell4_stiffness(xyz, props)[0]
    want = oracle.shell4_stiffness(3.0e7, 0.2, 0.4, xyz[0])
    assert np.array_equal(got, want)


def test_gpu_shell4_refuses_degenerate_geometry():
    from strut_amd._native import StrutGpuError
    xyz = np.zeros((2, 4, 3))
    xyz[0] = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]]
    with pytest.raises(StrutGpuError, match="shell 1"):
        shell4_stiffness(xyz, np.array([[3.0e7, 0.2, 0.4]] * 2))

"""GPU parity of the modal analysis (sgpu_frame3d_modal, csrc/modal.hip)
with the oracle's restatement of run_modal_eigen (or_tr3d_modal) on 3D
frames with rigid diaphragms and shell walls.

  * Run to convergence, the eigenvalues agree within 1e-10 and the modes
    solve the generalized problem of the collapsed system
    (tests/test_oracle_modal.py builds it) within the same bar as the
    oracle's.
  * With the reference's 100-rotation cap, on a model whose masses are
    uneven (so no two off-diagonal entries tie by symmetry and both sides
    rotate the same pairs), the unconverged eigenvalues and modes agree with
    the oracle's within 1e-9 and the rotation counts are equal."""

import copy

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.frame3d import load_frame3d_case
from strut_amd.transient3d import Frame3dRun

from test_oracle_modal import _system

pytestmark = pytest.mark.gpu


def _uneven(case, seed=3):
    c = copy.deepcopy(case)
    rng = np.random.default_rng(seed)
    for m in c["masses"]:
        m["value"] *= float(rng.uniform(0.6, 1.4))
    return c


MODELS = {
    "frame": lambda: cases.rc_frame3d_case(2, 2, 1, steps=2, n_core=4, n_bars=4),
    "tower": lambda: cases.rc_tower_case(2, 2, steps=2, core=(1, 1), wall_h_div=2, wall_v_div=2, n_core=4,
                                         n_bars=4),
}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_gpu_modal_converged_matches_oracle(name):
    fc = load_frame3d_case(MODELS[name]())
    Kr, Mr, _, free = _system(fc)
    n = len(free)
    want, _, rot_o = oracle.modal3d(fc, 6, max_rotations=40 * n * n)
    run = Frame3dRun(fc)
    got, modes, rot = run.modal(6, max_rotations=40 * n * n)
    run.close()
    print(f"{name}: n {n} rotations {rot}/{rot_o} eig {got[:3]} / {want[:3]}", flush=True)
    np.testing.assert_allclose(got, want, rtol=1e-10)
    for lam, phi in zip(got, modes):
        pf = phi[free]
        assert np.abs(Kr @ pf - lam * (Mr @ pf)).max() <= 1e-7 * np.abs(Kr @ pf).max()
    Phi = modes[:, free].T
    np.testing.assert_allclose(Phi.T @ Mr @ Phi, np.eye(6), atol=1e-9)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_gpu_modal_with_the_reference_rotation_cap_matches_oracle(name):
    fc = load_frame3d_case(_uneven(MODELS[name]()))
    want, want_modes, rot_o = oracle.modal3d(fc, 5)
    run = Frame3dRun(fc)
    got, modes, rot = run.modal(5)
    run.close()
    assert rot == rot_o == 100
    np.testing.assert_allclose(got, want, rtol=1e-9)
    assert np.abs(modes - want_modes).max() <= 1e-9 * np.abs(want_modes).max()

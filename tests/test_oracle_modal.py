"""Pins the oracle's modal analysis (or_tr3d_modal, a restatement of
run_modal_eigen, modal_eigen.mojo:143-362, with jacobi_symmetric_eigen,
eigen.mojo:15-79) on 3D frames with rigid diaphragms and shell walls.

  * Run to convergence (enough Jacobi rotations), its eigenvalues are the
    lowest generalized eigenvalues of (T^T K T, M_ff) from scipy, with K the
    stiffness at rest of the oracle's elements and shells (pinned by their
    own tests), T the diaphragm expansion and M_ff the reduced mass with the
    reference's floors (a zero diagonal gets K_ii 1e-10); its modes solve
    K phi = lambda M phi.
  * After any number of rotations (the reference stops at 100), the modes
    are M-orthonormal and phi^T K phi is the returned eigenvalue: the
    rotations are orthogonal similarity transforms of L^-1 K L^-T, whatever
    pair each one picks, and the returned values are that matrix's diagonal.
  * With the reference's 100-rotation cap a model with more than a few
    dozen free DOFs is far from converged (the reference's limit, kept):
    the cap is checked to bind.
"""

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle
from strut_amd import cases
from strut_amd.constraints import expansion_matrix
from strut_amd.frame3d import load_frame3d_case


def _system(fc):
    _, elems = oracle.frame3d_elements(fc)
    nt = fc.total_dofs
    K = np.zeros((nt, nt))
    for el, (i, j) in zip(elems, fc.elem_nodes):
        k, _ = el.state(np.zeros(12))
        dofs = list(range(6 * i, 6 * i + 6)) + list(range(6 * j, 6 * j + 6))
        K[np.ix_(dofs, dofs)] += k
    for nodes, props in zip(fc.shell_nodes, fc.shell_props):
        dofs = [6 * nd + k for nd in nodes for k in range(6)]
        K[np.ix_(dofs, dofs)] += oracle.shell4_stiffness(*props, fc.coords[nodes])
    free = np.flatnonzero(~fc.constrained)
    T = expansion_matrix(fc.row_offsets, fc.dof_pool, fc.coeff_pool)[:, free]
    Kr, Mr = T.T @ K @ T, T.T @ np.diag(fc.mass) @ T
    d = np.diag(Mr).copy()
    floor = max(d.max() * 1e-12, 1e-12)
    d = np.where(d == 0.0, np.diag(Kr) * 1e-10, d)
    d = np.where(d <= 0.0, floor, d)
    Mr[np.diag_indices_from(Mr)] = d
    return Kr, Mr, T, free


CASES = {
    "frame": lambda: cases.rc_frame3d_case(1, 1, 1, steps=2, n_core=4, n_bars=4),
    "tower": lambda: cases.rc_tower_case(1, 1, steps=2, core=(1, 1), wall_h_div=1, wall_v_div=2, n_core=4,
                                         n_bars=4),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_modal_converges_to_the_generalized_eigenvalues(name):
    fc = load_frame3d_case(CASES[name]())
    Kr, Mr, T, free = _system(fc)
    n = len(free)
    eig, modes, rot = oracle.modal3d(fc, 4, max_rotations=200 * n * n, tol=1e-12)
    assert rot < 200 * n * n  # stopped by the tolerance
    want = scipy.linalg.eigh(Kr, Mr, eigvals_only=True)[:4]
    np.testing.assert_allclose(eig, want, rtol=1e-8)
    for lam, phi in zip(eig, modes):
        pf = phi[free]
        r = Kr @ pf - lam * (Mr @ pf)
        assert np.abs(r).max() <= 1e-7 * np.abs(Kr @ pf).max()
        np.testing.assert_allclose(T @ pf, phi, atol=1e-12 * np.abs(phi).max())  # slaves follow the floors


@pytest.mark.parametrize("rotations", [0, 1, 2, 7, 100])
def test_every_rotation_keeps_an_orthogonal_similarity(rotations):
    fc = load_frame3d_case(CASES["tower"]())
    Kr, Mr, T, free = _system(fc)
    n = len(free)
    eig, modes, rot = oracle.modal3d(fc, n, max_rotations=rotations)
    assert rot == min(rotations, rot) and rot <= rotations
    Phi = modes[:, free].T
    np.testing.assert_allclose(Phi.T @ Mr @ Phi, np.eye(n), atol=1e-9)
    kk = Phi.T @ Kr @ Phi
    np.testing.assert_allclose(np.diag(kk), eig, rtol=1e-9, atol=1e-9 * np.abs(eig).max())
    assert np.all(np.diff(eig) >= 0.0)  # ascending


def test_the_reference_rotation_cap_binds_on_a_larger_model():
    fc = load_frame3d_case(cases.rc_frame3d_case(2, 2, 1, steps=2, n_core=4, n_bars=4))
    Kr, Mr, _, free = _system(fc)
    assert len(free) > 40
    eig, _, rot = oracle.modal3d(fc, 3)
    assert rot == 100
    want = scipy.linalg.eigh(Kr, Mr, eigvals_only=True)[:3]
    assert np.abs(eig - want).max() > 1e-3 * want.max()

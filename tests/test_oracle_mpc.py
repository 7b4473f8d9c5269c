"""Pins the oracle's Transformation-handler collapses (oracle/mpc.c:
helpers.mojo:3288-3397) and the host's table build (strut_amd.constraints,
loader.mojo:2230-2372) before they are trusted:
  * collapse_matrix = T^T K T, collapse_vector = T^T F, with T the expansion
    matrix of the tables (numpy, independent of the loop order);
  * rigid floor kinematics: a floor retained node moved by (ux, uy, rz) puts
    every floor node at (ux - dy rz, uy + dx rz, rz) -- the closed form of
    a rigid in-plane motion, which fixes the signs of the diaphragm
    coefficients from the geometry;
  * the collapsed system keeps the stiffness of a rigid-floor model: a
    uniform floor translation collapsed is the sum of the floor's rows.
"""

import numpy as np

from oracle import oracle
from strut_amd import cases
from strut_amd.constraints import expansion_matrix


def _model():
    return cases.frame3d_diaphragm_model(stories=2, bays_x=2, bays_y=1)


def test_collapse_is_tt_k_t():
    m = _model()
    tables = m["tables"][:3]
    K = cases.synthetic_frame_stiffness(m, seed=3)
    T = expansion_matrix(*tables)
    want = T.T @ K @ T
    got = oracle.mpc_collapse_matrix(K, tables)
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
    F = np.random.default_rng(4).normal(size=K.shape[0])
    np.testing.assert_allclose(oracle.mpc_collapse_vector(F, tables), T.T @ F, rtol=1e-13, atol=1e-12)


def test_rigid_floor_kinematics():
    m = _model()
    ro, dp, cp, slave, constrained = m["tables"]
    n = len(ro) - 1
    u = np.zeros(n)
    ni = m["node_index"]
    for k, mid in m["masters"].items():
        b = ni[mid] * 6
        u[b + 0], u[b + 1], u[b + 5] = 0.01 * k, -0.02 * k, 0.003 * k
    got = oracle.mpc_enforce(u, (ro, dp, cp), constrained, slave)
    for k, mid in m["masters"].items():
        xm, ym, _ = m["coords"][mid]
        b = ni[mid] * 6
        for nid, (x, y, z) in m["coords"].items():
            if abs(z - 3.0 * k) > 1e-12 or nid == mid:
                continue
            c = ni[nid] * 6
            dx, dy = x - xm, y - ym
            assert abs(got[c] - (u[b] - dy * u[b + 5])) <= 1e-15
            assert abs(got[c + 1] - (u[b + 1] + dx * u[b + 5])) <= 1e-15
            assert got[c + 5] == u[b + 5]
    # single-point fixed DOFs (the base) are zeroed
    assert np.all(got[constrained & ~slave] == 0.0)


def test_uniform_floor_translation_collapses_to_floor_sum():
    m = _model()
    ro, dp, cp, slave, constrained = m["tables"]
    K = cases.synthetic_frame_stiffness(m, seed=5)
    Kc = oracle.mpc_collapse_matrix(K, (ro, dp, cp))
    ni = m["node_index"]
    mid = m["masters"][1]
    floor = [ni[n] * 6 for n, (x, y, z) in m["coords"].items() if abs(z - 3.0) < 1e-12 and n != mid]
    # row ux of the retained node = sum of the floor nodes' ux rows (ux columns)
    want = sum(K[a, b] for a in floor for b in floor)
    got = Kc[ni[mid] * 6, ni[mid] * 6]
    assert abs(got - want) <= 1e-12 * abs(want)

"""Pins the oracle's 3D fiber beam-column elements (forceBeamColumn3d,
force_beam_column3d.mojo:2549-2990; dispBeamColumn3d,
beam_column3d_nonlinear.mojo:357-575) before they are trusted as the checker.

  * closed forms: an elastic prismatic element (every fiber Elastic) has the
    textbook 12 x 12 frame stiffness (EA/L, 12EI/L^3, 6EI/L^2, 4EI/L,
    2EI/L, GJ/L) for both formulations, rotated by the element axes of
    elements/utils.mojo:77-139 (checked against hand-built axes), plus the
    P-Delta terms N/L;
  * invariants of the force-based element after nonlinear steps: section
    forces equal b(xi) q (equilibrium, within the element tolerance) and
    the basic deformations equal sum w L b^T e (compatibility);
  * the force-based 3D element bent in one plane against the 2D one
    (force_beam_column2d.mojo, pinned separately) on the same fibers;
  * the Gauss-Jordan 5 x 5 inverse against numpy.
A planted error in the element's b(xi) or in a transformation fails them.
"""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.model import MaterialTable, beam_integration_rule, discretize_section

XI, W = (np.array(a) for a in beam_integration_rule("Lobatto", 5))


def _elastic_section(E=30.0e6, b=0.3, h=0.5, ny=10, nz=6):
    mats = MaterialTable([{"id": 9, "type": "Elastic", "params": {"E": E}}])
    sec = {"id": 1, "type": "FiberSection3d", "params": {"patches": [
        {"type": "rect", "material": 9, "num_subdiv_y": ny, "num_subdiv_z": nz,
         "y_i": -h / 2, "z_i": -b / 2, "y_j": h / 2, "z_j": b / 2}]}}
    cells = discretize_section(sec, mats)
    return mats, cells


def _props(cells, E):
    A = float(np.sum(cells["area"]))
    Iz = float(np.sum(cells["area"] * cells["y"] ** 2))  # bending in the local x-y plane
    Iy = float(np.sum(cells["area"] * cells["z"] ** 2))
    return E * A, E * Iz, E * Iy


def _frame_stiffness_local(L, EA, EIz, EIy, GJ):
    """Textbook 12 x 12 elastic frame stiffness, DOFs (u, v, w, tx, ty, tz) at
    each end, local axes."""
    k = np.zeros((12, 12))
    a = EA / L
    k[np.ix_([0, 6], [0, 6])] = a * np.array([[1, -1], [-1, 1]])
    t = GJ / L
    k[np.ix_([3, 9], [3, 9])] = t * np.array([[1, -1], [-1, 1]])
    # bending in x-y (v, tz): EIz
    c = EIz / L ** 3
    idx = [1, 5, 7, 11]
    k[np.ix_(idx, idx)] = c * np.array([[12, 6 * L, -12, 6 * L], [6 * L, 4 * L * L, -6 * L, 2 * L * L],
                                        [-12, -6 * L, 12, -6 * L], [6 * L, 2 * L * L, -6 * L, 4 * L * L]])
    # bending in x-z (w, ty): EIy, opposite rotation sign convention
    c = EIy / L ** 3
    idx = [2, 4, 8, 10]
    k[np.ix_(idx, idx)] = c * np.array([[12, -6 * L, -12, -6 * L], [-6 * L, 4 * L * L, 6 * L, 2 * L * L],
                                        [-12, 6 * L, 12, 6 * L], [-6 * L, 2 * L * L, 6 * L, 4 * L * L]])
    return k


def _T(R):
    T = np.zeros((12, 12))
    for b in range(4):
        T[3 * b:3 * b + 3, 3 * b:3 * b + 3] = R
    return T


def _axes(x1, x2, vecxz):
    lx = np.asarray(x2, float) - np.asarray(x1, float)
    lx /= np.linalg.norm(lx)
    y = np.cross(vecxz, lx)
    y /= np.linalg.norm(y)
    return np.array([lx, y, np.cross(lx, y)])


@pytest.mark.parametrize("kind", ["fbc", "dbc"])
@pytest.mark.parametrize("geom", [((0, 0, 0), (0, 0, 3.0), None), ((1.0, 2.0, 0.5), (5.0, 2.5, 1.5), (0.3, 0.2, 1.0)),
                                  ((0, 0, 0), (6.0, 0, 0), None)])
def test_elastic_element_is_textbook_frame_stiffness(kind, geom):
    E, G, J = 30.0e6, 12.5e6, 2.1e-3
    mats, cells = _elastic_section(E)
    t, p = mats.arrays()
    sec = oracle.Section3d(cells, t, p)
    sec.alloc(5)
    x1, x2, vz = geom
    el = oracle.Elem3d(kind, sec, 0, XI, W, x1, x2, vecxz=vz, G=G, J=J)
    rng = np.random.default_rng(3)
    u = rng.normal(scale=1e-4, size=12)
    k, f = el.state(u)
    L = float(np.linalg.norm(np.subtract(x2, x1)))
    EA, EIz, EIy = _props(cells, E)
    R = oracle.beam3d_rotation(x1, x2, vz)
    default = (1.0, 0.0, 0.0) if abs((np.subtract(x2, x1) / L)[0]) < 0.9 else (0.0, 1.0, 0.0)
    np.testing.assert_allclose(R, _axes(x1, x2, vz if vz is not None else default), atol=1e-15)
    kl = _frame_stiffness_local(L, EA, EIz, EIy, G * J)
    kg = _T(R).T @ kl @ _T(R)
    scale = np.abs(kg).max()
    assert np.abs(k - kg).max() <= 1e-9 * scale
    np.testing.assert_allclose(f, kg @ u, rtol=0, atol=1e-9 * np.abs(kg @ u).max())


def test_pdelta_adds_axial_force_over_length():
    E = 30.0e6
    mats, cells = _elastic_section(E)
    t, p = mats.arrays()
    L = 3.0
    out = {}
    for transf in ("Linear", "PDelta"):
        sec = oracle.Section3d(cells, t, p)
        sec.alloc(5)
        el = oracle.Elem3d("dbc", sec, 0, XI, W, (0, 0, 0), (L, 0, 0), vecxz=(0, 0, 1), transf=transf)
        u = np.zeros(12)
        u[6] = -1e-4  # axial shortening only
        out[transf] = el.state(u)
    EA, _, _ = _props(cells, E)
    N = EA * (-1e-4) / L
    d = out["PDelta"][0] - out["Linear"][0]
    want = np.zeros((12, 12))
    for i, j in ((1, 7), (2, 8)):
        want[i, i] = want[j, j] = N / L
        want[i, j] = want[j, i] = -N / L
    np.testing.assert_allclose(d, want, atol=1e-9 * abs(N / L))


def _rc3d():
    mats = MaterialTable(cases.rc_materials())
    sec = cases.rc_section(1, dim=3, b=0.4, h=0.4, n_core=8, n_bars=8)
    return mats, discretize_section(sec, mats)


def test_force_based_element_equilibrium_and_compatibility():
    mats, cells = _rc3d()
    t, p = mats.arrays()
    sec = oracle.Section3d(cells, t, p)
    sec.alloc(5)
    L = 3.0
    el = oracle.Elem3d("fbc", sec, 0, XI, W, (0, 0, 0), (0, 0, L), G=1.0e7, J=2e-3, transf="PDelta")
    R = oracle.beam3d_rotation((0, 0, 0), (0, 0, L))
    k0, _ = el.state(np.zeros(12))
    for step in range(1, 9):
        a = step / 8.0
        u = np.zeros(12)
        u[6] = 0.03 * a   # top drift x
        u[7] = -0.02 * a  # top drift y
        u[8] = -0.002 * a  # shortening
        u[11] = 0.004 * a  # twist
        k, _ = el.state(u)
        q = el.q
        nip = 5
        e0, ky, kz = q[5:5 + nip], q[5 + nip:5 + 2 * nip], q[5 + 2 * nip:5 + 3 * nip]
        for ip in range(nip):
            xi = XI[ip]
            r = sec.set_trial(ip, e0[ip], ky[ip], kz[ip])
            N, My, Mz = r[0], r[1], r[2]
            scale = max(abs(q[0]), abs(q[1]), abs(q[3]), 1.0)
            assert abs(N - q[0]) <= 1e-6 * scale
            assert abs(Mz - ((xi - 1.0) * q[1] + xi * q[2])) <= 1e-6 * scale
            assert abs(My - ((xi - 1.0) * q[3] + xi * q[4])) <= 1e-6 * scale
        ul = _T(R) @ u
        cz, cy = (ul[7] - ul[1]) / L, (ul[8] - ul[2]) / L
        vb = np.array([ul[6] - ul[0], ul[5] - cz, ul[11] - cz, ul[4] + cy, ul[10] + cy])
        wl = W * L
        v = np.array([np.sum(wl * e0), np.sum(wl * (XI - 1) * kz), np.sum(wl * XI * kz), np.sum(wl * (XI - 1) * ky),
                      np.sum(wl * XI * ky)])
        np.testing.assert_allclose(v, vb, rtol=0, atol=1e-10)
        el.commit()
    # the run went far into the nonlinear range: the lateral tangent has
    # softened to a fraction of the initial one
    assert k[6, 6] < 0.2 * k0[6, 6]


def test_force_based_3d_in_plane_matches_2d():
    mats = MaterialTable(cases.rc_materials())
    s3 = cases.rc_section(1, dim=3, b=0.4, h=0.4, n_core=8, n_bars=8)
    c3 = discretize_section(s3, mats)
    s2 = cases.rc_section(1, dim=2, b=0.4, h=0.4, n_core=8, n_bars=8)
    c2 = discretize_section(s2, mats)
    t, p = mats.arrays()
    sec3 = oracle.Section3d(c3, t, p)
    sec3.alloc(5)
    sec2 = oracle.Section2d(c2, t, p)
    sec2.alloc(5)
    L = 3.0
    # local axes = global axes: x along the element, vecxz = +z
    el3 = oracle.Elem3d("fbc", sec3, 0, XI, W, (0, 0, 0), (L, 0, 0), vecxz=(0, 0, 1))
    q2 = np.zeros(3 + 2 * 5 + 3)
    for step in range(1, 7):
        a = step / 6.0
        u2 = np.array([0.0, 0.0, 0.0, -0.0015 * a, 0.025 * a, 0.004 * a])
        u3 = np.zeros(12)
        u3[[6, 7, 11]] = u2[[3, 4, 5]]
        k2, f2 = oracle.fbc2d_eval(sec2, 0, XI, W, 0.0, 0.0, L, 0.0, q2, u2)
        k3, f3 = el3.state(u3)
        idx = [0, 1, 5, 6, 7, 11]
        scale_k = np.abs(k2).max()
        assert np.abs(k3[np.ix_(idx, idx)] - k2).max() <= 1e-7 * scale_k
        assert np.abs(f3[idx] - f2).max() <= 1e-7 * np.abs(f2).max()
        # nothing out of plane
        assert np.abs(f3[[2, 3, 4, 8, 9, 10]]).max() <= 1e-7 * np.abs(f2).max()
        sec2.commit_all()
        el3.commit()


def test_invert5_matches_numpy():
    rng = np.random.default_rng(11)
    for _ in range(20):
        a = rng.normal(size=(5, 5)) + 5 * np.eye(5)
        ok, inv = oracle.invert5(a)
        assert ok
        np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=1e-12, atol=1e-14)
    ok, _ = oracle.invert5(np.zeros((5, 5)))
    assert not ok

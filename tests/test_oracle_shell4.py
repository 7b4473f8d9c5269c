"""Pins the oracle's elastic MITC4 shell (oracle/shell4.c, a restatement of
shell4_mindlin_stiffness, shell4.mojo:10-330) by closed forms of plate and
membrane theory, and shows that each check catches a plausible slip.

For a flat four-node element of area A with section (E, nu, h) and any
convex shape, the strain states below are reproduced exactly by the
bilinear membrane, the bilinear rotations and the MITC4 assumed shear, and
2 x 2 Gauss integrates their energies exactly, so u^T K u must equal:
  * membrane, u = ex x + g/2 y, v = g/2 x + ey y:  h A e^T Dm e with
    Dm = E / (1 - nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]],
    e = (ex, ey, g);
  * bending (Kirchhoff field w = kx x^2 / 2 + ky y^2 / 2 + kt x y, rotations
    rx = dw/dy, ry = -dw/dx, so the tied edge shears vanish):
    A k^T Db k with Db = E h^3 / (12 (1 - nu^2)) [...same pattern...],
    k = (kx, ky, 2 kt);
  * transverse shear (w = gx x + gy y, no rotation): A (5/6) G h (gx^2 + gy^2)
    on parallelograms (the reference's assumed shear takes its edge lengths
    and the rotation Rot at the element centre, shell4.mojo:230-246, which is
    exact for constant shear only when opposite edges are parallel; on other
    shapes the shear terms are pinned by the Kirchhoff states, whose tied
    shears must vanish, by the rigid-body motions and by invariance);
  * drilling (constant rz = c, nothing else): A G h c^2;
and the six rigid-body motions give zero force.  Every bilinear form
u1^T K u2 between two states of one kind follows the same closed form, and
the element is invariant under rotations of space: K(Q x) = Qb K(x) Qb^T.

The checks take the library they drive, so they also run against copies of
the oracle with one slip planted in shell4.c, compiled here with gcc: every
planted slip must fail a check.
"""

import ctypes
import os
import shutil
import subprocess
from ctypes import POINTER, c_double, c_int

import numpy as np
import pytest

from oracle import oracle

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

E, NU, H = 30.0e6, 0.2, 0.4
G = E / (2.0 * (1.0 + NU))
QUADS = {
    "rectangle": [[0.0, 0.0], [2.0, 0.0], [2.0, 1.5], [0.0, 1.5]],
    "trapezoid": [[0.0, 0.0], [3.0, 0.0], [2.2, 1.4], [0.6, 1.4]],
    "general": [[0.1, -0.2], [2.4, 0.1], [2.0, 1.9], [-0.3, 1.3]],
    "parallelogram": [[0.0, 0.0], [2.5, 0.4], [3.1, 2.0], [0.6, 1.6]],
}
PARALLELOGRAMS = ("rectangle", "parallelogram")


def _stiffness(lib):
    lib.or_shell4_stiffness.restype = c_int
    lib.or_shell4_stiffness.argtypes = [c_double] * 3 + [POINTER(c_double)] * 4

    def k(xyz, e=E, nu=NU, h=H):
        p = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).T)
        out = np.zeros(576)
        dp = lambda a: a.ctypes.data_as(POINTER(c_double))  # noqa: E731
        assert lib.or_shell4_stiffness(e, nu, h, dp(p[0]), dp(p[1]), dp(p[2]), dp(out)) == 0
        return out.reshape(24, 24)

    return k


def _area(xy):
    x, y = np.asarray(xy).T
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _flat(xy):
    return np.column_stack([np.asarray(xy, float), np.zeros(4)])


def _field(xy, fn):
    """Nodal DOF vector (ux uy uz rx ry rz per node) of fn(x, y) -> 6 values."""
    return np.concatenate([fn(x, y) for x, y in xy])


def _membrane(e):
    ex, ey, g = e
    return lambda x, y: np.array([ex * x + 0.5 * g * y, 0.5 * g * x + ey * y, 0.0, 0.0, 0.0, 0.0])


def _bending(k):
    kx, ky, kt = k  # w = kx x^2/2 + ky y^2/2 + kt x y
    return lambda x, y: np.array([0.0, 0.0, 0.5 * kx * x * x + 0.5 * ky * y * y + kt * x * y,
                                  ky * y + kt * x, -(kx * x + kt * y), 0.0])


def _shear(gxy):
    gx, gy = gxy
    return lambda x, y: np.array([0.0, 0.0, gx * x + gy * y, 0.0, 0.0, 0.0])


def _drill(s):
    return lambda x, y: np.array([0.0, 0.0, 0.0, 0.0, 0.0, s[0]])


def _plane_matrix(scale, nu):
    return scale * np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, 0.5 * (1.0 - nu)]])


STATES = {
    "membrane": ([(1e-3, 0.0, 0.0), (0.0, 2e-3, 0.0), (0.0, 0.0, 3e-3), (1e-3, -2e-3, 5e-4)], _membrane,
                 lambda A: H * A * _plane_matrix(E / (1.0 - NU * NU), NU)),
    "bending": ([(1e-3, 0.0, 0.0), (0.0, 2e-3, 0.0), (0.0, 0.0, 1.5e-3), (1e-3, -5e-4, 7e-4)], _bending,
                lambda A: A * _plane_matrix(E * H ** 3 / (12.0 * (1.0 - NU * NU)), NU)),
    "shear": ([(1e-3, 0.0), (0.0, 2e-3), (1e-3, -7e-4)], _shear,
              lambda A: A * (5.0 / 6.0) * G * H * np.eye(2)),
    "drill": ([(1e-3,), (-2e-3,)], _drill, lambda A: A * G * H * np.eye(1)),
}


def _strain_vector(kind, s):
    if kind == "bending":
        return np.array([s[0], s[1], 2.0 * s[2]])
    return np.array(s, dtype=float)


def check_energies(stiff, quad):
    """Largest relative error of u1^T K u2 against the closed form over every
    pair of strain states of each kind."""
    xy = QUADS[quad]
    K = stiff(_flat(xy))
    A = _area(xy)
    worst = 0.0
    for kind, (states, make, mat) in STATES.items():
        if kind == "shear" and quad not in PARALLELOGRAMS:
            continue
        D = mat(A)
        us = [_field(xy, make(s)) for s in states]
        for a, sa in enumerate(states):
            for b, sb in enumerate(states):
                want = _strain_vector(kind, sa) @ D @ _strain_vector(kind, sb)
                got = us[a] @ K @ us[b]
                scale = np.sqrt(abs(_strain_vector(kind, sa) @ D @ _strain_vector(kind, sa)) *
                                abs(_strain_vector(kind, sb) @ D @ _strain_vector(kind, sb)))
                worst = max(worst, abs(got - want) / scale)
    return worst


def check_rigid_body(stiff, quad):
    """Largest |K r| over the six rigid-body motions, relative to |K| |r|."""
    xy = QUADS[quad]
    K = stiff(_flat(xy))
    c = np.mean(_flat(xy), axis=0)
    worst = 0.0
    for axis in np.eye(3):
        r = np.concatenate([np.concatenate([axis, np.zeros(3)]) for _ in range(4)])
        worst = max(worst, np.abs(K @ r).max() / (np.abs(K).max() * np.abs(r).max()))
        rot = np.concatenate([np.concatenate([np.cross(axis, p - c), axis]) for p in _flat(xy)])
        worst = max(worst, np.abs(K @ rot).max() / (np.abs(K).max() * np.abs(rot).max()))
    return worst


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q if np.linalg.det(q) > 0 else -q


def check_rotation_invariance(stiff, quad, warp=0.0, seed=3):
    """Relative difference between K of the rotated element and the rotated
    K (a warped element when warp != 0)."""
    xyz = _flat(QUADS[quad])
    xyz[:, 2] = [0.0, warp, 0.0, -warp]
    Q = _rotation(seed)
    Qb = np.kron(np.eye(8), Q)
    K0 = stiff(xyz)
    K1 = stiff(xyz @ Q.T)
    return np.abs(K1 - Qb @ K0 @ Qb.T).max() / np.abs(K0).max()


@pytest.mark.parametrize("quad", sorted(QUADS))
def test_shell4_energies_follow_plate_and_membrane_theory(quad):
    assert check_energies(_stiffness(oracle.load()), quad) < 1e-11


@pytest.mark.parametrize("quad", sorted(QUADS))
def test_shell4_rigid_body_motions_give_no_force(quad):
    assert check_rigid_body(_stiffness(oracle.load()), quad) < 1e-13


@pytest.mark.parametrize("quad,warp", [("general", 0.0), ("trapezoid", 0.0), ("general", 0.15),
                                       ("parallelogram", 0.1)])
def test_shell4_is_invariant_under_rotations_of_space(quad, warp):
    assert check_rotation_invariance(_stiffness(oracle.load()), quad, warp) < 1e-12


def test_shell4_is_symmetric_with_six_zero_modes():
    K = oracle.shell4_stiffness(E, NU, H, _flat(QUADS["general"]))
    assert np.abs(K - K.T).max() <= 1e-15 * np.abs(K).max()
    w = np.linalg.eigvalsh(K)
    assert np.all(np.abs(w[:6]) < 1e-12 * w[-1]) and w[6] > 1e-4 * w[-1]


def test_shell4_refuses_degenerate_geometry():
    with pytest.raises(ValueError):
        oracle.shell4_stiffness(E, NU, H, np.zeros((4, 3)))


MUTATIONS = [
    ("membrane shear modulus (1 + nu)", "{0.0, 0.0, factor * (1.0 - nu) / 2.0}}", "{0.0, 0.0, factor * (1.0 + nu) / 2.0}}"),
    ("bending h^2", "double bend_factor = E * h * h * h / (12.0 * (1.0 - nu * nu));",
     "double bend_factor = E * h * h / (12.0 * (1.0 - nu * nu));"),
    ("shear correction 1", "double kappa = 5.0 / 6.0;", "double kappa = 1.0;"),
    ("drilling stiffness G h / 2", "double Ktt = G * h;", "double Ktt = 0.5 * G * h;"),
    ("tying row sign", "Gm[0][2] = dx41 * q4;", "Gm[0][2] = -dx41 * q4;"),
    ("tying edge 3-4 swapped", "double dx34 = xl[0][2] - xl[0][3], dy34 = xl[1][2] - xl[1][3];",
     "double dx34 = xl[0][3] - xl[0][2], dy34 = xl[1][3] - xl[1][2];"),
    ("shear rotation angle", "double beta = 0.5 * 3.141592653589793 - atan2(Cx, Cy);",
     "double beta = 0.5 * 3.141592653589793 - atan2(Cy, Cx);"),
    ("Jacobian transposed", "sx01 = -xs[0][1] * jinv, sx10 = -xs[1][0] * jinv;",
     "sx01 = -xs[1][0] * jinv, sx10 = -xs[0][1] * jinv;"),
    ("bending block not negated", "for (int q = 3; q < 6; ++q) BJ[p][q] *= -1.0;",
     "for (int q = 3; q < 6; ++q) BJ[p][q] *= 1.0;"),
    ("drilling B2 sign", "double B1 = -0.5 * shp[1][j], B2 = 0.5 * shp[0][j], B6 = -shp[2][j];",
     "double B1 = -0.5 * shp[1][j], B2 = -0.5 * shp[0][j], B6 = -shp[2][j];"),
    ("g2 not orthogonalised", "v2x -= alpha * g1[0];", "v2x -= 0.0 * g1[0];"),
    ("bending Bb[1][0] sign", "Bb[1][0] = shp[1][j];", "Bb[1][0] = -shp[1][j];"),
]


def _mutant(tmp_path, old, new):
    src = tmp_path / "src"
    shutil.copytree(ORACLE_DIR, src, ignore=shutil.ignore_patterns("*.so", "_*", "__pycache__", "*.py"))
    text = (src / "shell4.c").read_text()
    assert text.count(old) == 1, f"mutation target not unique: {old}"
    (src / "shell4.c").write_text(text.replace(old, new))
    lib = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(lib),
                    str(src / "shell4.c"), "-lm"], check=True, capture_output=True)
    return ctypes.CDLL(str(lib))


@pytest.mark.parametrize("name,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_every_planted_shell_slip_fails_a_check(tmp_path, name, old, new):
    stiff = _stiffness(_mutant(tmp_path, old, new))
    worst = 0.0
    for quad in QUADS:
        try:
            worst = max(worst, check_energies(stiff, quad), check_rigid_body(stiff, quad),
                        check_rotation_invariance(stiff, quad), check_rotation_invariance(stiff, quad, 0.15))
        except AssertionError:  # the mutant refused the geometry
            worst = np.inf
    assert worst > 1e-6, f"{name}: no check saw the slip ({worst:.2e})"

"""Pins the oracle's hysteresis branches by closed forms of the models'
geometry, and shows that each check catches a plausible slip.

Steel02 (Menegotto-Pinto with Filippou's isotropic shift,
steel02.mojo:3-114): after a reversal at (eps_r, sig_r) the branch runs
toward the intersection (eps_0, sig_0) of the elastic line from the reversal
point (slope E0) and the opposite hardening asymptote (slope b E0 through
(-+eps_y shift, -+Fy shift), shift = 1 + a1|a3 (d1)^0.8, d1 = (eps_max -
eps_min) / (2 a2|a4 eps_y)); its curvature parameter is R = R0 (1 - cR1 xi
/ (cR2 + xi)) with xi = |eps_pl - eps_0| / eps_y, eps_pl the far end of the
previous excursion; in normalised coordinates e* = (eps - eps_r) / (eps_0 -
eps_r) the stress is sig_r + (sig_0 - sig_r) (b e* + (1 - b) e* / (1 +
|e*|^R)^(1/R)).

Concrete02 (concrete02.mojo:3-109): unloading from the largest compression
(eps_m, sig_m) heads for the focal point where the initial-modulus line
sig = Ec0 eps meets the line of slope rat Ec0 through the ultimate point
(eps_cu, f_cu); the secant to it has slope er and crosses zero stress at
ept = eps_m - sig_m / er; the unloading stress is the elastic step
sig_c + Ec0 d_eps held between sig_m + er (eps - eps_m) and 0.5 er (eps -
ept); past ept the tension envelope starts at ept, and unloading from a
cracked tension strain follows the secant through (ept, 0).

Each check is a function of the library it drives, so it also runs against
copies of the oracle with one slip planted in materials.c (cR1 and cR2
swapped, the shift's exponent, a1 for a3, the focal point, the 0.5 er
slope, ept's sign, the tension secant), compiled here with gcc: every
planted slip must fail its check.
"""

import ctypes
import os
import shutil
import subprocess
from ctypes import POINTER, c_double, c_int

import numpy as np
import pytest

from oracle import oracle
from strut_amd.model import CONCRETE02, STEEL02, make_uniaxial_def

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

STEEL_ISO = {"type": "Steel02", "params": {"Fy": 420.0e3, "E0": 200.0e6, "b": 0.01, "R0": 18.0, "cR1": 0.925,
                                           "cR2": 0.15, "a1": 0.039, "a2": 1.0, "a3": 0.029, "a4": 1.0}}
STEEL_KIN = {"type": "Steel02", "params": {"Fy": 420.0e3, "E0": 200.0e6, "b": 0.01, "R0": 18.0, "cR1": 0.925,
                                           "cR2": 0.15, "a1": 0.0, "a2": 1.0, "a3": 0.0, "a4": 1.0}}
CONC = {"type": "Concrete02", "params": {"fpc": -34.0e3, "epsc0": -0.0025, "fpcu": -24.0e3, "epscu": -0.02}}


def _p(mat):
    return np.array(make_uniaxial_def(mat).p)


def _path(lib, mat_type, params, strains):
    """Stress and tangent of every point of a committed strain path (each
    point is a trial from the previous point's committed state)."""
    lib.or_material_path.restype = None
    lib.or_material_path.argtypes = [c_int, POINTER(c_double), c_int, POINTER(c_double), c_int,
                                     POINTER(c_double), POINTER(c_double)]
    p = np.ascontiguousarray(params, dtype=np.float64)
    s = np.ascontiguousarray(strains, dtype=np.float64)
    sig, tan = np.zeros(len(s)), np.zeros(len(s))
    dp = lambda a: a.ctypes.data_as(POINTER(c_double))  # noqa: E731
    lib.or_material_path(mat_type, dp(p), len(s), dp(s), 1, dp(sig), dp(tan))
    return sig, tan


def _trial_after(lib, mat_type, params, history, strain):
    """The trial response at `strain` after the committed `history`."""
    sig, tan = _path(lib, mat_type, params, list(history) + [strain])
    return sig[-1], tan[-1], sig[:-1]


# ------------------------------------------------------------------- Steel02
def _mp_branch(eps, eps_r, sig_r, eps_0, sig_0, b, R):
    e = (eps - eps_r) / (eps_0 - eps_r)
    return sig_r + (sig_0 - sig_r) * (b * e + (1.0 - b) * e / (1.0 + np.abs(e) ** R) ** (1.0 / R))


def _asymptote_meet(eps_r, sig_r, E0, Esh, eps_pt, sig_pt):
    """Intersection of the line through (eps_r, sig_r) of slope E0 and the
    line through (eps_pt, sig_pt) of slope Esh."""
    eps_0 = (sig_pt - Esh * eps_pt - sig_r + E0 * eps_r) / (E0 - Esh)
    return eps_0, sig_pt + Esh * (eps_0 - eps_pt)


def check_steel02_reversals(lib, mat):
    """Largest relative error of the two reversal branches (tension ->
    compression, then compression -> tension) against the closed form."""
    p = _p(mat)
    Fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4 = p[:10]
    epsy, Esh = Fy / E0, b * E0
    up = list(np.linspace(0.5 * epsy, 5.0 * epsy, 9))
    eps_a = up[-1]
    worst = 0.0
    # branch 1: compression after the tension excursion
    eps_max, eps_min = max(epsy, eps_a), -epsy
    shift = 1.0 + a1 * ((eps_max - eps_min) / (2.0 * a2 * epsy)) ** 0.8
    _, _, hist = _trial_after(lib, STEEL02, p, up, eps_a)
    sig_a = hist[-1]
    e0, s0 = _asymptote_meet(eps_a, sig_a, E0, Esh, -epsy * shift, -Fy * shift)
    xi = abs(eps_min - e0) / epsy
    R = R0 * (1.0 - cR1 * xi / (cR2 + xi))
    down = list(np.linspace(eps_a - 0.4 * epsy, -3.0 * epsy, 8))
    for eps in down:
        got, _, _ = _trial_after(lib, STEEL02, p, up, eps)
        want = _mp_branch(eps, eps_a, sig_a, e0, s0, b, R)
        worst = max(worst, abs(got - want) / Fy)
    # branch 2: tension after the compression excursion to eps_c
    eps_c = down[-1]
    hist2 = up + down
    _, _, h2 = _trial_after(lib, STEEL02, p, hist2, eps_c)
    sig_c = h2[-1]
    eps_max, eps_min = eps_a, min(-epsy, eps_c)
    shift = 1.0 + a3 * ((eps_max - eps_min) / (2.0 * a4 * epsy)) ** 0.8
    e0, s0 = _asymptote_meet(eps_c, sig_c, E0, Esh, epsy * shift, Fy * shift)
    xi = abs(eps_max - e0) / epsy
    R = R0 * (1.0 - cR1 * xi / (cR2 + xi))
    for eps in np.linspace(eps_c + 0.4 * epsy, 6.0 * epsy, 8):
        got, _, _ = _trial_after(lib, STEEL02, p, hist2, eps)
        want = _mp_branch(eps, eps_c, sig_c, e0, s0, b, R)
        worst = max(worst, abs(got - want) / Fy)
    return worst


@pytest.mark.parametrize("mat", [STEEL_KIN, STEEL_ISO], ids=["kinematic", "isotropic"])
def test_steel02_reversal_branches_follow_the_closed_form(mat):
    assert check_steel02_reversals(oracle.load(), mat) < 1e-12


def test_steel02_branch_reaches_its_hardening_asymptote():
    """Far along a reversal branch the stress approaches the hardening line
    through the asymptotes' intersection (slope b E0)."""
    p = _p(STEEL_KIN)
    Fy, E0, b = p[0], p[1], p[2]
    epsy = Fy / E0
    up = list(np.linspace(0.5 * epsy, 5.0 * epsy, 9))
    sig_a = _trial_after(oracle.load(), STEEL02, p, up, up[-1])[2][-1]
    e0, s0 = _asymptote_meet(up[-1], sig_a, E0, b * E0, -epsy, -Fy)
    far = e0 - 100.0 * epsy
    got, tan, _ = _trial_after(oracle.load(), STEEL02, p, up, far)
    line = s0 + b * E0 * (far - e0)
    assert abs(got - line) < 1e-3 * Fy
    assert abs(tan / (b * E0) - 1.0) < 0.02


# ---------------------------------------------------------------- Concrete02
def _c02(p):
    fc, epsc0, fcu, epscu, rat, ft, Ets = p[:7]
    Ec0 = 2.0 * fc / epsc0
    return fc, epsc0, fcu, epscu, rat, ft, Ets, Ec0


def _c02_envelope(p, eps):
    fc, epsc0, fcu, epscu, _, _, _, _ = _c02(p)
    if eps >= epsc0:
        return fc * (eps / epsc0) * (2.0 - eps / epsc0)
    if eps > epscu:
        return fc + (fcu - fc) * (eps - epsc0) / (epscu - epsc0)
    return fcu


def _c02_tension(p, eps):
    _, _, _, _, _, ft, Ets, Ec0 = _c02(p)
    eps0 = ft / Ec0
    if eps <= eps0:
        return Ec0 * eps
    if eps <= ft * (1.0 / Ets + 1.0 / Ec0):
        return ft - Ets * (eps - eps0)
    return 0.0


def _c02_unload_geometry(p, eps_m):
    fc, epsc0, fcu, epscu, rat, ft, Ets, Ec0 = _c02(p)
    # focal point: sig = Ec0 eps meets the line of slope rat Ec0 through (epscu, fcu)
    eps_f = (fcu - rat * Ec0 * epscu) / (Ec0 - rat * Ec0)
    sig_f = Ec0 * eps_f
    sig_m = _c02_envelope(p, eps_m)
    er = (sig_m - sig_f) / (eps_m - eps_f)
    ept = eps_m - sig_m / er
    return sig_m, er, ept, Ec0


def check_concrete02_unload_reload(lib, mat):
    """Largest error (relative to |fc|) of the unloading, clamped-reloading,
    tension-after-compression and tension-secant branches."""
    p = _p(mat)
    fc = p[0]
    worst = 0.0
    for eps_m in (-0.0015, -0.004, -0.012):
        down = list(np.linspace(-1e-4, eps_m, 12))
        sig_m, er, ept, Ec0 = _c02_unload_geometry(p, eps_m)
        assert abs(_path(lib, CONCRETE02, p, down)[0][-1] - sig_m) <= 1e-12 * abs(fc)
        # one unloading step from (eps_m, sig_m): the elastic step held below
        # 0.5 er (eps - ept) (and above the secant back to (eps_m, sig_m))
        for eps in np.linspace(eps_m + 0.02 * (ept - eps_m), ept, 9):
            got, _, _ = _trial_after(lib, CONCRETE02, p, down, eps)
            trial = sig_m + Ec0 * (eps - eps_m)
            want = max(min(trial, 0.5 * er * (eps - ept)), sig_m + er * (eps - eps_m))
            worst = max(worst, abs(got - want) / abs(fc))
        # reloading from a partly unloaded committed point takes the secant
        # back to (eps_m, sig_m)
        eps_u = eps_m + 0.6 * (ept - eps_m)
        hist = down + [eps_u]
        sig_u = _path(lib, CONCRETE02, p, hist)[0][-1]
        for eps in np.linspace(eps_u - 0.1 * (eps_u - eps_m), eps_m + 1e-6, 5):
            got, _, _ = _trial_after(lib, CONCRETE02, p, hist, eps)
            trial = sig_u + Ec0 * (eps - eps_u)
            want = max(min(trial, 0.5 * er * (eps - ept)), sig_m + er * (eps - eps_m))
            worst = max(worst, abs(got - want) / abs(fc))
        # past ept the tension envelope starts at ept
        eps0 = p[5] / Ec0
        for off in (0.3 * eps0, eps0, 2.5 * eps0):
            got, _, _ = _trial_after(lib, CONCRETE02, p, down, ept + off)
            worst = max(worst, abs(got - _c02_tension(p, off)) / abs(fc))
        # cracked at ept + 2.5 eps0 and committed, then unloading toward ept:
        # the secant through (ept, 0) and the cracked point
        dept = 2.5 * eps0
        hist = down + [ept + dept]
        for frac in (0.8, 0.5, 0.1):
            got, tan, _ = _trial_after(lib, CONCRETE02, p, hist, ept + frac * dept)
            slope = _c02_tension(p, dept) / dept
            worst = max(worst, abs(got - slope * frac * dept) / abs(fc), abs(tan - slope) / abs(Ec0))
    return worst


def test_concrete02_unload_reload_and_tension_branches_follow_the_closed_form():
    assert check_concrete02_unload_reload(oracle.load(), CONC) < 1e-12


# ------------------------------------------------------------ planted slips
MUTATIONS = [
    ("R: cR1 and cR2 swapped", "(cR1 * xi) / (cR2 + xi)", "(cR2 * xi) / (cR1 + xi)", "steel"),
    ("isotropic shift exponent 0.8 -> 1", "double shft = 1.0 + a1 * pow(d1, 0.8);",
     "double shft = 1.0 + a1 * pow(d1, 1.0);", "steel"),
    ("a1 used for the tension shift", "double shft = 1.0 + a3 * pow(d1, 0.8);",
     "double shft = 1.0 + a1 * pow(d1, 0.8);", "steel"),
    ("focal point: 1 + rat", "(Ec0 * (1.0 - rat))", "(Ec0 * (1.0 + rat))", "concrete"),
    ("sigmax slope er", "double sigmax = 0.5 * er * (strain - ept);", "double sigmax = er * (strain - ept);",
     "concrete"),
    ("ept sign", "double ept = s->c2_ecmin_t - sigmm / er;", "double ept = s->c2_ecmin_t + sigmm / er;", "concrete"),
    ("tension secant Ec0", "if (s->c2_dept_t != 0.0) e = sicn / s->c2_dept_t;", "if (s->c2_dept_t != 0.0) e = Ec0;",
     "concrete"),
]


def _mutant(tmp_path, old, new):
    src = tmp_path / "src"
    shutil.copytree(ORACLE_DIR, src, ignore=shutil.ignore_patterns("*.so", "_*", "__pycache__", "*.py"))
    text = (src / "materials.c").read_text()
    assert text.count(old) == 1, f"mutation target not unique: {old}"
    (src / "materials.c").write_text(text.replace(old, new))
    lib = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(lib),
                    str(src / "materials.c"), "-lm"], check=True, capture_output=True)
    return ctypes.CDLL(str(lib))


@pytest.mark.parametrize("name,old,new,law", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_every_planted_slip_fails_its_check(tmp_path, name, old, new, law):
    lib = _mutant(tmp_path, old, new)
    if law == "steel":
        worst = max(check_steel02_reversals(lib, STEEL_KIN), check_steel02_reversals(lib, STEEL_ISO))
    else:
        worst = check_concrete02_unload_reload(lib, CONC)
    assert worst > 1e-6, f"{name}: the check did not see the slip ({worst:.2e})"

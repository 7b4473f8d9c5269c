"""Seeded synthetic workloads in Strut's JSON case schema.

The reference ships no generator for its RC-frame benchmarks (its only one,
scripts/gen_frame_case.py, writes elastic static frames), so the sections,
frames and motions used by the tests and the benchmark are generated here,
deterministically, from the configuration text they implement.

Units: kN, m, t, s (stress in kPa).

Sections ("RC fiber section"): a rectangular b x h section with a confined
core meshed n x n with Concrete02 (core), a cover two fibers thick on every
side with unconfined Concrete02 (cover), and straight layers of Steel02 bars
split between the top and bottom faces.  The cover thickness to the core
edge is 40 mm.
"""

from __future__ import annotations

import cmath
import math
from typing import Dict, List, Tuple

import numpy as np

MPA = 1000.0  # kPa per MPa

CORE_CONCRETE = {"fpc": -34.0 * MPA, "epsc0": -0.0025, "fpcu": -24.0 * MPA, "epscu": -0.02}
# The unconfined cover's strain at peak is not given by the configuration
# text; -0.002 is the usual unconfined value.
COVER_CONCRETE = {"fpc": -28.0 * MPA, "epsc0": -0.002, "fpcu": 0.0, "epscu": -0.005}
STEEL = {"Fy": 420.0 * MPA, "E0": 200.0e6, "b": 0.01, "R0": 18.0, "cR1": 0.925, "cR2": 0.15,
         "a1": 0.0, "a2": 1.0, "a3": 0.0, "a4": 1.0}
BAR_AREA_20MM = math.pi * 0.020 ** 2 / 4.0


def rc_materials(core_id: int = 1, cover_id: int = 2, steel_id: int = 3) -> List[dict]:
    return [
        {"id": core_id, "type": "Concrete02", "params": dict(CORE_CONCRETE)},
        {"id": cover_id, "type": "Concrete02", "params": dict(COVER_CONCRETE)},
        {"id": steel_id, "type": "Steel02", "params": dict(STEEL)},
    ]


def rc_section(sec_id: int, b: float, h: float, n_core: int, n_bars: int, dim: int = 2,
               core_id: int = 1, cover_id: int = 2, steel_id: int = 3, cover: float = 0.04) -> dict:
    """Rectangular RC fiber section: y is the bending (depth h) axis, z the
    width b axis.  Patch order: core, cover bottom, top, left, right; then
    the bottom and top bar layers."""
    yc, zc = h / 2.0 - cover, b / 2.0 - cover
    n_side = n_core + 2
    patches = [
        {"type": "rect", "material": core_id, "num_subdiv_y": n_core, "num_subdiv_z": n_core,
         "y_i": -yc, "z_i": -zc, "y_j": yc, "z_j": zc},
        {"type": "rect", "material": cover_id, "num_subdiv_y": 2, "num_subdiv_z": n_side,
         "y_i": -h / 2.0, "z_i": -b / 2.0, "y_j": -yc, "z_j": b / 2.0},
        {"type": "rect", "material": cover_id, "num_subdiv_y": 2, "num_subdiv_z": n_side,
         "y_i": yc, "z_i": -b / 2.0, "y_j": h / 2.0, "z_j": b / 2.0},
        {"type": "rect", "material": cover_id, "num_subdiv_y": n_side, "num_subdiv_z": 2,
         "y_i": -yc, "z_i": -b / 2.0, "y_j": yc, "z_j": -zc},
        {"type": "rect", "material": cover_id, "num_subdiv_y": n_side, "num_subdiv_z": 2,
         "y_i": -yc, "z_i": zc, "y_j": yc, "z_j": b / 2.0},
    ]
    per_face = n_bars // 2
    layers = [
        {"type": "straight", "material": steel_id, "num_bars": per_face, "bar_area": BAR_AREA_20MM,
         "y_start": -yc, "z_start": -zc, "y_end": -yc, "z_end": zc},
        {"type": "straight", "material": steel_id, "num_bars": n_bars - per_face, "bar_area": BAR_AREA_20MM,
         "y_start": yc, "z_start": -zc, "y_end": yc, "z_end": zc},
    ]
    params: Dict = {"patches": patches, "layers": layers}
    stype = "FiberSection2d" if dim == 2 else "FiberSection3d"
    if dim == 3:
        # elastic torsion for forceBeamColumn3d (G = E/2.4, J of the gross section)
        ec = 2.0 * abs(CORE_CONCRETE["fpc"]) / abs(CORE_CONCRETE["epsc0"])
        a_, b_ = max(b, h), min(b, h)
        params["G"] = ec / 2.4
        params["J"] = a_ * b_ ** 3 * (1.0 / 3.0 - 0.21 * (b_ / a_) * (1.0 - b_ ** 4 / (12.0 * a_ ** 4)))
    return {"id": sec_id, "type": stype, "params": params}


def frame3d_section_population(stories: int = 20, bays_x: int = 5, bays_y: int = 5,
                               nip: int = 5) -> Dict[str, int]:
    """Section counts of the 3D RC frame workload (20 stories x 5 x 5 bays,
    forceBeamColumn columns and dispBeamColumn beams, 5 integration points
    each): columns and beams per floor, times stories and nip."""
    cols = (bays_x + 1) * (bays_y + 1) * stories
    beams = (bays_x * (bays_y + 1) + bays_y * (bays_x + 1)) * stories
    return {"columns": cols, "beams": beams, "column_sections": cols * nip, "beam_sections": beams * nip}


# Config-1 section shapes: columns 400 x 400, beams 300 x 500, 12 x 12
# confined core, 12 bars.
FRAME3D_COLUMN = dict(b=0.4, h=0.4, n_core=12, n_bars=12)
FRAME3D_BEAM = dict(b=0.3, h=0.5, n_core=12, n_bars=12)
# Config-0 section shapes: 10 x 10 core, 8 bars.
FRAME2D_COLUMN = dict(b=0.4, h=0.4, n_core=10, n_bars=8)
FRAME2D_BEAM = dict(b=0.3, h=0.5, n_core=10, n_bars=8)


def synthetic_section_deformation(n: int, step: int, seed: int, dim: int) -> Tuple:
    """Deterministic cyclic section deformations for n section points at
    a given step: axial strain and curvatures on independent phases with
    amplitudes that take the cover past its peak and the bars past yield.
    Returns numpy arrays (eps0, kappa_z, kappa_y)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    amp_k = rng.uniform(0.002, 0.014, size=n)
    amp_e = rng.uniform(1.0e-4, 6.0e-4, size=n)
    ph = rng.uniform(0.0, 2.0 * math.pi, size=(3, n))
    w = rng.uniform(0.15, 0.6, size=n)
    t = float(step)
    eps0 = -2.0e-4 + amp_e * np.sin(w * t + ph[0])
    kz = amp_k * np.sin(w * t + ph[1]) * (1.0 + 0.25 * math.sin(0.05 * t))
    ky = 0.0 * kz if dim == 2 else 0.6 * amp_k * np.sin(0.8 * w * t + ph[2])
    return eps0, kz, ky


# ---------------------------------------------------------------- motions
_MASK64 = (1 << 64) - 1


def splitmix64(seed: int):
    """SplitMix64 stream (Steele, Lea, Flood 2014): portable across numpy
    versions, so the same seed yields the same record everywhere."""
    state = seed & _MASK64
    while True:
        state = (state + 0x9E3779B97F4A7C15) & _MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        yield z ^ (z >> 31)


def gaussian_noise(seed: int, n: int):
    """Box-Muller normals from SplitMix64 uniforms on (0, 1]."""
    import numpy as np

    gen = splitmix64(seed)
    out = np.empty(n)
    for i in range(0, n, 2):
        u1 = ((next(gen) >> 11) + 1) * (1.0 / 9007199254740992.0)
        u2 = ((next(gen) >> 11) + 1) * (1.0 / 9007199254740992.0)
        r = math.sqrt(-2.0 * math.log(u1))
        out[i] = r * math.cos(2.0 * math.pi * u2)
        if i + 1 < n:
            out[i + 1] = r * math.sin(2.0 * math.pi * u2)
    return out


def butterworth_bandpass_sos(order: int, f_lo: float, f_hi: float, fs: float):
    """Digital Butterworth band-pass of the given prototype order as
    second-order sections (b0, b1, b2, a1, a2), a0 = 1: the analog
    prototype's left-half-plane poles exp(i pi (2k + N + 1) / 2N), the
    low-pass to band-pass transform s -> (s^2 + w1 w2) / ((w2 - w1) s) with
    edges pre-warped to w = 2 fs tan(pi f / fs), then the bilinear
    transform z = (2 fs + s) / (2 fs - s).  The N analog zeros at s = 0 go
    to z = 1 and the N at infinity to z = -1, one of each per section; the
    gain sits in the first section, and sections are ordered by pole
    radius, the pair nearest the unit circle last."""
    fs2 = 2.0 * fs
    w1 = fs2 * math.tan(math.pi * f_lo / fs)
    w2 = fs2 * math.tan(math.pi * f_hi / fs)
    bw, w0sq = w2 - w1, w1 * w2
    poles = []
    for k in range(order):
        half = cmath.exp(1j * math.pi * (2 * k + order + 1) / (2 * order)) * (0.5 * bw)
        disc = cmath.sqrt(half * half - w0sq)
        poles += [half + disc, half - disc]
    gain = complex(bw ** order * fs2 ** order)
    for p in poles:
        gain /= fs2 - p
    zp = [(fs2 + p) / (fs2 - p) for p in poles]
    upper = sorted((z for z in zp if z.imag > 0.0), key=abs)
    if len(upper) != order:
        raise ValueError("band edges too close for a conjugate-pair design")
    sos = []
    for i, z in enumerate(upper):
        g = gain.real if i == 0 else 1.0
        sos.append((g, 0.0, -g, -2.0 * z.real, z.real * z.real + z.imag * z.imag))
    return sos


def _sos_pass(sos, x, x0):
    """One causal pass through the sections (transposed direct form II),
    each section starting in its steady state for a constant input equal
    to x0 (the filtfilt initial conditions of Gustafsson 1996)."""
    y = list(x)
    v = x0
    for b0, b1, b2, a1, a2 in sos:
        y_ss = v * (b0 + b1 + b2) / (1.0 + a1 + a2)
        z2 = b2 * v - a2 * y_ss
        z1 = b1 * v - a1 * y_ss + z2
        for i, xi in enumerate(y):
            yi = b0 * xi + z1
            z1 = b1 * xi - a1 * yi + z2
            z2 = b2 * xi - a2 * yi
            y[i] = yi
        v = y_ss
    return y


def sos_filtfilt(sos, x):
    """Zero-phase filtering: odd extension by 3 (2 sections + 1) samples at
    each end (at most len - 1), a forward and a backward pass, the
    extension cut off again."""
    x = [float(v) for v in x]
    n = len(x)
    pad = min(3 * (2 * len(sos) + 1), n - 1)
    ext = [2.0 * x[0] - x[k] for k in range(pad, 0, -1)] + x + [2.0 * x[-1] - x[n - 2 - k] for k in range(pad)]
    fwd = _sos_pass(sos, ext, ext[0])
    back = _sos_pass(sos, fwd[::-1], fwd[-1])[::-1]
    return back[pad:pad + n]


def ground_motion(seed: int, dt: float, n_steps: int, pga_g: float, f_lo: float = 0.1, f_hi: float = 25.0,
                  rise: float = 2.0, strong: float = 6.0, decay: float = 2.0):
    """Band-limited white noise (4th-order Butterworth band-pass applied
    forward and backward) under a trapezoidal envelope, scaled to a peak
    ground acceleration of pga_g * 9.81 m/s^2.  Returns n_steps + 1 samples
    at spacing dt starting at t = 0.  Python floats and the math module
    only, so the record does not depend on a numerical library's version."""
    import numpy as np

    n = n_steps + 1
    raw = gaussian_noise(seed, n)
    fs = 1.0 / dt
    hi = min(f_hi, 0.45 * fs)
    band = np.array(sos_filtfilt(butterworth_bandpass_sos(4, f_lo, hi, fs), raw))
    t = np.arange(n) * dt
    env = np.clip(np.minimum(t / rise, (rise + strong + decay - t) / decay), 0.0, 1.0)
    acc = band * env
    peak = np.max(np.abs(acc))
    return acc * (pga_g * 9.81 / peak)


# ---------------------------------------------------------------- frames
def _section_flex_elastic(cells, mats):
    """Initial section stiffness terms (EA, ES, EI) about the centroid."""
    import numpy as np

    a = cells["area"]
    e = np.array([_init_tangent(mats.defs[m]) for m in cells["mat"]])
    yb = np.sum(a * cells["y"]) / np.sum(a)
    yr = cells["y"] - yb
    ks = e * a
    return ks.sum(), (-ks * yr).sum(), (ks * yr * yr).sum()


def _init_tangent(d):
    from .model import CONCRETE01, CONCRETE02, ELASTIC, STEEL01, STEEL02

    if d.mat_type == ELASTIC:
        return d.p[0]
    if d.mat_type in (STEEL01, STEEL02):
        return d.p[1]
    return 2.0 * d.p[0] / d.p[1]


def rc_frame2d_case(stories: int, bays: int, steps: int, dt: float, seed: int, pga_g: float,
                    story_h: float = 3.0, bay_w: float = 6.0, floor_mass: float = 50.0, nip: int = 5,
                    strong: float = 6.0, tol: float = 1.0e-8, max_iters: int = 25, name: str = "",
                    column_type: str = "forceBeamColumn2d", beam_type: str = "forceBeamColumn2d") -> dict:
    """2D RC moment frame in Strut's JSON schema: forceBeamColumn2d columns
    and beams (Lobatto, nip points, Linear), fixed bases, lumped floor masses
    on ux and uy of every floor node, Rayleigh 5 % on modes 1 and 3
    (alphaM + betaKInit, from the initial stiffness), Newmark 1/2-1/4,
    Newton / NormUnbalance / BandGeneral, UniformExcitation in x."""
    import numpy as np

    from .model import MaterialTable, discretize_section

    nodes = []
    nid = 1
    node_at = {}
    for f in range(stories + 1):
        for c in range(bays + 1):
            node = {"id": nid, "x": c * bay_w, "y": f * story_h}
            if f == 0:
                node["constraints"] = [1, 2, 3]
            nodes.append(node)
            node_at[(f, c)] = nid
            nid += 1
    elements = []
    eid = 1
    for f in range(stories):
        for c in range(bays + 1):
            elements.append({"id": eid, "type": column_type, "nodes": [node_at[(f, c)], node_at[(f + 1, c)]],
                             "section": 1, "geomTransf": "Linear", "integration": "Lobatto", "num_int_pts": nip})
            eid += 1
    for f in range(1, stories + 1):
        for c in range(bays):
            elements.append({"id": eid, "type": beam_type, "nodes": [node_at[(f, c)], node_at[(f, c + 1)]],
                             "section": 2, "geomTransf": "Linear", "integration": "Lobatto", "num_int_pts": nip})
            eid += 1
    masses = []
    for f in range(1, stories + 1):
        for c in range(bays + 1):
            for dof in (1, 2):
                masses.append({"node": node_at[(f, c)], "dof": dof, "value": floor_mass})
    materials = rc_materials()
    sections = [rc_section(1, dim=2, **FRAME2D_COLUMN), rc_section(2, dim=2, **FRAME2D_BEAM)]
    # Rayleigh from the initial (elastic-section) stiffness, modes 1 and 3
    mt = MaterialTable(materials)
    sec_k = {s["id"]: _section_flex_elastic(discretize_section(s, mt), mt) for s in sections}
    n_nodes = len(nodes)
    K = np.zeros((3 * n_nodes, 3 * n_nodes))
    idx = {n["id"]: i for i, n in enumerate(nodes)}
    for el in elements:
        i, j = idx[el["nodes"][0]], idx[el["nodes"][1]]
        xi, yi, xj, yj = nodes[i]["x"], nodes[i]["y"], nodes[j]["x"], nodes[j]["y"]
        L = math.hypot(xj - xi, yj - yi)
        c, s = (xj - xi) / L, (yj - yi) / L
        EA, ES, EI = sec_k[el["section"]]
        det = EA * EI - ES * ES
        f00, f01, f11 = EI / det, -ES / det, EA / det
        # uniform section flexibility integrates exactly: F = L * int b^T f b
        F = np.array([[f00 * L, f01 * L * (-0.5), f01 * L * 0.5],
                      [f01 * L * (-0.5), f11 * L / 3.0, -f11 * L / 6.0],
                      [f01 * L * 0.5, -f11 * L / 6.0, f11 * L / 3.0]])
        kb = np.linalg.inv(F)
        A = np.array([[-1, 0, 0, 1, 0, 0], [0, 1 / L, 1, 0, -1 / L, 0], [0, 1 / L, 0, 0, -1 / L, 1]], dtype=float)
        kl = A.T @ kb @ A
        T = np.zeros((6, 6))
        for b in (0, 3):
            T[b:b + 2, b:b + 2] = [[c, s], [-s, c]]
            T[b + 2, b + 2] = 1.0
        kg = T.T @ kl @ T
        dofs = [3 * i, 3 * i + 1, 3 * i + 2, 3 * j, 3 * j + 1, 3 * j + 2]
        K[np.ix_(dofs, dofs)] += kg
    Mv = np.zeros(3 * n_nodes)
    for m in masses:
        Mv[3 * idx[m["node"]] + m["dof"] - 1] += m["value"]
    free = [k for k in range(3 * n_nodes) if k // 3 >= bays + 1]
    Kf = K[np.ix_(free, free)]
    mf = Mv[free]
    dyn = mf > 0
    # static condensation of the massless rotations
    Kdd = Kf[np.ix_(dyn, dyn)] - Kf[np.ix_(dyn, ~dyn)] @ np.linalg.solve(Kf[np.ix_(~dyn, ~dyn)], Kf[np.ix_(~dyn, dyn)])
    Md = mf[dyn]
    lam = np.sort(np.linalg.eigvalsh(Kdd / np.sqrt(np.outer(Md, Md))))
    w1, w3 = math.sqrt(lam[0]), math.sqrt(lam[2])
    zeta = 0.05
    alpha_m = 2.0 * zeta * w1 * w3 / (w1 + w3)
    beta_k = 2.0 * zeta / (w1 + w3)
    acc = ground_motion(seed, dt, steps, pga_g, strong=strong)
    return {
        "schema_version": "1.0",
        "metadata": {"name": name or f"rc_frame2d_{stories}x{bays}_seed{seed}", "units": "kN-m-t-s"},
        "model": {"ndm": 2, "ndf": 3},
        "nodes": nodes,
        "materials": materials,
        "sections": sections,
        "elements": elements,
        "masses": masses,
        "time_series": [{"type": "Path", "tag": 1, "dt": dt, "values": [float(v) for v in acc], "factor": 1.0}],
        "pattern": {"type": "UniformExcitation", "tag": 1, "direction": 1, "accel": 1},
        "rayleigh": {"alphaM": alpha_m, "betaKInit": beta_k},
        "analysis": {"type": "transient_nonlinear", "steps": steps, "dt": dt, "tol": tol, "max_iters": max_iters,
                     "system": "BandGeneral", "numberer": "RCM", "constraints": "Plain", "algorithm": "Newton",
                     "test_type": "NormUnbalance", "integrator": {"type": "Newmark", "gamma": 0.5, "beta": 0.25}},
        "recorders": [{"type": "node_displacement", "nodes": [node_at[(stories, 0)]], "dofs": [1, 2, 3],
                       "output": "roof"}],
    }


def config0_case(steps: int = 1000, seed: int = 0) -> dict:
    """3 stories x 1 bay, 3 m / 6 m, 50 t floor masses, dt 0.01 s, PGA 0.4 g,
    2 s rise / 6 s strong / 2 s decay."""
    return rc_frame2d_case(3, 1, steps=steps, dt=0.01, seed=seed, pga_g=0.4, strong=6.0)


# ------------------------------------------------ uneven ensemble (config 2)
CONFIG2_FRAMES = ((4, 3), (8, 4), (12, 6))  # stories x bays of the three frame variants
CONFIG2_PGA_GRID = tuple(0.1 * (15.0 ** (k / 15.0)) for k in range(16))  # 16-point geometric grid, 0.1..1.5 g


def config2_motion_params(seed: int, dt: float = 0.01) -> dict:
    """Parameters of one ground motion of the uneven ensemble: corner
    frequency drawn uniformly from 2-15 Hz, strong-motion duration from
    5-20 s, PGA from the 16-point geometric grid over 0.1-1.5 g; record
    length = 2 s rise + strong + 2 s decay + 1 s of free vibration, clamped
    to 1000-3000 steps."""
    gen = splitmix64(seed ^ 0x5DEECE66D)
    u = [((next(gen) >> 11) + 0.5) * (1.0 / 9007199254740992.0) for _ in range(3)]
    f_corner = 2.0 + 13.0 * u[0]
    strong = 5.0 + 15.0 * u[1]
    pga_g = CONFIG2_PGA_GRID[min(15, int(u[2] * 16))]
    steps = int(round((2.0 + strong + 2.0 + 1.0) / dt))
    steps = max(1000, min(3000, steps))
    return {"seed": seed, "f_corner": f_corner, "strong": strong, "pga_g": pga_g, "steps": steps, "dt": dt}


def config2_motion(seed: int, dt: float = 0.01, max_steps: int = 3000) -> dict:
    """config2_motion_params plus the record (values at t = 0, dt, ...);
    max_steps caps the record length for short runs."""
    m = config2_motion_params(seed, dt)
    m["steps"] = min(m["steps"], max_steps)
    m["values"] = ground_motion(seed, dt, m["steps"], m["pga_g"], f_hi=m["f_corner"], strong=m["strong"])
    return m


def config2_frame_case(stories: int, bays: int, steps: int = 1000) -> dict:
    """A frame variant of the uneven ensemble (sections as in config 0);
    its own record is a placeholder: instances bring config2_motion records."""
    return rc_frame2d_case(stories, bays, steps=steps, dt=0.01, seed=0, pga_g=0.4,
                           name=f"config2_frame_{stories}x{bays}")


def peak_story_drift(u_hist, stories: int, bays: int, story_h: float = 3.0):
    """Peak inter-story drift ratio of each story (max over time and column
    lines of |u_x(floor k) - u_x(floor k - 1)| / h) from a displacement
    history over every DOF of an rc_frame2d_case frame (node f * (bays + 1) + c,
    ux = DOF 3 * node)."""
    import numpy as np

    u = np.asarray(u_hist)
    cols = bays + 1
    ux = u[:, 0::3].reshape(u.shape[0], stories + 1, cols)
    drift = np.abs(np.diff(ux, axis=1)) / story_h
    return drift.max(axis=(0, 2))


# ------------------------------------- 3D frame with rigid floor diaphragms
def frame3d_diaphragm_model(stories: int = 2, bays_x: int = 2, bays_y: int = 1, h: float = 3.0, lx: float = 6.0,
                            ly: float = 5.0) -> dict:
    """Node grid of a 3D frame (6 DOFs per node) with a retained node at each
    floor's centre joined to the floor's nodes by a rigidDiaphragm
    (perpDirn 3), the base fixed and each retained node's uz, rx, ry fixed, as
    OpenSees-style models set them up.  Also lists the column and beam
    elements (node index pairs) for assembling a stiffness."""
    from .constraints import mpc_tables, rigid_diaphragm

    coords, grid = {}, {}
    nid = 1
    for k in range(stories + 1):
        for j in range(bays_y + 1):
            for i in range(bays_x + 1):
                coords[nid] = (i * lx, j * ly, k * h)
                grid[(i, j, k)] = nid
                nid += 1
    masters = {}
    for k in range(1, stories + 1):
        coords[nid] = (bays_x * lx / 2.0, bays_y * ly / 2.0, k * h)
        masters[k] = nid
        nid += 1
    ids = sorted(coords)
    node_index = {n: idx for idx, n in enumerate(ids)}
    ndf = 6
    sp = np.zeros(len(ids) * ndf, dtype=bool)
    for j in range(bays_y + 1):
        for i in range(bays_x + 1):
            sp[node_index[grid[(i, j, 0)]] * ndf:node_index[grid[(i, j, 0)]] * ndf + 6] = True
    mps = []
    for k, m in masters.items():
        for d in (3, 4, 5):
            sp[node_index[m] * ndf + d - 1] = True
        floor = [grid[(i, j, k)] for j in range(bays_y + 1) for i in range(bays_x + 1)]
        mps += rigid_diaphragm(coords, m, floor, 3)
    elems = []
    for k in range(stories):
        for j in range(bays_y + 1):
            for i in range(bays_x + 1):
                elems.append((node_index[grid[(i, j, k)]], node_index[grid[(i, j, k + 1)]]))
    for k in range(1, stories + 1):
        for j in range(bays_y + 1):
            for i in range(bays_x):
                elems.append((node_index[grid[(i, j, k)]], node_index[grid[(i + 1, j, k)]]))
        for j in range(bays_y):
            for i in range(bays_x + 1):
                elems.append((node_index[grid[(i, j, k)]], node_index[grid[(i, j + 1, k)]]))
    tables = mpc_tables(len(ids), ndf, sp, mps, node_index)
    return {"coords": coords, "node_index": node_index, "ndf": ndf, "sp": sp, "mp_constraints": mps,
            "elements": elems, "masters": masters, "tables": tables}


def synthetic_frame_stiffness(model: dict, seed: int = 0) -> np.ndarray:
    """A dense symmetric stiffness with the frame's element sparsity: each
    element adds a seeded symmetric positive definite 12 x 12 block on its
    two nodes' DOFs (a stand-in for the assembled K in the collapse tests and
    timing)."""
    n = len(model["node_index"]) * model["ndf"]
    rng = np.random.default_rng(seed)
    K = np.zeros((n, n))
    for a, b in model["elements"]:
        m = rng.normal(size=(12, 12))
        ke = m @ m.T + 12.0 * np.eye(12)
        dofs = list(range(6 * a, 6 * a + 6)) + list(range(6 * b, 6 * b + 6))
        K[np.ix_(dofs, dofs)] += ke
    return K


# ------------------------------------- 3D RC frame transient (rigid floors)
def rc_frame3d_case(stories: int, bays_x: int, bays_y: int, steps: int, dt: float = 0.005,
                    seeds: Tuple[int, int] = (0, 1), pgas_g: Tuple[float, float] = (0.5, 0.35),
                    story_h: float = 3.0, bay_w: float = 6.0, floor_mass: float = 50.0, nip: int = 5,
                    n_core: int = 12, n_bars: int = 12, strong: float = 6.0, tol: float = 1.0e-8,
                    max_iters: int = 25, diaphragms: bool = True, bidirectional: bool = True,
                    column_type: str = "forceBeamColumn3d", beam_type: str = "dispBeamColumn3d",
                    column_transf: str = "PDelta", name: str = "") -> dict:
    """3D RC moment frame in Strut's JSON schema: forceBeamColumn3d columns
    (400 x 400 mm, PDelta) and dispBeamColumn3d beams (300 x 500 mm, depth
    vertical, Linear), nip Lobatto points, n_core x n_core confined core,
    two-fiber cover and n_bars Steel02 bars; fixed bases; lumped masses on
    ux, uy, uz of every floor node; per floor a retained node at the plan
    centre (uz, rx, ry fixed) tied to the floor's nodes by rigidDiaphragm
    (perpDirn 3) under the Transformation handler; two UniformExcitation
    records (x, y; seeded band-limited noise at pgas_g) as the ``patterns``
    list; Newmark 1/2-1/4, Newton / NormUnbalance / BandGeneral.

    Rayleigh 5 % on modes 1 and 3 from the estimates T1 = 0.1 s x stories
    and T3 = T1 / 5 (alphaM + betaKInit)."""
    from .constraints import rigid_diaphragm

    nodes, node_at, nid = [], {}, 1
    for k in range(stories + 1):
        for j in range(bays_y + 1):
            for i in range(bays_x + 1):
                node = {"id": nid, "x": i * bay_w, "y": j * bay_w, "z": k * story_h}
                if k == 0:
                    node["constraints"] = [1, 2, 3, 4, 5, 6]
                nodes.append(node)
                node_at[(i, j, k)] = nid
                nid += 1
    masters = {}
    if diaphragms:
        for k in range(1, stories + 1):
            nodes.append({"id": nid, "x": bays_x * bay_w / 2.0, "y": bays_y * bay_w / 2.0, "z": k * story_h,
                          "constraints": [3, 4, 5]})
            masters[k] = nid
            nid += 1
    elements, eid = [], 1
    for k in range(stories):
        for j in range(bays_y + 1):
            for i in range(bays_x + 1):
                elements.append({"id": eid, "type": column_type, "nodes": [node_at[(i, j, k)], node_at[(i, j, k + 1)]],
                                 "section": 1, "geomTransf": column_transf, "vecxz": [1.0, 0.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
    for k in range(1, stories + 1):
        for j in range(bays_y + 1):
            for i in range(bays_x):
                elements.append({"id": eid, "type": beam_type, "nodes": [node_at[(i, j, k)], node_at[(i + 1, j, k)]],
                                 "section": 2, "geomTransf": "Linear", "vecxz": [0.0, 1.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
        for j in range(bays_y):
            for i in range(bays_x + 1):
                elements.append({"id": eid, "type": beam_type, "nodes": [node_at[(i, j, k)], node_at[(i, j + 1, k)]],
                                 "section": 2, "geomTransf": "Linear", "vecxz": [1.0, 0.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
    masses = [{"node": node_at[(i, j, k)], "dof": d, "value": floor_mass}
              for k in range(1, stories + 1) for j in range(bays_y + 1) for i in range(bays_x + 1) for d in (1, 2, 3)]
    coords = {n["id"]: (n["x"], n["y"], n["z"]) for n in nodes}
    mps = []
    for k, m in masters.items():
        floor = [node_at[(i, j, k)] for j in range(bays_y + 1) for i in range(bays_x + 1)]
        mps += rigid_diaphragm(coords, m, floor, 3)
    sections = [rc_section(1, dim=3, b=0.4, h=0.4, n_core=n_core, n_bars=n_bars),
                rc_section(2, dim=3, b=0.3, h=0.5, n_core=n_core, n_bars=n_bars)]
    t1 = 0.1 * stories
    w1, w3 = 2.0 * math.pi / t1, 2.0 * math.pi / (t1 / 5.0)
    zeta = 0.05
    series, patterns = [], []
    for tag, (seed, pga, direction) in enumerate(zip(seeds, pgas_g, (1, 2)), start=1):
        if not bidirectional and tag > 1:
            break
        acc = ground_motion(seed, dt, steps, pga, strong=strong)
        series.append({"type": "Path", "tag": tag, "dt": dt, "values": [float(v) for v in acc], "factor": 1.0})
        patterns.append({"type": "UniformExcitation", "tag": tag, "direction": direction, "accel": tag})
    case = {
        "schema_version": "1.0",
        "metadata": {"name": name or f"rc_frame3d_{stories}x{bays_x}x{bays_y}", "units": "kN-m-t-s"},
        "model": {"ndm": 3, "ndf": 6},
        "nodes": nodes,
        "materials": rc_materials(),
        "sections": sections,
        "elements": elements,
        "masses": masses,
        "time_series": series,
        "patterns": patterns,
        "rayleigh": {"alphaM": 2.0 * zeta * w1 * w3 / (w1 + w3), "betaKInit": 2.0 * zeta / (w1 + w3)},
        "analysis": {"type": "transient_nonlinear", "steps": steps, "dt": dt, "tol": tol, "max_iters": max_iters,
                     "system": "BandGeneral", "numberer": "Plain",
                     "constraints": "Transformation" if diaphragms else "Plain", "algorithm": "Newton",
                     "test_type": "NormUnbalance", "integrator": {"type": "Newmark", "gamma": 0.5, "beta": 0.25}},
        "recorders": [{"type": "node_displacement",
                       "nodes": [masters[stories] if diaphragms else node_at[(0, 0, stories)]],
                       "dofs": [1, 2, 6], "output": "roof"}],
    }
    # inter-story drift of the corner column line in x and y, and the floor
    # envelopes (the retained nodes with diaphragms)
    line = [node_at[(0, 0, k)] for k in range(stories + 1)]
    for k in range(1, stories + 1):
        for d in (1, 2):
            case["recorders"].append({"type": "drift", "i_node": line[k - 1], "j_node": line[k], "dof": d,
                                      "perp_dirn": 3, "output": f"drift_{k}_{d}"})
    case["recorders"].append({"type": "envelope_node_displacement",
                              "nodes": [masters[k] for k in range(1, stories + 1)] if diaphragms else line[1:],
                              "dofs": [1, 2, 6], "output": "floor_envelope"})
    if mps:
        case["mp_constraints"] = mps
    return case


def config1_case(steps: int = 4000) -> dict:
    """20 stories x 5 x 5 bays, forceBeamColumn3d columns and
    dispBeamColumn3d beams (5 IPs, 268-fiber sections: 12 x 12 core, 12
    bars), rigid diaphragms, two records at 0.5 g and 0.35 g, dt 0.005 s."""
    return rc_frame3d_case(20, 5, 5, steps=steps, dt=0.005, seeds=(0, 1), pgas_g=(0.5, 0.35))


# ------------------------------- 3D RC tower with an elastic shell core wall
def rc_tower_case(stories: int, bays: int, steps: int, dt: float = 0.005, core: Tuple[int, int] = (2, 2),
                  wall_h_div: int = 4, wall_v_div: int = 2, seeds: Tuple[int, int] = (0, 1),
                  pgas_g: Tuple[float, float] = (0.5, 0.35), story_h: float = 3.0, bay_w: float = 6.0,
                  floor_mass: float = 50.0, nip: int = 5, n_core: int = 8, n_bars: int = 8,
                  wall_E: float = 30.0e6, wall_nu: float = 0.2, wall_t: float = 0.4, strong: float = 6.0,
                  tol: float = 1.0e-8, max_iters: int = 25, name: str = "") -> dict:
    """3D RC frame tower with an elastic shell core wall, in Strut's JSON
    schema.

    Frame: bays x bays plan, forceBeamColumn3d columns (400 x 400 mm, PDelta)
    at every grid point and dispBeamColumn3d beams (300 x 500 mm) along both
    grid directions, nip Lobatto points, n_core x n_core confined core and
    n_bars Steel02 bars (8 and 8: 152 fibers per section).  Core wall: the
    perimeter of the core[0] x core[1] central bays, 0.4 m thick elastic
    ``shell`` elements (ElasticMembranePlateSection, E 30 GPa, nu 0.2, no
    mass), each bay face split into wall_h_div panels along the plan and
    every story into wall_v_div panels up the height.  Floors: fixed base;
    per floor a retained node at the plan centre (uz, rx, ry fixed) tied by
    rigidDiaphragm (perpDirn 3) to every node of the floor level (grid and
    wall-edge nodes); the wall's nodes between floors keep six DOFs.  Nodes
    are numbered level by level (the Plain numberer then keeps the band to
    about one story's DOFs).  Lumped floor_mass on ux, uy, uz of every grid
    node; two UniformExcitation records (x, y); Rayleigh 5 % on modes 1 and
    3 from T1 = 0.1 s x stories and T3 = T1 / 5; Newmark 1/2-1/4, Newton,
    NormUnbalance, BandGeneral, Transformation."""
    from .constraints import rigid_diaphragm

    i0 = (bays - core[0]) // 2
    j0 = (bays - core[1]) // 2
    i1, j1 = i0 + core[0], j0 + core[1]
    # perimeter points of the core (plan coordinates as multiples of bay_w / wall_h_div), counter-clockwise
    per = []
    for i in range(i0 * wall_h_div, i1 * wall_h_div):
        per.append((i, j0 * wall_h_div))
    for j in range(j0 * wall_h_div, j1 * wall_h_div):
        per.append((i1 * wall_h_div, j))
    for i in range(i1 * wall_h_div, i0 * wall_h_div, -1):
        per.append((i, j1 * wall_h_div))
    for j in range(j1 * wall_h_div, j0 * wall_h_div, -1):
        per.append((i0 * wall_h_div, j))
    d = bay_w / wall_h_div
    nodes, at, nid = [], {}, 1
    masters = {}
    levels = []  # (z index in units of story_h / wall_v_div, is floor)
    for k in range(stories + 1):
        levels.append((k * wall_v_div, True))
        if k < stories:
            levels.extend(((k * wall_v_div + m, False) for m in range(1, wall_v_div)))
    for zi, floor in levels:
        z = zi * story_h / wall_v_div
        if floor:
            for j in range(bays + 1):
                for i in range(bays + 1):
                    node = {"id": nid, "x": i * bay_w, "y": j * bay_w, "z": z}
                    if zi == 0:
                        node["constraints"] = [1, 2, 3, 4, 5, 6]
                    nodes.append(node)
                    at[(i * wall_h_div, j * wall_h_div, zi)] = nid
                    nid += 1
        for (pi, pj) in per:
            if (pi, pj, zi) in at:
                continue
            node = {"id": nid, "x": pi * d, "y": pj * d, "z": z}
            if zi == 0:
                node["constraints"] = [1, 2, 3, 4, 5, 6]
            nodes.append(node)
            at[(pi, pj, zi)] = nid
            nid += 1
        if floor and zi > 0:
            nodes.append({"id": nid, "x": bays * bay_w / 2.0, "y": bays * bay_w / 2.0, "z": z,
                          "constraints": [3, 4, 5]})
            masters[zi // wall_v_div] = nid
            nid += 1
    g = lambda i, j, k: at[(i * wall_h_div, j * wall_h_div, k * wall_v_div)]  # noqa: E731
    elements, eid = [], 1
    for k in range(stories):
        for j in range(bays + 1):
            for i in range(bays + 1):
                elements.append({"id": eid, "type": "forceBeamColumn3d", "nodes": [g(i, j, k), g(i, j, k + 1)],
                                 "section": 1, "geomTransf": "PDelta", "vecxz": [1.0, 0.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
    for k in range(1, stories + 1):
        for j in range(bays + 1):
            for i in range(bays):
                elements.append({"id": eid, "type": "dispBeamColumn3d", "nodes": [g(i, j, k), g(i + 1, j, k)],
                                 "section": 2, "geomTransf": "Linear", "vecxz": [0.0, 1.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
        for j in range(bays):
            for i in range(bays + 1):
                elements.append({"id": eid, "type": "dispBeamColumn3d", "nodes": [g(i, j, k), g(i, j + 1, k)],
                                 "section": 2, "geomTransf": "Linear", "vecxz": [1.0, 0.0, 0.0],
                                 "integration": "Lobatto", "num_int_pts": nip})
                eid += 1
    for zi in range(stories * wall_v_div):
        for a in range(len(per)):
            pa, pb = per[a], per[(a + 1) % len(per)]
            elements.append({"id": eid, "type": "shell", "section": 3,
                             "nodes": [at[(pa[0], pa[1], zi)], at[(pb[0], pb[1], zi)],
                                       at[(pb[0], pb[1], zi + 1)], at[(pa[0], pa[1], zi + 1)]]})
            eid += 1
    masses = [{"node": g(i, j, k), "dof": dd, "value": floor_mass}
              for k in range(1, stories + 1) for j in range(bays + 1) for i in range(bays + 1) for dd in (1, 2, 3)]
    coords = {n["id"]: (n["x"], n["y"], n["z"]) for n in nodes}
    mps = []
    for k, m in masters.items():
        floor = [nd["id"] for nd in nodes if nd["z"] == k * story_h and nd["id"] != m]
        mps += rigid_diaphragm(coords, m, floor, 3)
    sections = [rc_section(1, dim=3, b=0.4, h=0.4, n_core=n_core, n_bars=n_bars),
                rc_section(2, dim=3, b=0.3, h=0.5, n_core=n_core, n_bars=n_bars),
                {"id": 3, "type": "ElasticMembranePlateSection",
                 "params": {"E": wall_E, "nu": wall_nu, "h": wall_t, "rho": 0.0}}]
    t1 = 0.1 * stories
    w1, w3 = 2.0 * math.pi / t1, 2.0 * math.pi / (t1 / 5.0)
    zeta = 0.05
    series, patterns = [], []
    for tag, (seed, pga, direction) in enumerate(zip(seeds, pgas_g, (1, 2)), start=1):
        acc = ground_motion(seed, dt, steps, pga, strong=strong)
        series.append({"type": "Path", "tag": tag, "dt": dt, "values": [float(v) for v in acc], "factor": 1.0})
        patterns.append({"type": "UniformExcitation", "tag": tag, "direction": direction, "accel": tag})
    line = [g(0, 0, k) for k in range(stories + 1)]
    recorders = [{"type": "node_displacement", "nodes": [masters[stories]], "dofs": [1, 2, 6], "output": "roof"}]
    for k in range(1, stories + 1):
        for dd in (1, 2):
            recorders.append({"type": "drift", "i_node": line[k - 1], "j_node": line[k], "dof": dd, "perp_dirn": 3,
                              "output": f"drift_{k}_{dd}"})
    recorders.append({"type": "envelope_node_displacement", "nodes": [masters[k] for k in range(1, stories + 1)],
                      "dofs": [1, 2, 6], "output": "floor_envelope"})
    return {
        "schema_version": "1.0",
        "metadata": {"name": name or f"rc_tower_{stories}x{bays}x{bays}_core{core[0]}x{core[1]}", "units": "kN-m-t-s"},
        "model": {"ndm": 3, "ndf": 6},
        "nodes": nodes,
        "materials": rc_materials(),
        "sections": sections,
        "elements": elements,
        "masses": masses,
        "mp_constraints": mps,
        "time_series": series,
        "patterns": patterns,
        "rayleigh": {"alphaM": 2.0 * zeta * w1 * w3 / (w1 + w3), "betaKInit": 2.0 * zeta / (w1 + w3)},
        "analysis": {"type": "transient_nonlinear", "steps": steps, "dt": dt, "tol": tol, "max_iters": max_iters,
                     "system": "BandGeneral", "numberer": "Plain", "constraints": "Transformation",
                     "algorithm": "Newton", "test_type": "NormUnbalance",
                     "integrator": {"type": "Newmark", "gamma": 0.5, "beta": 0.25}},
        "recorders": recorders,
    }


def config4_case(steps: int = 6000, seeds: Tuple[int, int] = (0, 1), pgas_g: Tuple[float, float] = (0.6, 0.42)) -> dict:
    """The large single-GPU case: 60 stories x 6 x 6 bays (2940 forceBeamColumn3d
    columns, 5040 dispBeamColumn3d beams, 5 IPs, 152-fiber sections: about
    6.1 M fibers) around a 2 x 2-bay elastic shell core (each bay face in 6
    panels, each story in 2: 5760 shells), rigid floors, two records,
    dt 0.005 s: 33,480 free DOFs, half bandwidth 560."""
    return rc_tower_case(60, 6, steps=steps, dt=0.005, core=(2, 2), wall_h_div=6, wall_v_div=2, seeds=seeds,
                         pgas_g=pgas_g)

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

import math
from typing import Dict, List, Tuple

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

"""Case-schema pieces of the fiber-section path: uniaxial material
normalisation and fiber discretisation.

Both follow the reference's loader exactly so the definitions handed to the
device library are the ones Strut builds from the same JSON:

* materials: ``src/mojo/solver/run_case/loader.mojo:915-1096`` (defaults,
  sign normalisation, parameter-group rules and aborts become ValueError);
* fibers: rect/quadr patches, straight layers and inline fibers in
  ``src/mojo/sections/fiber2d.mojo:3504-3878`` (2D) and
  ``src/mojo/sections/fiber3d.mojo:258-470`` (3D).  Cell order and the
  floating-point expressions are the reference's, because the section
  centroid and every fiber coordinate enter the reductions.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy as np

# tag_types.mojo:136-142
ELASTIC, STEEL01, CONCRETE01, STEEL02, CONCRETE02 = 0, 1, 2, 3, 4
TYPE_TAGS = {
    "Elastic": ELASTIC,
    "Steel01": STEEL01,
    "Concrete01": CONCRETE01,
    "Steel02": STEEL02,
    "Concrete02": CONCRETE02,
}


@dataclass
class UniMaterialDef:
    """UniMaterialDef (materials/uniaxial/core.mojo:48-118): tag + p0..p10."""

    mat_type: int
    p: Tuple[float, ...]


def make_uniaxial_def(mat: dict) -> UniMaterialDef:
    """Normalise one ``materials[]`` entry as loader.mojo:915-1096 does."""
    mtype = mat["type"]
    prm = mat.get("params", {})
    has = lambda k: k in prm  # noqa: E731
    if mtype == "Elastic":
        E = float(prm["E"])
        if E <= 0.0:
            raise ValueError("Elastic material E must be > 0")
        return UniMaterialDef(ELASTIC, (E,) + (0.0,) * 10)
    if mtype == "Steel01":
        Fy, E0, b = float(prm["Fy"]), float(prm["E0"]), float(prm["b"])
        if Fy <= 0.0:
            raise ValueError("Steel01 Fy must be > 0")
        if E0 <= 0.0:
            raise ValueError("Steel01 E0 must be > 0")
        if b < 0.0 or b >= 1.0:
            raise ValueError("Steel01 b must be in [0, 1)")
        return UniMaterialDef(STEEL01, (Fy, E0, b) + (0.0,) * 8)
    if mtype in ("Concrete01", "Concrete02"):
        fpc, epsc0 = float(prm["fpc"]), float(prm["epsc0"])
        fpcu, epscu = float(prm["fpcu"]), float(prm["epscu"])
        if fpc > 0.0:
            fpc = -fpc
        if epsc0 > 0.0:
            epsc0 = -epsc0
        if fpcu > 0.0:
            fpcu = -fpcu
        if epscu > 0.0:
            epscu = -epscu
        if fpc >= 0.0:
            raise ValueError(f"{mtype} fpc must be < 0")
        if epsc0 >= 0.0:
            raise ValueError(f"{mtype} epsc0 must be < 0")
        if epscu >= 0.0:
            raise ValueError(f"{mtype} epscu must be < 0")
        if epscu >= epsc0:
            raise ValueError(f"{mtype} epscu must be < epsc0")
        if fpcu > 0.0 or fpcu < fpc:
            raise ValueError(f"{mtype} fpcu must be between fpc and 0")
        if mtype == "Concrete01":
            return UniMaterialDef(CONCRETE01, (fpc, epsc0, fpcu, epscu) + (0.0,) * 7)
        if has("rat") != has("ft") or has("rat") != has("Ets"):
            raise ValueError("Concrete02 requires rat, ft, Ets together")
        rat = 0.1
        ft = 0.1 * fpc
        if ft < 0.0:
            ft = -ft
        Ets = 0.1 * fpc / epsc0
        if has("rat"):
            rat, ft, Ets = float(prm["rat"]), float(prm["ft"]), float(prm["Ets"])
        if rat == 1.0:
            raise ValueError("Concrete02 rat must not be 1")
        if ft < 0.0:
            raise ValueError("Concrete02 ft must be >= 0")
        if Ets <= 0.0:
            raise ValueError("Concrete02 Ets must be > 0")
        return UniMaterialDef(CONCRETE02, (fpc, epsc0, fpcu, epscu, rat, ft, Ets, 0.0, 0.0, 0.0, 0.0))
    if mtype == "Steel02":
        Fy, E0, b = float(prm["Fy"]), float(prm["E0"]), float(prm["b"])
        if Fy < 0.0:
            raise ValueError("Steel02 Fy must be >= 0")
        if E0 <= 0.0:
            raise ValueError("Steel02 E0 must be > 0")
        if b < 0.0 or b >= 1.0:
            raise ValueError("Steel02 b must be in [0, 1)")
        hr, hc1, hc2 = has("R0"), has("cR1"), has("cR2")
        ha = [has(k) for k in ("a1", "a2", "a3", "a4")]
        if hr != hc1 or hr != hc2:
            raise ValueError("Steel02 requires R0, cR1, cR2 together")
        if any(x != ha[0] for x in ha):
            raise ValueError("Steel02 requires a1, a2, a3, a4 together")
        if ha[0] and not hr:
            raise ValueError("Steel02 a1-a4 require R0, cR1, cR2")
        if has("sigInit") and not ha[0]:
            raise ValueError("Steel02 sigInit requires a1-a4 and R0/cR1/cR2")
        R0, cR1, cR2 = 15.0, 0.925, 0.15
        if hr:
            R0, cR1, cR2 = float(prm["R0"]), float(prm["cR1"]), float(prm["cR2"])
        if R0 <= 0.0:
            raise ValueError("Steel02 R0 must be > 0")
        if cR2 <= 0.0:
            raise ValueError("Steel02 cR2 must be > 0")
        a1, a2, a3, a4 = 0.0, 1.0, 0.0, 1.0
        if ha[0]:
            a1, a2, a3, a4 = (float(prm[k]) for k in ("a1", "a2", "a3", "a4"))
        if a2 <= 0.0:
            raise ValueError("Steel02 a2 must be > 0")
        if a4 <= 0.0:
            raise ValueError("Steel02 a4 must be > 0")
        sig_init = float(prm["sigInit"]) if has("sigInit") else 0.0
        return UniMaterialDef(STEEL02, (Fy, E0, b, R0, cR1, cR2, a1, a2, a3, a4, sig_init))
    raise ValueError(f"unsupported uniaxial material: {mtype}")


class MaterialTable:
    """Material id -> definition index table (loader.mojo:905-915)."""

    def __init__(self, materials: Sequence[dict]):
        self.defs: List[UniMaterialDef] = []
        self.index_by_id: Dict[int, int] = {}
        for mat in materials:
            if mat["type"] not in TYPE_TAGS:
                continue
            self.index_by_id[int(mat["id"])] = len(self.defs)
            self.defs.append(make_uniaxial_def(mat))

    def arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        types = np.array([d.mat_type for d in self.defs], dtype=np.int32)
        params = np.array([list(d.p) for d in self.defs], dtype=np.float64).reshape(-1, 11)
        return types, params

    def resolve(self, mat_id: int) -> int:
        # fiber2d.mojo:3131-3138 (_resolve_uniaxial_def_index)
        if mat_id < 0:
            raise ValueError("FiberSection material id must be >= 0")
        if mat_id not in self.index_by_id:
            raise ValueError("FiberSection requires uniaxial material for all fibers")
        return self.index_by_id[mat_id]


def _rect_cells(p: dict, mat: int, out: list) -> None:
    """fiber2d.mojo:3504-3550 / fiber3d.mojo:258-307."""
    ny, nz = int(p["num_subdiv_y"]), int(p["num_subdiv_z"])
    if ny <= 0 or nz <= 0:
        raise ValueError("patch rect requires num_subdiv_y and num_subdiv_z > 0")
    y_i, z_i, y_j, z_j = float(p["y_i"]), float(p["z_i"]), float(p["y_j"]), float(p["z_j"])
    y_min, y_max = (y_j, y_i) if y_i > y_j else (y_i, y_j)
    z_min, z_max = (z_j, z_i) if z_i > z_j else (z_i, z_j)
    dy_total, dz_total = y_max - y_min, z_max - z_min
    if dy_total <= 0.0 or dz_total <= 0.0:
        raise ValueError("patch rect must have non-zero side lengths")
    dy, dz = dy_total / float(ny), dz_total / float(nz)
    area = dy * dz
    for iy in range(ny):
        y = y_min + (float(iy) + 0.5) * dy
        for iz in range(nz):
            z = z_min + (float(iz) + 0.5) * dz
            out.append((y, z, area, mat))


def _bilinear(u: float, v: float, a: float, b: float, c: float, d: float) -> float:
    return (1.0 - u) * (1.0 - v) * a + u * (1.0 - v) * b + u * v * c + (1.0 - u) * v * d


def _quadr_cells(p: dict, mat: int, out: list) -> None:
    """fiber2d.mojo:3553-3610 / fiber3d.mojo:310-370."""
    ny, nz = int(p["num_subdiv_y"]), int(p["num_subdiv_z"])
    if ny <= 0 or nz <= 0:
        raise ValueError("patch quadr requires num_subdiv_y and num_subdiv_z > 0")
    ys = [float(p[k]) for k in ("y_i", "y_j", "y_k", "y_l")]
    zs = [float(p[k]) for k in ("z_i", "z_j", "z_k", "z_l")]
    for iy in range(ny):
        u0, u1 = float(iy) / float(ny), float(iy + 1) / float(ny)
        uc = 0.5 * (u0 + u1)
        for iz in range(nz):
            v0, v1 = float(iz) / float(nz), float(iz + 1) / float(nz)
            vc = 0.5 * (v0 + v1)
            y00, z00 = _bilinear(u0, v0, *ys), _bilinear(u0, v0, *zs)
            y10, z10 = _bilinear(u1, v0, *ys), _bilinear(u1, v0, *zs)
            y11, z11 = _bilinear(u1, v1, *ys), _bilinear(u1, v1, *zs)
            y01, z01 = _bilinear(u0, v1, *ys), _bilinear(u0, v1, *zs)
            twice = (y00 * z10 + y10 * z11 + y11 * z01 + y01 * z00
                     - z00 * y10 - z10 * y11 - z11 * y01 - z01 * y00)
            area = abs(twice) * 0.5
            if area <= 0.0:
                raise ValueError("patch quadr generated zero-area cell")
            out.append((_bilinear(uc, vc, *ys), _bilinear(uc, vc, *zs), area, mat))


def _straight_cells(p: dict, mat: int, out: list) -> None:
    """fiber2d.mojo:3613-3648 / fiber3d.mojo:373-412."""
    n = int(p["num_bars"])
    if n <= 0:
        raise ValueError("layer straight requires num_bars > 0")
    a = float(p["bar_area"])
    if a <= 0.0:
        raise ValueError("layer straight requires bar_area > 0")
    ys, zs, ye, ze = float(p["y_start"]), float(p["z_start"]), float(p["y_end"]), float(p["z_end"])
    if n == 1:
        out.append((0.5 * (ys + ye), 0.5 * (zs + ze), a, mat))
        return
    for i in range(n):
        t = float(i) / float(n - 1)
        out.append((ys + (ye - ys) * t, zs + (ze - zs) * t, a, mat))


def discretize_section(section: dict, materials: MaterialTable) -> np.ndarray:
    """Fiber cells of a FiberSection2d/3d in the reference's order: patches,
    then layers, then inline fibers (fiber2d.mojo:3881-3943).  Returns a
    structured array with fields y, z, area, mat (definition index)."""
    params = section.get("params", {})
    cells: list = []
    patches, layers, fibers = params.get("patches", []), params.get("layers", []), params.get("fibers", [])
    if not patches and not layers and not fibers:
        raise ValueError("FiberSection requires at least one patch, layer, or fiber")
    for patch in patches:
        ptype = patch["type"]
        if ptype == "quad":
            ptype = "quadr"
        mat = materials.resolve(int(patch["material"]))
        if ptype == "rect":
            _rect_cells(patch, mat, cells)
        elif ptype == "quadr":
            _quadr_cells(patch, mat, cells)
        else:
            raise ValueError(f"unsupported FiberSection patch type: {ptype}")
    for layer in layers:
        if layer["type"] != "straight":
            raise ValueError(f"unsupported FiberSection layer type: {layer['type']}")
        _straight_cells(layer, materials.resolve(int(layer["material"])), cells)
    for f in fibers:
        area = float(f["area"])
        if area <= 0.0:
            raise ValueError("FiberSection fiber area must be > 0")
        cells.append((float(f["y"]), float(f.get("z", 0.0)), area, materials.resolve(int(f["material"]))))
    out = np.zeros(len(cells), dtype=[("y", "f8"), ("z", "f8"), ("area", "f8"), ("mat", "i4")])
    for i, c in enumerate(cells):
        out[i] = c
    return out


# ---------------------------------------------------------------- beam integration
def _legendre_p(n: int, x: float):
    """(P_n, P_{n-1}, P'_n) as beam_integration.mojo:21-45 computes them."""
    if n == 0:
        return 1.0, 0.0, 0.0
    if n == 1:
        return x, 1.0, 1.0
    pnm2, pnm1 = 1.0, x
    for k in range(2, n + 1):
        kf = float(k)
        pk = ((2.0 * kf - 1.0) * x * pnm1 - (kf - 1.0) * pnm2) / kf
        pnm2, pnm1 = pnm1, pk
    pn, pn1 = pnm1, pnm2
    denom = x * x - 1.0
    if abs(denom) <= 1.0e-18:
        x_off = x - 1.0e-12 if x >= 0.0 else x + 1.0e-12
        return pn, pn1, _legendre_p(n, x_off)[2]
    return pn, pn1, float(n) * (x * pn - pn1) / denom


def beam_integration_rule(integration: str, n: int):
    """Integration points on [0, 1] and weights (halved), as
    beam_integration_rule (beam_integration.mojo:224-241) builds them: the
    3- and 5-point Lobatto rules are the decimal literals of :63-91."""
    if integration == "Lobatto":
        if n < 2:
            raise ValueError("Lobatto integration requires num_int_pts >= 2")
        if n == 3:
            x, w = [-1.0, 0.0, 1.0], [1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0]
        elif n == 5:
            x = [-1.0, -0.6546536707079772, 0.0, 0.6546536707079772, 1.0]
            w = [0.1, 0.5444444444444444, 0.7111111111111111, 0.5444444444444444, 0.1]
        else:
            m, nf = n - 1, float(n)
            x, w = [0.0] * n, [0.0] * n
            x[0], x[-1] = -1.0, 1.0
            w[0] = 2.0 / (nf * (nf - 1.0))
            w[-1] = w[0]
            for i in range(1, n - 1):
                root = math.cos(3.141592653589793 * float(i) / float(n - 1))
                for _ in range(80):
                    pm, _, dpm = _legendre_p(m, root)
                    denom = 1.0 - root * root
                    if abs(denom) <= 1.0e-18:
                        break
                    ddpm = (2.0 * root * dpm - float(m * (m + 1)) * pm) / denom
                    if abs(ddpm) <= 1.0e-24:
                        break
                    delta = dpm / ddpm
                    root -= delta
                    if abs(delta) <= 1.0e-14:
                        break
                x[i] = root
                pm_root = _legendre_p(m, root)[0]
                w[i] = 2.0 / (nf * (nf - 1.0) * pm_root * pm_root)
            order = sorted(range(n), key=lambda k: x[k])
            x, w = [x[k] for k in order], [w[k] for k in order]
    elif integration == "Legendre":
        if n < 1:
            raise ValueError("Legendre integration requires num_int_pts >= 1")
        if n == 1:
            x, w = [0.0], [2.0]
        else:
            x, w = [0.0] * n, [0.0] * n
            for i in range((n + 1) // 2):
                root = math.cos(3.141592653589793 * (float(i) + 0.75) / (float(n) + 0.5))
                dpn = 0.0
                for _ in range(80):
                    pn, _, dpn = _legendre_p(n, root)
                    if abs(dpn) <= 1.0e-24:
                        break
                    delta = pn / dpn
                    root -= delta
                    if abs(delta) <= 1.0e-14:
                        break
                one_minus = 1.0 - root * root
                weight = 2.0 / (one_minus * dpn * dpn)
                x[i], w[i] = -root, weight
                x[n - 1 - i], w[n - 1 - i] = root, weight
            order = sorted(range(n), key=lambda k: x[k])
            x, w = [x[k] for k in order], [w[k] for k in order]
    else:
        raise ValueError(f"unsupported beam integration: {integration}")
    return [0.5 * (xi + 1.0) for xi in x], [wi * 0.5 for wi in w]


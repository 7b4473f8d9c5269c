"""Load a 3D fiber-frame transient case (Strut's JSON schema) into flat
arrays: the part of solver/run_case/loader.mojo:758-2460 the 3D
fiber beam-column transient path needs -- node index map, plain DOF
numbering (6 DOFs per node), single-point constraints, lumped masses, the
Transformation handler's rigidDiaphragm / equalDOF expansion tables
(loader.mojo:2230-2372), forceBeamColumn3d / dispBeamColumn3d elements on
FiberSection3d sections (geomTransf Linear or PDelta, optional vecxz, elastic
torsion G J from the section's params), elastic four-node ``shell`` elements
on ElasticMembranePlateSection sections (E, nu, h; rho lumps rho h area / 4
on ux, uy, uz of each node, loader.mojo:2060-2100) -- and the analysis /
pattern / Rayleigh records of input_types.mojo:1821-2290.

Excitation: one ``pattern`` (Strut's schema) or a ``patterns`` list of
UniformExcitation records whose inertia loads are summed in order (the
bidirectional extension: two horizontal records at once).

Cases outside that path raise ValueError rather than run differently.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .constraints import mpc_tables
from .model import MaterialTable, beam_integration_rule, discretize_section

MAX_NIP = 10
NDF = 6
ELEMENT_TYPES_3D = {"forceBeamColumn3d": 0, "dispBeamColumn3d": 1}
TRANSFS = {"Linear": 0, "PDelta": 1}


@dataclass
class Pattern:
    direction: int           # 1-based global direction
    values: np.ndarray       # Path values
    dt: float
    factor: float


@dataclass
class Frame3dCase:
    n_nodes: int
    coords: np.ndarray               # (n_nodes, 3)
    sp: np.ndarray                   # 6 n_nodes int32 single-point flags
    mass: np.ndarray                 # 6 n_nodes
    materials: MaterialTable
    section_ids: List[int]
    section_cells: Dict[int, np.ndarray]
    section_gj: Dict[int, tuple]     # section id -> (G, J)
    elem_nodes: np.ndarray           # (n_elem, 2) node indices
    elem_kind: np.ndarray            # 0 forceBeamColumn3d, 1 dispBeamColumn3d
    elem_section: np.ndarray
    elem_transf: np.ndarray          # 0 Linear, 1 PDelta
    elem_vecxz: np.ndarray           # (n_elem, 3); NaN row = none given
    elem_nip: np.ndarray
    elem_xi: np.ndarray              # (n_elem, MAX_NIP)
    elem_w: np.ndarray
    has_mpc: bool
    row_offsets: np.ndarray
    dof_pool: np.ndarray
    coeff_pool: np.ndarray
    slave: np.ndarray                # bool per DOF
    constrained: np.ndarray          # bool per DOF (single-point or slave)
    dt: float
    steps: int
    gamma: float
    beta: float
    tol: float
    max_iters: int
    alpha_m: float
    beta_k: float
    beta_k_init: float
    patterns: List[Pattern] = field(default_factory=list)
    record_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    envelope_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    drift: List[tuple] = field(default_factory=list)  # (i_dof, j_dof, dx) per drift recorder
    shell_nodes: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), dtype=np.int32))
    shell_props: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))  # E, nu, h per shell

    @property
    def total_dofs(self) -> int:
        return NDF * self.n_nodes

    @property
    def n_elem(self) -> int:
        return len(self.elem_kind)


def load_frame3d_case(case: dict, steps: int | None = None) -> Frame3dCase:
    if case.get("model", {}).get("ndm") != 3 or case.get("model", {}).get("ndf") != NDF:
        raise ValueError("3D frame cases need ndm = 3, ndf = 6")
    nodes = case["nodes"]
    idx = {int(n["id"]): i for i, n in enumerate(nodes)}
    n_nodes = len(nodes)
    coords = np.array([[float(n["x"]), float(n["y"]), float(n.get("z", 0.0))] for n in nodes])
    sp = np.zeros(NDF * n_nodes, dtype=np.int32)
    for i, n in enumerate(nodes):
        for d in n.get("constraints", []):
            sp[NDF * i + int(d) - 1] = 1
    mass = np.zeros(NDF * n_nodes)
    for m in case.get("masses", []):
        mass[NDF * idx[int(m["node"])] + int(m["dof"]) - 1] += float(m["value"])
    mats = MaterialTable(case["materials"])
    sections = {int(s["id"]): s for s in case["sections"]}
    shells = [e for e in case["elements"] if e["type"] == "shell"]
    elems = [e for e in case["elements"] if e["type"] != "shell"]
    used: List[int] = []
    for e in elems:
        if e["type"] not in ELEMENT_TYPES_3D:
            raise ValueError(f"element type {e['type']} is outside the 3D fiber-frame path")
        if e.get("geomTransf", "Linear") not in TRANSFS:
            raise ValueError(f"geomTransf {e.get('geomTransf')} is outside the 3D fiber-frame path")
        sid = int(e["section"])
        if sections[sid]["type"] != "FiberSection3d":
            raise ValueError(f"{e['type']} needs a FiberSection3d on the 3D fiber-frame path")
        if sid not in used:
            used.append(sid)
    shell_nodes = np.zeros((len(shells), 4), dtype=np.int32)
    shell_props = np.zeros((len(shells), 3))
    for k, e in enumerate(shells):
        sec = sections[int(e["section"])]
        if sec["type"] != "ElasticMembranePlateSection":  # layered shells are outside the path
            raise ValueError(f"shell {e.get('id')} needs an ElasticMembranePlateSection on the 3D frame path")
        if len(e["nodes"]) != 4:
            raise ValueError("shell requires 4 nodes")
        pr = sec["params"]
        shell_nodes[k] = [idx[int(v)] for v in e["nodes"]]
        shell_props[k] = [float(pr.get("E", 0.0)), float(pr.get("nu", 0.0)), float(pr.get("h", 0.0))]
        rho_area = float(pr.get("rho", 0.0)) * float(pr.get("h", 0.0))
        if rho_area != 0.0:  # loader.mojo:2060-2100
            p = coords[shell_nodes[k]]
            a = (p[1] - p[0]) + (p[2] - p[3])
            b = (p[3] - p[0]) + (p[2] - p[1])
            cr = (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])
            area = 0.25 * float(np.sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]))
            if area <= 0.0:
                raise ValueError("shell area must be > 0")
            lumped = rho_area * area / 4.0
            for d in range(3):
                for nd in shell_nodes[k]:
                    mass[NDF * nd + d] += lumped
    cells = {sid: discretize_section(sections[sid], mats) for sid in used}
    gj = {sid: (float(sections[sid]["params"].get("G", 0.0)), float(sections[sid]["params"].get("J", 0.0)))
          for sid in used}
    ne = len(elems)
    elem_nodes = np.array([[idx[int(e["nodes"][0])], idx[int(e["nodes"][1])]] for e in elems], dtype=np.int32)
    elem_kind = np.array([ELEMENT_TYPES_3D[e["type"]] for e in elems], dtype=np.int32)
    elem_section = np.array([int(e["section"]) for e in elems], dtype=np.int32)
    elem_transf = np.array([TRANSFS[e.get("geomTransf", "Linear")] for e in elems], dtype=np.int32)
    elem_vecxz = np.full((ne, 3), np.nan)
    for k, e in enumerate(elems):
        if "vecxz" in e:
            elem_vecxz[k] = [float(v) for v in e["vecxz"]]
    elem_nip = np.array([int(e.get("num_int_pts", 3)) for e in elems], dtype=np.int32)
    elem_xi = np.zeros((ne, MAX_NIP))
    elem_w = np.zeros((ne, MAX_NIP))
    for k, e in enumerate(elems):
        xi, w = beam_integration_rule(e.get("integration", "Lobatto"), int(elem_nip[k]))
        if len(xi) > 5:
            raise ValueError("3D fiber elements take at most 5 integration points on this path")
        elem_xi[k, :len(xi)] = xi
        elem_w[k, :len(w)] = w
    an = case["analysis"]
    if an["type"] != "transient_nonlinear":
        raise ValueError("only transient_nonlinear analyses are on the 3D fiber-frame path")
    for key, allowed in (("system", ("BandGeneral",)), ("constraints", ("Plain", "Transformation")),
                         ("algorithm", ("Newton",)), ("test_type", ("NormUnbalance",))):
        if an.get(key, allowed[0]) not in allowed:
            raise ValueError(f"analysis {key} {an.get(key)} is outside the 3D fiber-frame path")
    if "solver_chain" in an:  # the reference retries steps over the chain (transient_nonlinear.mojo:758-827)
        raise ValueError("analysis solver_chain is outside the 3D fiber-frame path")
    for key in ("loads", "element_loads", "dampings"):  # nodal / element loads and damping records
        if case.get(key):
            raise ValueError(f"{key} are outside the 3D fiber-frame path")
    mps = case.get("mp_constraints", [])
    if mps and an.get("constraints", "Plain") != "Transformation":
        raise ValueError("mp_constraints need the Transformation constraint handler on this path")
    ro, dp, cp, slave, constrained = mpc_tables(n_nodes, NDF, sp.astype(bool), mps, idx)
    integ = an.get("integrator", {})
    ray = case.get("rayleigh", {})
    if float(ray.get("betaKComm", 0.0)) != 0.0:
        raise ValueError("betaKComm is outside the 3D fiber-frame path")
    ts = {int(t["tag"]): t for t in case.get("time_series", [])}
    pats = case.get("patterns", [case["pattern"]] if "pattern" in case else [])
    patterns = []
    for pat in pats:
        if pat.get("type") != "UniformExcitation":
            raise ValueError("only UniformExcitation patterns are on the 3D fiber-frame path")
        series = ts[int(pat["accel"])]
        if series["type"] != "Path" or "dt" not in series or "time" in series:
            raise ValueError("the excitation must be a Path time series with dt")
        if float(series.get("start_time", 0.0)) != 0.0 or series.get("use_last", False):
            raise ValueError("Path start_time / use_last are outside the 3D fiber-frame path")
        patterns.append(Pattern(int(pat["direction"]), np.array(series["values"], dtype=np.float64),
                                float(series["dt"]), float(series.get("factor", 1.0))))
    if len(patterns) > 3:
        raise ValueError("at most three excitation patterns")
    rec, env, drift = [], [], []
    for r in case.get("recorders", []):
        if r["type"] in ("node_displacement", "envelope_node_displacement"):
            for nid in r["nodes"]:
                for d in r["dofs"]:
                    (rec if r["type"] == "node_displacement" else env).append(NDF * idx[int(nid)] + int(d) - 1)
        elif r["type"] == "drift":  # helpers.mojo:3113-3142
            i, j = idx[int(r["i_node"])], idx[int(r["j_node"])]
            d, perp = int(r["dof"]), int(r["perp_dirn"])
            if not 1 <= d <= NDF or not 1 <= perp <= 3:
                raise ValueError("drift recorder dof / perp_dirn out of range")
            dx = float(coords[j, perp - 1] - coords[i, perp - 1])
            if dx == 0.0:
                raise ValueError("drift denominator is zero")
            drift.append((NDF * i + d - 1, NDF * j + d - 1, dx))
    return Frame3dCase(
        n_nodes=n_nodes, coords=coords, sp=sp, mass=mass, materials=mats, section_ids=used, section_cells=cells,
        section_gj=gj, elem_nodes=elem_nodes, elem_kind=elem_kind, elem_section=elem_section,
        elem_transf=elem_transf, elem_vecxz=elem_vecxz, elem_nip=elem_nip, elem_xi=elem_xi, elem_w=elem_w,
        has_mpc=bool(mps), row_offsets=ro, dof_pool=dp, coeff_pool=cp, slave=slave, constrained=constrained,
        dt=float(an["dt"]), steps=int(steps if steps is not None else an["steps"]),
        gamma=float(integ.get("gamma", 0.5)), beta=float(integ.get("beta", 0.25)), tol=float(an.get("tol", 1.0e-10)),
        max_iters=int(an.get("max_iters", 20)), alpha_m=float(ray.get("alphaM", 0.0)),
        beta_k=float(ray.get("betaK", 0.0)), beta_k_init=float(ray.get("betaKInit", 0.0)), patterns=patterns,
        record_dofs=np.array(rec, dtype=np.int32), envelope_dofs=np.array(env, dtype=np.int32), drift=drift,
        shell_nodes=shell_nodes, shell_props=shell_props)


def recorder_plan(fc: Frame3dCase):
    """The device recorder of a case's envelope and drift recorders: the DOFs
    to record (envelope DOFs first, then the drift recorders' DOFs not
    already there) and each drift channel as (column i, column j, dx) over
    them."""
    cols = [int(d) for d in fc.envelope_dofs]
    col_i, col_j, dx = [], [], []
    for i_dof, j_dof, d in fc.drift:
        for dof in (i_dof, j_dof):
            if dof not in cols:
                cols.append(dof)
        col_i.append(cols.index(i_dof))
        col_j.append(cols.index(j_dof))
        dx.append(d)
    return np.array(cols, dtype=np.int32), np.array(col_i, dtype=np.int32), np.array(col_j, dtype=np.int32), \
        np.array(dx, dtype=np.float64)

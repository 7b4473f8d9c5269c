"""Load a 2D fiber-frame transient case (Strut's JSON schema) into flat
arrays: the part of solver/run_case/loader.mojo:758-2460 the
forceBeamColumn2d transient path needs (node index map, plain DOF
numbering, constraints, lumped masses, per-element section instances) and
the analysis / pattern / Rayleigh records of input_types.mojo:1821-2290.

Cases outside that path (other element types, MPC constraints, other
algorithms or systems) raise ValueError rather than run differently.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .model import MaterialTable, beam_integration_rule, discretize_section

MAX_NIP = 10


@dataclass
class Frame2dCase:
    n_nodes: int
    x: np.ndarray
    y: np.ndarray
    constrained: np.ndarray          # 3 * n_nodes int32
    mass: np.ndarray                 # 3 * n_nodes
    materials: MaterialTable
    section_ids: List[int]           # section definition order
    section_cells: Dict[int, np.ndarray]
    elem_nodes: np.ndarray           # (n_elem, 2) int32 node indices
    elem_section: np.ndarray         # section id per element
    elem_nip: np.ndarray
    elem_xi: np.ndarray              # (n_elem, MAX_NIP)
    elem_w: np.ndarray
    dt: float
    steps: int
    gamma: float
    beta: float
    tol: float
    max_iters: int
    alpha_m: float
    beta_k: float
    beta_k_init: float
    exc_dir: int
    ts_values: np.ndarray
    ts_dt: float
    ts_factor: float
    ts_start: float
    ts_use_last: int
    record_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    elem_type: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))  # 0 FBC, 1 DBC

    @property
    def total_dofs(self) -> int:
        return 3 * self.n_nodes


# Element types of the fiber-frame path and their codes in the device tables.
ELEMENT_TYPES = {"forceBeamColumn2d": 0, "dispBeamColumn2d": 1}


def load_frame2d_case(case: dict, steps: int | None = None) -> Frame2dCase:
    if case.get("model", {}).get("ndm") != 2 or case.get("model", {}).get("ndf") != 3:
        raise ValueError("2D frame cases need ndm = 2, ndf = 3")
    nodes = case["nodes"]
    idx = {int(n["id"]): i for i, n in enumerate(nodes)}
    n_nodes = len(nodes)
    x = np.array([float(n["x"]) for n in nodes])
    y = np.array([float(n["y"]) for n in nodes])
    constrained = np.zeros(3 * n_nodes, dtype=np.int32)
    for i, n in enumerate(nodes):
        for d in n.get("constraints", []):
            constrained[3 * i + int(d) - 1] = 1
    mass = np.zeros(3 * n_nodes)
    for m in case.get("masses", []):
        mass[3 * idx[int(m["node"])] + int(m["dof"]) - 1] += float(m["value"])
    mats = MaterialTable(case["materials"])
    sections = {int(s["id"]): s for s in case["sections"]}
    elems = case["elements"]
    used = []
    for e in elems:
        if e["type"] not in ELEMENT_TYPES:
            raise ValueError(f"element type {e['type']} is outside the fiber-frame path")
        if e.get("geomTransf", "Linear") != "Linear":
            raise ValueError("only the Linear geometric transformation is on the fiber-frame path")
        sid = int(e["section"])
        if sections[sid]["type"] != "FiberSection2d":
            raise ValueError(f"{e['type']} needs a FiberSection2d on the fiber-frame path")
        if sid not in used:
            used.append(sid)
    cells = {sid: discretize_section(sections[sid], mats) for sid in used}
    ne = len(elems)
    elem_nodes = np.array([[idx[int(e["nodes"][0])], idx[int(e["nodes"][1])]] for e in elems], dtype=np.int32)
    elem_section = np.array([int(e["section"]) for e in elems], dtype=np.int32)
    elem_nip = np.array([int(e.get("num_int_pts", 5)) for e in elems], dtype=np.int32)
    elem_type = np.array([ELEMENT_TYPES[e["type"]] for e in elems], dtype=np.int32)
    elem_xi = np.zeros((ne, MAX_NIP))
    elem_w = np.zeros((ne, MAX_NIP))
    for k, e in enumerate(elems):
        xi, w = beam_integration_rule(e.get("integration", "Lobatto"), int(elem_nip[k]))
        if len(xi) > MAX_NIP:
            raise ValueError("too many integration points")
        elem_xi[k, :len(xi)] = xi
        elem_w[k, :len(w)] = w
    an = case["analysis"]
    if an["type"] != "transient_nonlinear":
        raise ValueError("only transient_nonlinear analyses are on the fiber-frame path")
    for key, allowed in (("system", ("BandGeneral",)), ("constraints", ("Plain",)), ("algorithm", ("Newton",)),
                         ("test_type", ("NormUnbalance",))):
        if an.get(key, allowed[0]) not in allowed:
            raise ValueError(f"analysis {key} {an.get(key)} is outside the fiber-frame path")
    if "solver_chain" in an:  # the reference retries steps over the chain (transient_nonlinear.mojo:758-827)
        raise ValueError("analysis solver_chain is outside the fiber-frame path")
    for key in ("loads", "element_loads", "dampings"):  # nodal / element loads and damping records
        if case.get(key):
            raise ValueError(f"{key} are outside the fiber-frame path")
    integ = an.get("integrator", {})
    ray = case.get("rayleigh", {})
    if float(ray.get("betaKComm", 0.0)) != 0.0:
        raise ValueError("betaKComm is outside the fiber-frame path")
    pat = case.get("pattern", {})
    ts_values, ts_dt, ts_factor, ts_start, ts_use_last, exc_dir = np.zeros(0), 1.0, 1.0, 0.0, 0, 0
    if pat:
        if pat.get("type") != "UniformExcitation":
            raise ValueError("only UniformExcitation patterns are on the fiber-frame path")
        exc_dir = int(pat["direction"])
        ts = {int(t["tag"]): t for t in case.get("time_series", [])}[int(pat["accel"])]
        if ts["type"] not in ("Path", "PathFile") or "dt" not in ts or "time" in ts:
            raise ValueError("the excitation must be a Path time series with dt")
        ts_values = np.array(ts["values"], dtype=np.float64)
        ts_dt = float(ts["dt"])
        ts_factor = float(ts.get("factor", 1.0))
        ts_start = float(ts.get("start_time", 0.0))
        ts_use_last = 1 if ts.get("use_last", False) else 0
    rec = []
    for r in case.get("recorders", []):
        if r["type"] == "node_displacement":
            for nid in r["nodes"]:
                for d in r["dofs"]:
                    rec.append(3 * idx[int(nid)] + int(d) - 1)
    return Frame2dCase(
        n_nodes=n_nodes, x=x, y=y, constrained=constrained, mass=mass, materials=mats, section_ids=used,
        section_cells=cells, elem_nodes=elem_nodes, elem_section=elem_section, elem_nip=elem_nip,
        elem_xi=elem_xi, elem_w=elem_w, dt=float(an["dt"]), steps=int(steps if steps is not None else an["steps"]),
        gamma=float(integ.get("gamma", 0.5)), beta=float(integ.get("beta", 0.25)), tol=float(an.get("tol", 1.0e-10)),
        max_iters=int(an.get("max_iters", 20)), alpha_m=float(ray.get("alphaM", 0.0)),
        beta_k=float(ray.get("betaK", 0.0)), beta_k_init=float(ray.get("betaKInit", 0.0)), exc_dir=exc_dir,
        ts_values=ts_values, ts_dt=ts_dt, ts_factor=ts_factor, ts_start=ts_start, ts_use_last=ts_use_last,
        record_dofs=np.array(rec, dtype=np.int32), elem_type=elem_type)

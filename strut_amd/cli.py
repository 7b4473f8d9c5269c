"""The solver binary's process boundary over the device path.

``strut`` honours the reference binary's command line (strut.mojo:13-82,
strut_io.mojo:326-347): ``--input <case.json>`` or ``--input-bin
<case.bin>`` (the STRUT_JSON_CACHE_V1 file), ``--write-input-bin
<case.bin>``, ``--output <dir>``, ``--batch <manifest.json>``, ``--profile
<speedscope.json>`` and ``--compute-only``.  Into the output directory it
writes what the reference writes after a transient run
(main.mojo:160-2290, transient_nonlinear.mojo:1553-1695 and 3228-3432):
the recorder files, ``analysis_time_us.txt``, ``phase_times_us.json`` and
``run_progress.json``; a batch run adds ``case_time_us.txt`` per case.

``section_path`` is the fiber-section path driver (section_path.mojo:349-
567): ``--input <json> --output <csv>``, one row per deformation point, the
section's trial state committed after each.

The analysis runs on the GPU through libstrut_gpu (strut_amd.transient
for 2D frames, strut_amd.transient3d for 3D frames); a case outside that
path stops with a message and exit status 2, and a run that stops early
(no convergence, a singular band, an element failure) with the
reference's abort message and exit status 1, without output files, as the
reference's abort leaves none.
"""

from __future__ import annotations

import json
import re
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._native import StrutGpuError
from .json_native import CaseSourceInfo, load_case_document, load_json_native

USAGE = ("usage: strut.mojo -- (--input <case.json> | --input-bin <case.bin>) --output <dir> "
         "[--batch <manifest.json>] [--profile <speedscope.json>] [--compute-only] [--write-input-bin <case.bin>]")


class CaseOutsidePath(ValueError):
    """The case asks for something the device path does not run."""


class RunAborted(RuntimeError):
    """The analysis stopped where the reference aborts."""


# ---------------------------------------------------------------- arguments

def arg_value(argv: Sequence[str], name: str) -> str:
    """strut_io.mojo:36-41: the word after the first occurrence of name."""
    for i in range(len(argv) - 1):
        if argv[i] == name:
            return argv[i + 1]
    return ""


@dataclass
class Args:
    input: str = ""
    input_bin: str = ""
    write_input_bin: str = ""
    output: str = ""
    batch: str = ""
    profile: str = ""
    compute_only: bool = False


def parse_args(argv: Sequence[str]) -> Args:
    """strut_io.mojo:326-347 (argv[0] is the program, as in the reference)."""
    return Args(arg_value(argv, "--input"), arg_value(argv, "--input-bin"), arg_value(argv, "--write-input-bin"),
                arg_value(argv, "--output"), arg_value(argv, "--batch"), arg_value(argv, "--profile"),
                "--compute-only" in argv)


# ---------------------------------------------------------------- time series files

_NUMBER_IN_TEXT = re.compile(r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eEdD][+-]?[0-9]+)?")


def parse_numeric_values_text(text: str, values_path: str) -> List[float]:
    """input_types.mojo:1603-1676: every number in the text, in order;
    Fortran D exponents read as E; anything else separates."""
    values = [float(m.group(0).replace("D", "E").replace("d", "e")) for m in _NUMBER_IN_TEXT.finditer(text)]
    if not values:
        raise ValueError("Path time_series values_path had no numeric values: " + values_path)
    return values


def resolve_time_series(case: dict, source: CaseSourceInfo) -> dict:
    """Path series given by a file (values_path / path, relative to the
    case's directory when that file exists there; input_types.mojo:1679-
    1699, 1711-1734) become inline values; PathFile reads as Path."""
    for ts in case.get("time_series", []) or []:
        if ts.get("type") == "PathFile":
            ts["type"] = "Path"
        if ts.get("type") != "Path":
            continue
        has_file = "values_path" in ts or "path" in ts
        if "values" in ts and has_file:
            raise ValueError("Path time_series cannot specify both values and values_path/path")
        if "values" in ts:
            continue
        vp = ts.get("values_path", "") or ts.get("path", "")
        if not vp:
            raise ValueError("Path time_series missing values_path")
        resolved = vp
        if not vp.startswith("/") and source.case_dir:
            cand = str(Path(source.case_dir) / vp)
            if Path(cand).exists():
                resolved = cand
        if not Path(resolved).exists():
            raise ValueError("Path time_series values_path not found: " + vp)
        ts["values"] = parse_numeric_values_text(Path(resolved).read_text(), resolved)
        ts.pop("values_path", None)
        ts.pop("path", None)
    return case


# ---------------------------------------------------------------- recorders

def format_values_line(values) -> str:
    """helpers.mojo:2581-2588: space-separated shortest round-trip decimals."""
    return " ".join(map(repr, [float(v) for v in values])) + "\n"


@dataclass
class RecorderPlan:
    """The device recorder columns (full-DOF indices recorded after every
    converged step) and what each output file is made of."""
    columns: List[int] = field(default_factory=list)
    files: List[str] = field(default_factory=list)
    # per-step emitters in recorder order: ("row", file, [column positions]) or
    # ("drift", file, column i, column j, dx)
    emitters: List[tuple] = field(default_factory=list)
    envelopes: Dict[str, List[List[int]]] = field(default_factory=dict)  # file -> column-position lists, in order

    def column(self, dof: int) -> int:
        if dof not in self.columns:
            self.columns.append(dof)
        return self.columns.index(dof)

    def file(self, name: str) -> int:
        if name not in self.files:
            self.files.append(name)
        return self.files.index(name)


def recorder_plan(case: dict, ndf: int, has_z: bool) -> RecorderPlan:
    """The node displacement, envelope displacement and drift recorders of
    a case (input_types.mojo:2917-3019; file names and DOF indices as
    transient_nonlinear.mojo:1553-1695).  Other recorder types are outside
    the device path."""
    nodes = case["nodes"]
    idx = {int(n["id"]): i for i, n in enumerate(nodes)}
    plan = RecorderPlan()

    def node_index(nid, what):
        if int(nid) not in idx:
            raise ValueError(what)
        return idx[int(nid)]

    for rec in case.get("recorders", []) or []:
        rtype = rec.get("type", "")
        if rtype in ("node_displacement", "envelope_node_displacement"):
            out = rec.get("output", "node_disp" if rtype == "node_displacement" else "envelope_node_displacement")
            if rec.get("nodes") is None or rec.get("dofs") is None:
                raise ValueError("node recorder requires nodes and dofs")
            dofs = [int(d) for d in rec["dofs"]]
            for d in dofs:
                if d < 1 or d > ndf:
                    raise ValueError(f"recorder dof out of range: {d}")
            for nid in rec["nodes"]:
                i = node_index(nid, "recorder node not found")
                cols = [plan.column(ndf * i + d - 1) for d in dofs]
                name = f"{out}_node{int(nid)}.out"
                if rtype == "node_displacement":
                    plan.emitters.append(("row", plan.file(name), cols))
                else:
                    plan.file(name)
                    plan.envelopes.setdefault(name, []).append(cols)
        elif rtype == "drift":
            out = rec.get("output", "drift")
            if any(k not in rec for k in ("i_node", "j_node", "dof", "perp_dirn")):
                raise ValueError("drift recorder requires i_node, j_node, dof, perp_dirn")
            i_node, j_node = int(rec["i_node"]), int(rec["j_node"])
            i = node_index(i_node, "drift i_node not found")
            j = node_index(j_node, "drift j_node not found")
            d, perp = int(rec["dof"]), int(rec["perp_dirn"])
            if d < 1 or d > ndf:
                raise ValueError(f"drift recorder dof out of range: {d}")
            if perp < 1 or perp > 3:
                raise ValueError("drift perp_dirn must be in 1..3")
            if perp == 3 and not has_z:
                raise ValueError("drift perp_dirn=3 requires z coordinates")
            key = ("x", "y", "z")[perp - 1]
            dx = float(nodes[j].get(key, 0.0)) - float(nodes[i].get(key, 0.0))
            if dx == 0.0:
                raise ValueError("drift denominator is zero")
            name = f"{out}_i{i_node}_j{j_node}.out"
            plan.emitters.append(("drift", plan.file(name), plan.column(ndf * i + d - 1),
                                  plan.column(ndf * j + d - 1), dx))
        else:
            raise CaseOutsidePath(f"recorder type {rtype!r} is outside the device path")
    return plan


def recorder_outputs(plan: RecorderPlan, hist: np.ndarray) -> Dict[str, str]:
    """File name -> contents from the recorded history (steps x columns):
    a line per step and recorder for row and drift files, in recorder
    order; three lines (min, max, largest magnitude) for envelope files
    (helpers.mojo:3153-3225, the first step's value kept on ties)."""
    steps = hist.shape[0]
    per_file: Dict[int, List[List[str]]] = {}
    for em in plan.emitters:
        if em[0] == "row":
            lines = [format_values_line(r) for r in hist[:, em[2]].tolist()] if steps else []
        else:
            _, _, ci, cj, dx = em
            lines = [repr(float(v)) + "\n" for v in ((hist[:, cj] - hist[:, ci]) / dx).tolist()]
        per_file.setdefault(em[1], []).append(lines)
    out: Dict[str, str] = {name: "" for name in plan.files}  # every named file is written, even empty
    for f, streams in per_file.items():
        if len(streams) == 1:
            out[plan.files[f]] = "".join(streams[0])
        else:
            out[plan.files[f]] = "".join("".join(s[k] for s in streams) for k in range(steps))
    for name, col_lists in plan.envelopes.items():
        if steps == 0:
            continue
        mins = maxs = abss = None
        for cols in col_lists:  # a second recorder on the same file updates the first's envelope
            v = hist[:, cols]
            if mins is None:
                mins = v[np.argmin(v, axis=0), np.arange(v.shape[1])]
                maxs = v[np.argmax(v, axis=0), np.arange(v.shape[1])]
                abss = np.max(np.abs(v), axis=0)
            else:
                k = min(len(cols), len(mins))
                mins[:k] = np.minimum(mins[:k], v[:, :k].min(axis=0))
                maxs[:k] = np.maximum(maxs[:k], v[:, :k].max(axis=0))
                abss[:k] = np.maximum(abss[:k], np.abs(v[:, :k]).max(axis=0))
        out[name] += format_values_line(mins) + format_values_line(maxs) + format_values_line(abss)
    return out


# ---------------------------------------------------------------- running a case

PHASE_KEYS = ("case_load_parse_us", "model_build_dof_map_us", "analysis_us", "element_state_update_us",
              "predictor_section_eval_us", "corrector_section_eval_us", "local_flexibility_accumulation_us",
              "local_3x3_solve_us", "local_commit_revert_us", "solve_total_us", "global_nonlinear_iterations",
              "local_force_beam_column_iterations", "subdivision_fallback_iterations", "tangent_factorizations",
              "section_evaluations", "active_bandwidth", "active_nnz", "active_profile_size", "output_write_us",
              "total_case_us")


def phase_times_json(values: dict, type_timing: dict, type_calls: dict, step_iterations) -> str:
    """main.mojo:2175-2283's layout (keys in its order, integers), then
    the per-step Newton iteration counts of the run."""
    lines = ["{"]
    for k in PHASE_KEYS:
        lines.append(f'  "{k}": {int(values.get(k, 0))},')
    for name, table in (("element_type_timing_us", type_timing), ("element_type_call_counts", type_calls)):
        items = [f'    "{t}": {int(v)}' for t, v in table.items() if int(v) > 0]
        lines.append(f'  "{name}": {{')
        lines.append(",\n".join(items))
        lines.append("  },")
    lines.append('  "step_newton_iterations": [' + ", ".join(str(int(i)) for i in step_iterations) + "]")
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_run_progress(path: Path, status: str, analysis_type: str, step: int = 0, step_count: int = 0) -> None:
    """helpers.mojo:120-143."""
    path.write_text("{\n" f'  "status": "{status}",\n' f'  "analysis_type": "{analysis_type}",\n'
                    '  "stage_number": 0,\n  "stage_count": 0,\n' f'  "step_number": {step},\n'
                    f'  "step_count": {step_count}\n' "}\n")


def write_speedscope(path: str, spans: List[Tuple[str, int, int]], total_us: int) -> None:
    """profile.mojo:124-163: an evented speedscope file of the case's phases."""
    names = []
    events = []
    for name, a, b in spans:
        if name not in names:
            names.append(name)
        f = names.index(name)
        events.append({"type": "O", "frame": f, "at": int(a)})
        events.append({"type": "C", "frame": f, "at": int(b)})
    doc = {"$schema": "https://www.speedscope.app/file-format-schema.json",
           "shared": {"frames": [{"name": n} for n in names]},
           "profiles": [{"type": "evented", "name": "strut", "unit": "microseconds", "startValue": 0,
                         "endValue": int(total_us), "events": events}]}
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(doc, separators=(",", ":")) + "\n")


def _us(ns: int) -> int:
    return ns // 1000


@dataclass
class DeviceRun:
    hist: np.ndarray
    iters: np.ndarray
    status: int
    steps_done: int
    hbw: int
    element_types: Dict[str, int]


def _status_message(status: int, dim: int) -> str:
    if status > 0:
        return f"transient_nonlinear did not converge at step {status}"  # transient_nonlinear.mojo:2850
    if status == -11:
        return "BandGeneral matrix is singular or near-singular"  # linear_solver_backend.mojo:661-662
    if status in (-1, -12):
        return f"forceBeamColumn{dim}d element state failed"
    return f"device run stopped with status {status}"


def _run_frame2d(case: dict, plan: RecorderPlan, device: int, marks: dict) -> DeviceRun:
    from .frame2d import load_frame2d_case
    from .transient import FrameBatch

    try:
        fc = load_frame2d_case(case)
    except (ValueError, KeyError) as e:
        raise CaseOutsidePath(str(e)) from None
    fc.record_dofs = np.array(plan.columns, dtype=np.int32)
    fb = FrameBatch(fc.materials, device=device)
    try:
        f = fb.add_frame(fc)
        inst = fb.add_case_motion(f, fc)
        fb.finalize()
        hbw = fb.info(f)["hbw"]
        marks["model_built"] = time.perf_counter_ns()
        fb.run(rounds=64)
        res = fb.result(inst)
        marks["solved"] = time.perf_counter_ns()
    finally:
        fb.close()
    names = {0: "forceBeamColumn2d", 1: "dispBeamColumn2d"}
    counts: Dict[str, int] = {}
    for t in (fc.elem_type if len(fc.elem_type) else np.zeros(len(fc.elem_nip), dtype=np.int32)):
        counts[names[int(t)]] = counts.get(names[int(t)], 0) + 1
    return DeviceRun(res["u"], res["iters"], res["status"], res["steps_done"], hbw, counts)


def _run_frame3d(case: dict, plan: RecorderPlan, device: int, marks: dict) -> DeviceRun:
    from .frame3d import load_frame3d_case
    from .transient3d import Frame3dRun

    try:
        fc = load_frame3d_case(case)
    except (ValueError, KeyError) as e:
        raise CaseOutsidePath(str(e)) from None
    run = Frame3dRun(fc, device=device)
    try:
        if plan.columns:
            run.set_recorder(plan.columns, fc.steps)
        marks["model_built"] = time.perf_counter_ns()
        res = run.run(fc.steps, record_history=False)
        hist = run.recorded() if plan.columns else np.zeros((res["steps_done"], 0))
        marks["solved"] = time.perf_counter_ns()
    finally:
        run.close()
    names = {0: "forceBeamColumn3d", 1: "dispBeamColumn3d"}
    counts: Dict[str, int] = {}
    for t in fc.elem_kind:
        counts[names[int(t)]] = counts.get(names[int(t)], 0) + 1
    return DeviceRun(hist, res["iters"][:res["steps_done"]], res["status"], res["steps_done"], run.hbw, counts)


def run_case(doc, source: CaseSourceInfo, output_path: str, profile_path: str, case_load_us: int,
             include_recorders: bool, device: int = 0) -> dict:
    """main.mojo:160-2290 for a transient_nonlinear fiber-frame case, run
    on the device.  Returns the phase record written to
    phase_times_us.json."""
    t_start = time.perf_counter_ns()
    t0 = t_start - max(case_load_us, 0) * 1000
    out_dir = Path(output_path)
    out_dir.mkdir(parents=True, exist_ok=True)
    progress = out_dir / "run_progress.json"
    write_run_progress(progress, "loading", "case")
    case = doc.to_python() if not isinstance(doc, dict) else doc
    if not isinstance(case, dict):
        raise CaseOutsidePath("the case document must be a JSON object")
    resolve_time_series(case, source)
    if not include_recorders:
        case.pop("recorders", None)
    analysis = case.get("analysis", {}) or {}
    atype = analysis.get("type", "")
    if atype != "transient_nonlinear":
        raise CaseOutsidePath(f"analysis type {atype!r} is outside the device path (transient_nonlinear only)")
    model = case.get("model", {}) or {}
    ndm, ndf = model.get("ndm"), model.get("ndf")
    if (ndm, ndf) not in ((2, 3), (3, 6)):
        raise CaseOutsidePath(f"ndm {ndm} / ndf {ndf} is outside the device path (2/3 and 3/6 frames)")
    plan = recorder_plan(case, ndf, has_z=(ndm == 3))
    steps = int(analysis.get("steps", 0))
    write_run_progress(progress, "running", atype, 0, steps)
    marks: dict = {}
    dev = (_run_frame2d if ndm == 2 else _run_frame3d)(case, plan, device, marks)
    t_model, t_solve_end = marks["model_built"], marks["solved"]
    if dev.status != 0:
        raise RunAborted(_status_message(dev.status, ndm))
    analysis_us = _us(t_solve_end - t_model)
    write_run_progress(progress, "writing_output", atype)
    (out_dir / "analysis_time_us.txt").write_text(f"{analysis_us}\n")
    for name, text in recorder_outputs(plan, np.asarray(dev.hist)).items():
        (out_dir / name).write_text(text)
    t2 = time.perf_counter_ns()
    iters = np.asarray(dev.iters, dtype=np.int64)
    values = {
        "case_load_parse_us": max(case_load_us, 0),
        "model_build_dof_map_us": _us(t_model - t_start),
        "analysis_us": analysis_us,
        "solve_total_us": analysis_us,
        "global_nonlinear_iterations": int(iters.sum()),
        # NormUnbalance is tested before the factorization, so the converged iteration factors nothing
        "tangent_factorizations": int(np.maximum(iters - 1, 0).sum()),
        "active_bandwidth": 2 * dev.hbw + 1,
        "output_write_us": _us(t2 - t_solve_end),
        "total_case_us": _us(t2 - t0),
    }
    calls = {t: n * int(iters.sum()) for t, n in dev.element_types.items()}
    (out_dir / "phase_times_us.json").write_text(phase_times_json(values, {}, calls, iters))
    write_run_progress(progress, "completed", atype)
    if profile_path:
        spans = [("total", 0, values["total_case_us"]), ("case_load_parse", 0, values["case_load_parse_us"]),
                 ("model_build_dof_map", _us(t_start - t0), _us(t_model - t0)),
                 ("solve", _us(t_model - t0), _us(t_solve_end - t0)),
                 ("output", _us(t_solve_end - t0), values["total_case_us"])]
        write_speedscope(profile_path, spans, values["total_case_us"])
    return values


def _load_and_run(input_path, input_bin, write_bin, output, profile, compute_only, device) -> None:
    t = time.perf_counter_ns()
    doc, source = load_case_document(input_path, input_bin, write_bin)
    load_us = _us(time.perf_counter_ns() - t)
    run_case(doc, source, output, profile, load_us, not compute_only, device)


def main(argv: Optional[Sequence[str]] = None, device: int = 0) -> int:
    """strut.mojo:13-82.  Returns the process exit status."""
    argv = list(sys.argv if argv is None else argv)
    a = parse_args(argv)
    try:
        if a.batch:
            manifest = json.loads(Path(a.batch).read_text())
            for entry in manifest.get("cases", []):
                e_in, e_bin = str(entry.get("input", "")), str(entry.get("input_bin", ""))
                e_out, e_prof = str(entry.get("output", "")), str(entry.get("profile", ""))
                if (not e_in and not e_bin) or not e_out:
                    raise ValueError("batch entry missing input/output")
                t = time.perf_counter_ns()
                _load_and_run(e_in, e_bin, "", e_out, e_prof, a.compute_only, device)
                elapsed = _us(time.perf_counter_ns() - t)
                Path(e_out).mkdir(parents=True, exist_ok=True)
                (Path(e_out) / "case_time_us.txt").write_text(f"{elapsed}\n")
            return 0
        if not a.output:
            if a.write_input_bin:
                load_case_document(a.input, a.input_bin, a.write_input_bin)
                return 0
            print(USAGE, file=sys.stderr)
            return 2
        _load_and_run(a.input, a.input_bin, a.write_input_bin, a.output, a.profile, a.compute_only, device)
        return 0
    except CaseOutsidePath as e:
        print(f"strut: {e}", file=sys.stderr)
        return 2
    except RunAborted as e:
        print(f"ABORT: {e}", file=sys.stderr)
        return 1
    except (ValueError, StrutGpuError) as e:
        print(f"ABORT: {e}", file=sys.stderr)
        return 1


# ---------------------------------------------------------------- section path

def run_section_path(input_path: str, output_path: str, device: int = 0) -> str:
    """section_path.mojo:349-567 on the device: one section instance, each
    deformation point evaluated as a trial state and committed."""
    from .model import MaterialTable, discretize_section
    from .sections import SectionRuntime

    doc = load_json_native(input_path).to_python()
    for key in ("materials", "section", "deformation_path"):
        if key not in doc:
            raise ValueError("missing required key: " + key)
    for m in doc["materials"]:
        if int(m.get("id", -1)) < 0:
            raise ValueError("section_path materials require id")
    mats = MaterialTable(doc["materials"])
    section = doc["section"]
    stype = section.get("type", "")
    if "params" not in section:
        raise ValueError("missing required key: params")
    if stype not in ("FiberSection2d", "FiberSection3d"):
        raise ValueError("section_path requires FiberSection2d or FiberSection3d")
    dim = 2 if stype == "FiberSection2d" else 3
    cells = discretize_section(section, mats)
    path = doc["deformation_path"]
    if not isinstance(path, list):
        raise ValueError("deformation_path must be an array")

    def num(step, key, default):
        v = step.get(key, default) if isinstance(step, dict) else default
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise ValueError(key + " must be a number")
        return float(v)

    rt = SectionRuntime(mats, device=device)
    try:
        d = rt.add_definition(dim, cells)
        rt.alloc_instances(d, 1)
        rt.finalize()
        rows = []
        if dim == 2:
            rows.append("eps0,kappa,N,Mz,k11,k12,k22\n")
            for step in path:
                eps0, kappa = num(step, "eps0", 0.0), num(step, "kappa", 0.0)
                r = rt.set_trial(d, 0, eps0, kappa)
                rows.append(",".join(map(repr, [eps0, kappa] + [float(v) for v in r])) + "\n")
                rt.commit_all()
        else:
            rows.append("eps0,ky,kz,N,My,Mz,k11,k12,k13,k22,k23,k33\n")
            for step in path:
                eps0 = num(step, "eps0", 0.0)
                ky = num(step, "kappa_y", num(step, "ky", 0.0))
                kz = num(step, "kappa_z", num(step, "kz", num(step, "kappa", 0.0)))
                r = rt.set_trial(d, 0, eps0, kz, ky)
                rows.append(",".join(map(repr, [eps0, ky, kz] + [float(v) for v in r])) + "\n")
                rt.commit_all()
    finally:
        rt.close()
    text = "".join(rows)
    Path(output_path).write_text(text)
    return text


def section_path_main(argv: Optional[Sequence[str]] = None, device: int = 0) -> int:
    argv = list(sys.argv if argv is None else argv)
    inp, out = arg_value(argv, "--input"), arg_value(argv, "--output")
    if not inp or not out:
        print("usage: section_path.strut --input <json> --output <csv>", file=sys.stderr)
        return 2
    try:
        run_section_path(inp, out, device)
    except (ValueError, StrutGpuError) as e:
        print(f"ABORT: {e}", file=sys.stderr)
        return 1
    return 0

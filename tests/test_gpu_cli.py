"""GPU runs of the ``strut`` / ``section_path`` process boundary
(strut_amd/bin, strut_amd/cli.py) against the oracle on the same cases.

The recorder files are read back as numbers and compared with the oracle's
displacement history at the recorded DOFs (drift and envelope files with
the oracle's drift / envelope restatement applied to that history), with
the transient tests' bar: identical per-step Newton counts
(phase_times_us.json) and values within 1e-8 of the largest magnitude of
each column.  Runs from the JSON file and from its STRUT_JSON_CACHE_V1
file give the same bytes in every file.  The section path's CSV rows are
compared with the oracle's section restatement, point by point with commits
in between, within 1e-12 relative (exact for the strains written back).
"""

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases, cli
from strut_amd.frame2d import load_frame2d_case
from strut_amd.frame3d import NDF, load_frame3d_case
from strut_amd.model import MaterialTable, discretize_section

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parents[1]
REL = 1e-8


def _rows(path: Path) -> np.ndarray:
    return np.array([[float(v) for v in line.split()] for line in path.read_text().splitlines()])


def _close(got, want, rel=REL):
    scale = np.maximum(np.abs(want).max(axis=0), 1e-12)
    err = float(np.max(np.abs(got - want) / scale))
    assert err <= rel, err
    return err


@pytest.fixture(scope="module")
def config0(tmp_path_factory):
    d = tmp_path_factory.mktemp("config0")
    case = cases.config0_case(steps=300, seed=0)
    (d / "case.json").write_text(json.dumps(case))
    fc = load_frame2d_case(case)
    ref = oracle.run_transient2d(fc)
    assert ref["status"] == 0
    return d, case, fc, ref


def test_strut_binary_writes_the_reference_outputs(config0):
    d, case, fc, ref = config0
    out = d / "out_json"
    print("strut binary on config 0 (300 steps)", flush=True)
    proc = subprocess.run([str(REPO / "strut_amd" / "bin" / "strut"), "--input", str(d / "case.json"),
                           "--write-input-bin", str(d / "case.bin"), "--output", str(out)],
                          capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr
    rec = case["recorders"][0]
    nid = rec["nodes"][0]
    idx = [n["id"] for n in case["nodes"]].index(nid)
    got = _rows(out / f"{rec['output']}_node{nid}.out")
    want = ref["u"][:, [3 * idx + int(k) - 1 for k in rec["dofs"]]]
    assert got.shape == want.shape == (300, 3)
    err = _close(got, want)
    phases = json.loads((out / "phase_times_us.json").read_text())
    assert phases["step_newton_iterations"] == ref["iters"].tolist()
    assert phases["global_nonlinear_iterations"] == int(ref["iters"].sum())
    assert phases["active_bandwidth"] == 2 * ref["hbw"] + 1
    assert list(phases)[:len(cli.PHASE_KEYS)] == list(cli.PHASE_KEYS)
    assert int((out / "analysis_time_us.txt").read_text()) > 0
    assert json.loads((out / "run_progress.json").read_text())["status"] == "completed"
    print(f"roof rows: max relative error {err:.2e}; Newton iterations {phases['global_nonlinear_iterations']}")


def test_cache_input_batch_and_compute_only(config0, tmp_path):
    d, case, fc, ref = config0
    if not (d / "case.bin").exists():
        pytest.skip("the cache is written by the binary test")
    manifest = {"cases": [{"input": str(d / "case.json"), "output": str(tmp_path / "a")},
                          {"input_bin": str(d / "case.bin"), "output": str(tmp_path / "b")}]}
    (tmp_path / "batch.json").write_text(json.dumps(manifest))
    assert cli.main(["strut", "--batch", str(tmp_path / "batch.json")]) == 0
    files = sorted(p.name for p in (tmp_path / "a").iterdir())
    assert files == sorted(p.name for p in (tmp_path / "b").iterdir())
    assert "case_time_us.txt" in files
    roof = [f for f in files if f.endswith(".out")]
    assert len(roof) == 1
    assert (tmp_path / "a" / roof[0]).read_bytes() == (tmp_path / "b" / roof[0]).read_bytes() == \
        (d / "out_json" / roof[0]).read_bytes()
    assert cli.main(["strut", "--input-bin", str(d / "case.bin"), "--output", str(tmp_path / "c"),
                     "--compute-only", "--profile", str(tmp_path / "p" / "prof.json")]) == 0
    assert not list((tmp_path / "c").glob("*.out"))
    phases = json.loads((tmp_path / "c" / "phase_times_us.json").read_text())
    assert phases["step_newton_iterations"] == ref["iters"].tolist()
    prof = json.loads((tmp_path / "p" / "prof.json").read_text())
    assert prof["profiles"][0]["type"] == "evented"


def test_strut_3d_drift_and_envelope_files(tmp_path):
    case = cases.rc_frame3d_case(2, 2, 1, steps=80, dt=0.005, n_core=4, n_bars=4, pgas_g=(0.5, 0.35))
    (tmp_path / "case.json").write_text(json.dumps(case))
    print("strut on a 2-story 2 x 1 bay 3D frame (80 steps)", flush=True)
    assert cli.main(["strut", "--input", str(tmp_path / "case.json"), "--output", str(tmp_path / "o")]) == 0
    fc = load_frame3d_case(case)
    ref = oracle.run_transient3d(fc)
    assert ref["status"] == 0
    u = ref["u"]
    idx = {n["id"]: i for i, n in enumerate(case["nodes"])}
    coords = {n["id"]: (n["x"], n["y"], n.get("z", 0.0)) for n in case["nodes"]}
    phases = json.loads((tmp_path / "o" / "phase_times_us.json").read_text())
    assert phases["step_newton_iterations"] == ref["iters"].tolist()
    n_drift = n_env = 0
    for rec in case["recorders"]:
        if rec["type"] == "drift":
            i, j, dof = rec["i_node"], rec["j_node"], rec["dof"]
            cols = [NDF * idx[i] + dof - 1, NDF * idx[j] + dof - 1]
            dx = coords[j][rec["perp_dirn"] - 1] - coords[i][rec["perp_dirn"] - 1]
            want = ((u[:, cols[1]] - u[:, cols[0]]) / dx)[:, None]
            _close(_rows(tmp_path / "o" / f"{rec['output']}_i{i}_j{j}.out"), want)
            n_drift += 1
        elif rec["type"] == "envelope_node_displacement":
            for nid in rec["nodes"]:
                cols = [NDF * idx[nid] + k - 1 for k in rec["dofs"]]
                env, _ = oracle.envelopes(u[:, cols], u.shape[0])
                got = _rows(tmp_path / "o" / f"{rec['output']}_node{nid}.out")
                assert got.shape == (3, len(cols))
                _close(got.T, env)
                n_env += 1
        elif rec["type"] == "node_displacement":
            for nid in rec["nodes"]:
                cols = [NDF * idx[nid] + k - 1 for k in rec["dofs"]]
                _close(_rows(tmp_path / "o" / f"{rec['output']}_node{nid}.out"), u[:, cols])
    assert n_drift == 4 and n_env == 2


def test_strut_stops_where_the_reference_aborts(tmp_path, capsys):
    """A record scaled far past the frame's capacity: the reference aborts
    at the step Newton does not converge; the CLI exits 1 with that message
    and leaves no analysis time or recorder file."""
    case = cases.config0_case(steps=300, seed=0)
    case["time_series"][0]["factor"] = 40.0
    case["analysis"]["max_iters"] = 6
    fc = load_frame2d_case(case)
    ref = oracle.run_transient2d(fc, record_history=False)
    assert ref["status"] > 0
    (tmp_path / "case.json").write_text(json.dumps(case))
    assert cli.main(["strut", "--input", str(tmp_path / "case.json"), "--output", str(tmp_path / "o")]) == 1
    assert f"did not converge at step {ref['status']}" in capsys.readouterr().err
    assert not (tmp_path / "o" / "analysis_time_us.txt").exists() and not list((tmp_path / "o").glob("*.out"))


def _section_doc(dim):
    mats = [{"id": 1, "type": "Concrete02", "params": {"fpc": -30.0e3, "epsc0": -0.002, "fpcu": -6.0e3,
                                                        "epscu": -0.012, "rat": 0.1, "ft": 3.0e3, "Ets": 1.5e6}},
            {"id": 2, "type": "Steel02", "params": {"Fy": 420.0e3, "E0": 200.0e6, "b": 0.01}}]
    sec = {"id": 1, "type": "FiberSection2d" if dim == 2 else "FiberSection3d",
           "params": {"patches": [{"type": "rect", "material": 1, "num_subdiv_y": 10, "num_subdiv_z": 4,
                                   "y_i": -0.25, "z_i": -0.2, "y_j": 0.25, "z_j": 0.2}],
                      "layers": [{"type": "straight", "material": 2, "num_bars": 3, "bar_area": 3.1e-4,
                                  "y_start": -0.2, "z_start": -0.15, "y_end": -0.2, "z_end": 0.15},
                                 {"type": "straight", "material": 2, "num_bars": 3, "bar_area": 3.1e-4,
                                  "y_start": 0.2, "z_start": -0.15, "y_end": 0.2, "z_end": 0.15}]}}
    t = np.linspace(0.0, 6.0 * np.pi, 90)
    path = []
    for k, s in enumerate(t):
        amp = 1e-3 * (0.3 + s / 4.0)
        if dim == 2:
            path.append({"eps0": -2e-4 + 1e-4 * np.sin(0.5 * s), "kappa": 0.02 * amp * np.sin(s) / 1e-3})
        else:
            path.append({"eps0": -2e-4, "kappa_y": 0.01 * amp * np.cos(s) / 1e-3, "kz": 0.02 * amp * np.sin(s) / 1e-3})
    return {"materials": mats, "section": sec, "deformation_path": path}


@pytest.mark.parametrize("dim", [2, 3])
def test_section_path_csv_matches_the_oracle(tmp_path, dim):
    doc = _section_doc(dim)
    (tmp_path / "path.json").write_text(json.dumps(doc))
    print(f"section_path {dim}D, {len(doc['deformation_path'])} points", flush=True)
    proc = subprocess.run([str(REPO / "strut_amd" / "bin" / "section_path"), "--input", str(tmp_path / "path.json"),
                           "--output", str(tmp_path / "out.csv")], capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr
    lines = (tmp_path / "out.csv").read_text().splitlines()
    assert lines[0] == ("eps0,kappa,N,Mz,k11,k12,k22" if dim == 2 else "eps0,ky,kz,N,My,Mz,k11,k12,k13,k22,k23,k33")
    got = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    mats = MaterialTable(doc["materials"])
    t, p = mats.arrays()
    cells = discretize_section(doc["section"], mats)
    ref = (oracle.Section2d if dim == 2 else oracle.Section3d)(cells, t, p)
    inst = ref.alloc(1)
    want = []
    for step in doc["deformation_path"]:
        if dim == 2:
            r = ref.set_trial(inst, step["eps0"], step["kappa"])
            want.append([step["eps0"], step["kappa"]] + list(r))
        else:
            r = ref.set_trial(inst, step["eps0"], step["kappa_y"], step["kz"])
            want.append([step["eps0"], step["kappa_y"], step["kz"]] + list(r))
        ref.commit(inst)
    want = np.array(want)
    ncoord = 2 if dim == 2 else 3
    assert np.array_equal(got[:, :ncoord], want[:, :ncoord])
    err = _close(got[:, ncoord:], want[:, ncoord:], rel=1e-12)
    print(f"section_path {dim}D: max relative error {err:.2e}")

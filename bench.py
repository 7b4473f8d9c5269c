"""Benchmark: fiber updates/s of the fiber-section state determination path
on one MI355X.

Workload: the section population of the 3D RC frame configuration (20
stories x 5 x 5 bays; 720 forceBeamColumn columns and 1200 dispBeamColumn
beams, 5 integration points each, 12 x 12 confined core + 2-fiber cover +
12 Steel02 bars per section).  One step = every section evaluated once at a
seeded synthetic deformation (trial fiber states, section resultants,
tangents and flexibilities), then every fiber committed.  Inputs are
resident in HBM before the timed region.

Prints one JSON line (see the keys below).  `--gpus N` with N > 1 exits:
this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
BYTES_PER_UPDATE = {4: 80, 3: 168}  # Concrete02, Steel02: committed read + trial write


def build_population():
    from strut_amd import cases
    from strut_amd.model import MaterialTable, discretize_section

    pop = cases.frame3d_section_population()
    mats = MaterialTable(cases.rc_materials())
    shapes = [cases.FRAME3D_COLUMN, cases.FRAME3D_BEAM]
    cells = [discretize_section(cases.rc_section(i + 1, dim=3, **s), mats) for i, s in enumerate(shapes)]
    counts = [pop["column_sections"], pop["beam_sections"]]
    return mats, cells, counts, pop


def algorithmic_bytes(mats, cells, counts) -> int:
    total = 0
    for c, n in zip(cells, counts):
        per = sum(BYTES_PER_UPDATE.get(int(mats.defs[m].mat_type), 0) for m in c["mat"])
        total += per * n
    return total


def cpu_baseline(mats, cells, counts, budget_s: float = 12.0) -> dict:
    """The oracle (single-threaded C restatement) on a bounded sample: the
    first 48 section instances of the population (24 column, 24 beam), the
    same deformation sequence, evaluated step after step until the time
    budget is spent."""
    from oracle import oracle
    from strut_amd import cases

    t, p = mats.arrays()
    refs = []
    for c in cells:
        r = oracle.Section3d(c, t, p)
        r.alloc(24)
        refs.append((r, len(c)))
    n = 48
    steps = 0
    updates = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        e, kz, ky = cases.synthetic_section_deformation(n, steps, seed=1234, dim=3)
        for g, (r, nf) in enumerate(refs):
            sl = slice(24 * g, 24 * g + 24)
            r.eval_commit(np.arange(24), e[sl], ky[sl], kz[sl])
            updates += nf * 24
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": updates / dt, "unit": "fiber updates/s", "cores": 1, "kind": "port",
            "sample": f"48 section instances (24 column, 24 beam, {sum(len(c) for c in cells) // 2 * 2} "
                      f"fibers per pair) x {steps} steps of trial evaluation + commit, {dt:.1f} s, "
                      f"oracle C restatement, 1 thread"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(ensembles and models shard across GPUs without collectives; see the design notes)",
              file=sys.stderr)
        return 2

    import torch

    from strut_amd import cases
    from strut_amd._native import POINT_DTYPE, RESULT_DTYPE
    from strut_amd.sections import SectionRuntime

    mats, cells, counts, pop = build_population()
    rt = SectionRuntime(mats, device=0)
    defs = [rt.add_definition(3, c) for c in cells]
    for d, n in zip(defs, counts):
        rt.alloc_instances(d, n)
    rt.finalize()
    fibers, slots = rt.totals()
    n_pts = sum(counts)
    did = np.concatenate([np.full(n, d, dtype=np.int32) for d, n in zip(defs, counts)])
    inst = np.concatenate([np.arange(n, dtype=np.int32) for n in counts])
    total = args.warmup + args.steps
    host = np.zeros((total, n_pts), dtype=POINT_DTYPE)
    for s in range(total):
        e, kz, ky = cases.synthetic_section_deformation(n_pts, s, seed=2024, dim=3)
        host[s] = rt.points(did, inst, e, kz, ky)
    dev = torch.cuda.current_device()
    d_pts = torch.from_numpy(host.view(np.uint8).reshape(total, -1)).to(f"cuda:{dev}")
    d_res = torch.empty(n_pts * RESULT_DTYPE.itemsize, dtype=torch.uint8, device=f"cuda:{dev}")
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(rt.stream)

    for s in range(args.warmup):
        rt.eval_device(n_pts, d_pts[s].data_ptr(), d_res.data_ptr())
        rt.commit_all()
    rt.synchronize()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for k in range(args.steps):
        s = args.warmup + k
        ev[k][0].record(stream)
        rt.eval_device(n_pts, d_pts[s].data_ptr(), d_res.data_ptr())
        ev[k][1].record(stream)
        rt.commit_all()
    t_end.record(stream)
    rt.synchronize()
    torch.cuda.synchronize()
    elapsed_ms = t_start.elapsed_time(t_end)
    eval_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    res = d_res.cpu().numpy().view(RESULT_DTYPE)
    assert np.all(res["valid"] == 1), "a section flexibility was singular"
    ms_per_step = elapsed_ms / args.steps
    value = fibers * args.steps / (elapsed_ms / 1000.0)
    alg_bytes = algorithmic_bytes(mats, cells, counts)
    achieved = alg_bytes / (eval_ms / 1000.0) / 1e9
    out = {
        "metric": "fiber updates/s",
        "value": value,
        "unit": "fiber updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "section population of the 3D RC frame (20 stories x 5x5 bays): "
                        f"{pop['columns']} columns + {pop['beams']} beams x 5 IPs = {n_pts} 3D fiber sections, "
                        f"{fibers} fibers; step = trial evaluation of every section at a seeded cyclic "
                        "deformation + commit",
            "sections": n_pts,
            "fibers": int(fibers),
            "fibers_per_section": [len(c) for c in cells],
            "gm_steps_per_s": None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "section_eval_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_ms": eval_ms,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mats, cells, counts)
    rt.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

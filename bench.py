"""Benchmark of the fiber-section state determination path on one MI355X.

Two legs, one JSON line.

1. Fiber updates/s (``value``).  Workload: the section population of the 3D
   RC frame (20 stories x 5 x 5 bays; 720 forceBeamColumn columns and 1200
   dispBeamColumn beams, 5 integration points each, 12 x 12 confined core +
   2-fiber cover + 12 Steel02 bars per section = 268 fibers; 9600 sections,
   2.57 M fibers).  As Strut's fiber micro-benchmark does
   (benchmarks_micro.mojo:234-263), a preload history of committed steps
   comes first (the warmup steps: evaluate + commit), then every timed step
   evaluates every section once at a new seeded deformation from the
   committed state (trial fiber states, section resultants, tangents and
   flexibilities) without committing.  1 fiber update = one fiber's trial
   evaluation (benchmarks_micro.mojo:250).  Inputs are resident in HBM
   before the timed region.  The commit copy is timed separately.

2. Ground-motion steps/s (``config.gm_steps``).  The 2D RC frame of 3
   stories x 1 bay (forceBeamColumn, 5 Lobatto points, 204-fiber sections),
   1000 steps of dt = 0.01 s under seeded band-limited ground motions at
   PGA 0.4 g: one record alone (the plain path: records solved one by one)
   and N records batched in one launch on the same GPU, plus the oracle's
   single-core rate for one record.

3. The uneven ensemble (``config.ensemble``): the three frame variants
   (4 x 3, 8 x 4 and 12 x 6 bays, sections as above) x N ground motions from
   the ensemble generator (corner frequency 2-15 Hz, strong-motion duration
   5-20 s, PGA on the log grid 0.1-1.5 g, 1000-3000 steps; --ens-steps caps
   them), all in one FrameBatch (autotuned schedule), against the same
   records solved one by one with the plain path (four records of each
   frame timed alone; the one-by-one time is each record's Newton
   iterations at its frame's timed seconds per iteration).  Its GM-steps/s
   as a fraction of the HBM roofline counts the device's own section
   evaluations; the CPU baseline is the oracle in P pinned processes; the
   high-PGA records the device stopped earliest are re-run on the oracle
   and their stops compared.

GM-step roofline fraction of every GM leg: (section-evaluation bytes +
commit bytes + band bytes) / wall time / 8 TB/s, with 80 B (Concrete02) or
168 B (Steel02) per fiber evaluation and per fiber commit, and
n (2 hbw + 1) 8 B x 2 per band factorization plus x 1 per solve.

The single-core CPU baselines run one after another in one worker process
pinned to one CPU (as taskset -c would), after the GPU legs.

4. The learned Newton predictor (``config.learned``): on the 4 x 3 frame, a
   network trained on the device from the pairs the solver records on one
   seed range, then held-out records of another seed range solved with the
   plain Newmark predictor and with the learned one: Newton iterations per
   step, GM-steps/s and the peak inter-story drift error against the plain
   FP64 solution.

`--gpus N` with N > 1 exits: the benchmark measures one GPU.
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
# Algorithmic bytes per fiber update: committed history read + trial history write.
BYTES_PER_UPDATE = {4: 80, 3: 168, 1: 80, 2: 96}  # Concrete02, Steel02, Steel01, Concrete01
AUX_TYPES = (3, 2)  # laws evaluated by aux_eval_kernel (Steel02, Concrete01)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "section_eval_kernel_traffic.json")


def build_population():
    from strut_amd import cases
    from strut_amd.model import MaterialTable, discretize_section

    pop = cases.frame3d_section_population()
    mats = MaterialTable(cases.rc_materials())
    shapes = [cases.FRAME3D_COLUMN, cases.FRAME3D_BEAM]
    cells = [discretize_section(cases.rc_section(i + 1, dim=3, **s), mats) for i, s in enumerate(shapes)]
    counts = [pop["column_sections"], pop["beam_sections"]]
    return mats, cells, counts, pop


def algorithmic_bytes(mats, cells, counts, aux: bool) -> int:
    """Per evaluation pass, over the fibers one kernel updates: the aux laws
    (aux=True, aux_eval_kernel) or the others (section_eval_kernel)."""
    total = 0
    for c, n in zip(cells, counts):
        for m in c["mat"]:
            ty = int(mats.defs[m].mat_type)
            if (ty in AUX_TYPES) == aux:
                total += BYTES_PER_UPDATE.get(ty, 0) * n
    return total


def frame_bytes(fc) -> dict:
    """Algorithmic HBM bytes of a 2D frame's GM-step work: per section point
    evaluation (the frame's sections weighted by their integration points),
    per commit of every section after a converged step, and per band
    factorization (read + write) and solve (read)."""
    mats = fc.materials
    per_sec = {}
    for sid in fc.section_ids:
        per_sec[int(sid)] = sum(BYTES_PER_UPDATE.get(int(mats.defs[m].mat_type), 0) for m in fc.section_cells[sid]["mat"])
    nip = [int(n) for n in fc.elem_nip]
    commit = sum(n * per_sec[int(s)] for n, s in zip(nip, fc.elem_section))
    return {"per_section_eval": commit / max(sum(nip), 1), "per_commit": commit}


def gm_roofline(fbytes: dict, n_free: int, hbw: int, sec_evals: int, steps: int, iters: int,
                seconds: float) -> dict:
    """GM-steps/s as a fraction of the HBM roofline: (section-evaluation +
    commit + band bytes) / wall time / peak.  A step's converged iteration
    factors nothing (the residual test comes first), so factorizations =
    iterations - converged steps."""
    band = n_free * (2 * hbw + 1) * 8
    factorizations = max(iters - steps, 0)
    total = sec_evals * fbytes["per_section_eval"] + steps * fbytes["per_commit"] + factorizations * band * 3
    achieved = total / seconds / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": total, "section_evaluations": int(sec_evals),
            "bytes_per_section_eval": fbytes["per_section_eval"], "commit_bytes_per_step": fbytes["per_commit"],
            "band_factorizations": int(factorizations)}


def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "allowed_cpus": len(os.sched_getaffinity(0)), "lscpu_model": model}


_CPU_SLOT = None


def _pin_worker(q) -> None:
    """Pool initializer: this worker runs on one CPU of its own (taskset -c)."""
    global _CPU_SLOT
    _CPU_SLOT = q.get()
    os.sched_setaffinity(0, {_CPU_SLOT})


def cpu_pool(workers: int):
    """Worker processes started fresh (no GPU state), each pinned to its own
    CPU from this process's allowed set."""
    import concurrent.futures
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    cpus = sorted(os.sched_getaffinity(0))[:workers]
    q = ctx.Queue()
    for c in cpus:
        q.put(c)
    return concurrent.futures.ProcessPoolExecutor(max_workers=len(cpus), mp_context=ctx, initializer=_pin_worker,
                                                  initargs=(q,)), len(cpus)


def _cpu_task(name: str, *a):
    """Run one CPU-baseline or oracle-check task in a pinned worker."""
    if name == "sections":
        mats, cells, _, _ = build_population()
        out = cpu_baseline(mats, cells, *a)
    elif name == "gm":
        out = gm_cpu_baseline(*a)
    elif name == "frame3d":
        out = frame3d_cpu_baseline(*a)
    elif name == "ensemble_record":
        out = _oracle_record(*a)
    elif name == "start_sensitivity":
        out = _start_sensitivity(*a)
    else:
        raise ValueError(name)
    out["cpu"] = _CPU_SLOT
    return out


def _oracle_record(k: int, seed: int, max_steps: int, step_cap: int) -> dict:
    """The oracle on one record of the uneven ensemble (its first step_cap
    steps when step_cap > 0)."""
    from oracle import oracle
    from strut_amd import cases

    s, b = cases.CONFIG2_FRAMES[k]
    fc = _motion_case(cases.config2_frame_case(s, b, steps=100), seed, max_steps)
    if step_cap > 0:
        fc.steps = min(fc.steps, step_cap)
    t0 = time.perf_counter()
    r = oracle.run_transient2d(fc, record_history=False)
    dt = time.perf_counter() - t0
    n = int(r["steps_done"])
    return {"frame": k, "seed": seed, "steps": fc.steps, "status": int(r["status"]), "steps_done": n,
            "iters": [int(v) for v in r["iters"][:n]],
            "newton_iters": int(r["iters"][:fc.steps].sum()), "section_evals": int(r["section_evals"]),
            "seconds": dt}


def _start_sensitivity(seed: int, max_steps: int) -> dict:
    """Whether the oracle's own run of a 4 x 3 frame record changes its stop
    or moves its peak drift beyond 1e-6 relative when every step's Newton
    start is scaled by 1 +- 1e-4 or 1 +- 1e-2 on alternate DOFs."""
    from oracle import oracle
    from strut_amd import cases

    fc = _motion_case(cases.config2_frame_case(4, 3, steps=100), seed, max_steps)
    base = oracle.run_transient2d(fc)
    n = base["steps_done"]
    for eps in (1e-4, 1e-2):
        r = oracle.run_transient2d(fc, start_eps=eps)
        if r["status"] != base["status"] or r["steps_done"] != n:
            return {"evidence": f"start x(1+-{eps:g}): stop {r['status']} instead of {base['status']}"}
        if n:
            d0 = cases.peak_story_drift(base["u"][:n], 4, 3)
            e = float(np.max(np.abs(cases.peak_story_drift(r["u"][:n], 4, 3) - d0) / d0))
            if e > 1e-6:
                return {"evidence": f"start x(1+-{eps:g}): peak drift moves {e:.2e}"}
    return {"evidence": ""}


def cpu_baseline(mats, cells, budget_s: float = 12.0) -> dict:
    """The oracle (single-threaded C restatement) on a bounded sample of the
    same workload: 48 section instances (24 column, 24 beam), a 3-step
    committed preload, then trial evaluations at new seeded deformations
    without commit until the time budget is spent."""
    from oracle import oracle
    from strut_amd import cases

    t, p = mats.arrays()
    refs = []
    for c in cells:
        r = oracle.Section3d(c, t, p)
        r.alloc(24)
        refs.append((r, len(c)))
    n = 48
    insts = np.arange(24)
    for s in range(3):
        e, kz, ky = cases.synthetic_section_deformation(n, s, seed=1234, dim=3)
        for g, (r, _) in enumerate(refs):
            sl = slice(24 * g, 24 * g + 24)
            r.eval_many(insts, e[sl], ky[sl], kz[sl], commit=True)
    steps = 0
    updates = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        e, kz, ky = cases.synthetic_section_deformation(n, 3 + steps, seed=1234, dim=3)
        for g, (r, nf) in enumerate(refs):
            sl = slice(24 * g, 24 * g + 24)
            r.eval_many(insts, e[sl], ky[sl], kz[sl], commit=False)
            updates += nf * 24
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": updates / dt, "unit": "fiber updates/s", "cores": 1, "kind": "port",
            "sample": f"48 section instances (24 column + 24 beam, {len(cells[0])} fibers each) x {steps} "
                      f"trial evaluations after a 3-step committed preload, {dt:.1f} s, oracle C restatement, "
                      "1 thread"}


def section_leg(args) -> dict:
    import torch

    from strut_amd import _native, cases
    from strut_amd._native import POINT_DTYPE, RESULT_DTYPE
    from strut_amd.sections import SectionRuntime

    mats, cells, counts, pop = build_population()
    rt = SectionRuntime(mats, device=0)
    defs = [rt.add_definition(3, c) for c in cells]
    for d, n in zip(defs, counts):
        rt.alloc_instances(d, n)
    rt.finalize()
    fibers, _ = rt.totals()
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

    # preload history: committed steps
    for s in range(args.warmup):
        rt.eval_device(n_pts, d_pts[s].data_ptr(), d_res.data_ptr())
        rt.commit_all()
    rt.synchronize()
    torch.cuda.synchronize()

    # timed steps: the launches alone (no per-kernel events on the stream)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for k in range(args.steps):
        rt.eval_device(n_pts, d_pts[args.warmup + k].data_ptr(), d_res.data_ptr())
    t_end.record(stream)
    rt.synchronize()
    torch.cuda.synchronize()
    elapsed_ms = t_start.elapsed_time(t_end)
    res = d_res.cpu().numpy().view(RESULT_DTYPE)
    assert np.all(res["valid"] == 1), "a section flexibility was singular"
    # the same steps again with HIP events around each kernel (library
    # profiling on the context's stream): per-kernel launch durations
    rt.profile_enable(True)
    for k in range(args.steps):
        rt.eval_device(n_pts, d_pts[args.warmup + k].data_ptr(), d_res.data_ptr())
    rt.synchronize()
    main_ms, main_n = rt.profile_read(_native.KERNEL_SECTION_EVAL)
    aux_ms, aux_n = rt.profile_read(_native.KERNEL_AUX_EVAL)

    rt.profile_enable(True)
    for _ in range(args.steps):
        rt.commit_all()
    commit_ms, commit_n = rt.profile_read(_native.KERNEL_COMMIT)
    rt.profile_enable(False)
    rt.close()

    main_avg = main_ms / main_n
    aux_avg = aux_ms / aux_n if aux_n else 0.0
    main_bytes = algorithmic_bytes(mats, cells, counts, aux=False)
    aux_bytes = algorithmic_bytes(mats, cells, counts, aux=True)
    achieved = main_bytes / (main_avg / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            tr = json.load(fh)
        if tr.get("kernel") == "section_eval_kernel" and tr.get("sections") == n_pts:
            traffic = tr["bytes_per_launch"]
    return {
        "value": fibers * args.steps / (elapsed_ms / 1e3),
        "ms_per_step": elapsed_ms / args.steps,
        "n_pts": n_pts, "fibers": int(fibers), "pop": pop, "cells": cells,
        "roofline": {
            "bound": "hbm",
            "kernel": "section_eval_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": main_bytes,
            "avg_launch_ms": main_avg,
        },
        "aux": {"kernel": "aux_eval_kernel", "avg_launch_ms": aux_avg, "algorithmic_bytes_per_launch": aux_bytes,
                "achieved_GBps": aux_bytes / (aux_avg / 1e3) / 1e9 if aux_avg else None},
        "commit": {"kernel": "copy_state_kernel", "avg_launch_ms": commit_ms / commit_n,
                   "bytes_per_launch": main_bytes + aux_bytes,
                   "achieved_GBps": (main_bytes + aux_bytes) / (commit_ms / commit_n / 1e3) / 1e9},
        "mats": mats,
    }


def gm_leg(n_batch: int, steps: int) -> dict:
    """GM-steps/s of the 2D frame transient: one record alone, then n_batch
    records in one launch, same GPU."""
    from strut_amd import _native, cases
    from strut_amd.frame2d import load_frame2d_case
    from strut_amd.transient import FrameBatch

    fc = load_frame2d_case(cases.config0_case(steps=steps, seed=0))
    motions = [cases.ground_motion(s, fc.dt, steps, 0.4) for s in range(n_batch)]
    out = {}
    for label, ms in (("plain", motions[:1]), ("batched", motions)):
        fb = FrameBatch(fc.materials)
        fid = fb.add_frame(fc)
        ids = [fb.add_motion(fid, m, fc.dt, steps) for m in ms]
        fb.finalize()
        fb.rt.profile_enable(True)
        t0 = time.perf_counter()
        fb.run(rounds=64)
        dt = time.perf_counter() - t0
        kms, kn = fb.rt.profile_read(_native.KERNEL_FRAME2D)
        fb.rt.profile_enable(False)
        sched = fb.schedule_info()
        res = [fb.result(i) for i in ids]
        evals = int(fb.section_evals().sum())
        info = fb.info(fid)
        fb.close()
        done = sum(r["steps_done"] for r in res)
        iters = sum(int(r["iters"][:r["steps_done"]].sum()) for r in res)
        all_iters = sum(int(r["iters"].sum()) for r in res)
        out[label] = {"records": len(ms), "gm_steps": int(done), "seconds": dt, "gm_steps_per_s": done / dt,
                      "kernel_ms": kms, "launches": kn, "newton_iters": iters, "schedule": sched,
                      "stopped_unconverged": int(sum(1 for r in res if r["status"] != 0)),
                      "roofline": gm_roofline(frame_bytes(fc), info["free"], info["hbw"], evals, done, all_iters, dt)}
    out["batched_vs_plain"] = out["batched"]["gm_steps_per_s"] / out["plain"]["gm_steps_per_s"]
    return out


def gm_cpu_baseline(steps: int) -> dict:
    from oracle import oracle
    from strut_amd import cases
    from strut_amd.frame2d import load_frame2d_case

    fc = load_frame2d_case(cases.config0_case(steps=steps, seed=0))
    t0 = time.perf_counter()
    r = oracle.run_transient2d(fc, record_history=False)
    dt = time.perf_counter() - t0
    return {"value": r["steps_done"] / dt, "unit": "GM-steps/s", "cores": 1, "kind": "port",
            "sample": f"one record (seed 0), {r['steps_done']} steps, {dt:.2f} s, oracle C restatement, 1 thread"}


def frame3d_leg(steps: int, warmup: int, hbw_rule: int = 0, records: int = 32) -> dict:
    """Steps/s of the 3D RC frame transient (20 stories x 5 x 5 bays, rigid
    diaphragms, two simultaneous records) on the device: warmup steps, then
    `steps` timed steps of the same run; then `records` record pairs of the
    same frame run together (record-steps/s against one record alone)."""
    from strut_amd import _native, cases
    from strut_amd.frame3d import load_frame3d_case
    from strut_amd.transient3d import Frame3dRun

    fc = load_frame3d_case(cases.config1_case(steps=warmup + steps))
    t0 = time.perf_counter()
    run = Frame3dRun(fc, hbw_rule=hbw_rule)
    setup_s = time.perf_counter() - t0
    w = run.run(warmup, record_history=False)
    run.rt.profile_enable(True)
    t0 = time.perf_counter()
    r = run.run(steps, record_history=False)
    dt = time.perf_counter() - t0
    elem_ms, elem_n = run.rt.profile_read(_native.KERNEL_F3_ELEM)
    solve_ms, solve_n = run.rt.profile_read(_native.KERNEL_F3_SOLVE)
    run.rt.profile_enable(False)
    run.close()
    iters = int(r["iters"][:r["steps_done"]].sum())
    multi = None
    if records > 1:  # the same frame under `records` record pairs, run together
        from strut_amd.frame3d import Pattern

        recs = []
        for k in range(1, records):
            c = cases.rc_frame3d_case(20, 5, 5, steps=warmup + steps, dt=0.005, seeds=(2 * k, 2 * k + 1),
                                      pgas_g=(0.5, 0.35))
            recs.append([Pattern(p.direction, p.values, p.dt, p.factor) for p in load_frame3d_case(c).patterns])
        mrun = Frame3dRun(fc, hbw_rule=hbw_rule, records=recs)
        mrun.run_records(warmup)
        t0 = time.perf_counter()
        mr = mrun.run_records(steps)
        mdt = time.perf_counter() - t0
        mrun.close()
        done = int(mr["steps_done"].sum())
        multi = {"records": records, "record_steps": done, "seconds": mdt, "record_steps_per_s": done / mdt,
                 "newton_iters": int(mr["iters"].sum()), "stopped": int((mr["status"] != 0).sum()),
                 "vs_one_record": (done / mdt) / (r["steps_done"] / dt)}
    return {"frame": "3D RC frame 20 stories x 5 x 5 bays", "nodes": fc.n_nodes, "elements": fc.n_elem,
            "records_together": multi,
            "free_dofs": run.free_count, "hbw": run.hbw, "band_rule": ("element free map", "collapsed")[hbw_rule],
            "setup_seconds": setup_s, "warmup_steps": w["steps_done"], "steps": r["steps_done"],
            "status": r["status"], "seconds": dt, "steps_per_s": r["steps_done"] / dt, "newton_iters": iters,
            "ms_per_newton_iter": 1e3 * dt / max(iters, 1), "dense_fallbacks_of_the_reference": r["dense_fallbacks"],
            "element_state_ms_avg": elem_ms / max(elem_n, 1), "band_solve_ms_avg": solve_ms / max(solve_n, 1)}


def tower_leg(steps: int, warmup: int, records: int, modal_rotations: int) -> dict:
    """The large single-GPU case (60 stories x 6 x 6 bays around a 2 x 2-bay
    elastic shell core, 33,480 free DOFs, half bandwidth 560): `records`
    bidirectional seeded records together (PGA 0.3-0.9 g in x, 0.7 of it in
    y), the modal analysis of the frame at rest with the reference's
    Jacobi (`modal_rotations` rotations), then `warmup` and `steps` timed
    steps of every record; element state and band solve per Newton
    iteration."""
    from strut_amd import _native, cases
    from strut_amd.frame3d import Pattern, load_frame3d_case
    from strut_amd.transient3d import Frame3dRun

    pga = lambda k: 0.3 + 0.6 * k / max(records - 1, 1)  # noqa: E731
    n_steps = warmup + steps
    fc = load_frame3d_case(cases.config4_case(steps=n_steps, seeds=(0, 1), pgas_g=(pga(0), 0.7 * pga(0))))
    recs = []
    for k in range(1, records):
        c = cases.config4_case(steps=n_steps, seeds=(2 * k, 2 * k + 1), pgas_g=(pga(k), 0.7 * pga(k)))
        recs.append([Pattern(p.direction, p.values, p.dt, p.factor) for p in load_frame3d_case(c).patterns])
    t0 = time.perf_counter()
    run = Frame3dRun(fc, records=recs)
    setup_s = time.perf_counter() - t0
    modal = None
    if modal_rotations > 0:
        t0 = time.perf_counter()
        eig, _, rot = run.modal(6, max_rotations=modal_rotations)
        modal = {"rotations": rot, "seconds": time.perf_counter() - t0, "eigenvalues": [float(v) for v in eig],
                 "note": "the reference's classical Jacobi stops at 100 rotations, far from converged at this size"}
    run.run_records(warmup)
    run.rt.profile_enable(True)
    t0 = time.perf_counter()
    r = run.run_records(steps)
    dt = time.perf_counter() - t0
    elem_ms, elem_n = run.rt.profile_read(_native.KERNEL_F3_ELEM)
    solve_ms, solve_n = run.rt.profile_read(_native.KERNEL_F3_SOLVE)
    run.rt.profile_enable(False)
    run.close()
    done = int(r["steps_done"].sum())
    iters = int(r["iters"].sum())
    lockstep = int(np.max(r["iters"], axis=0).sum())  # the records iterate together
    n_fib = sum(len(fc.section_cells[int(fc.elem_section[e])]["y"]) * int(fc.elem_nip[e]) for e in range(fc.n_elem))
    return {"frame": "3D RC tower 60 stories x 6 x 6 bays, 2 x 2-bay elastic ShellMITC4 core",
            "nodes": fc.n_nodes, "beam_columns": fc.n_elem, "shells": len(fc.shell_nodes), "fibers": n_fib,
            "free_dofs": run.free_count, "hbw": run.hbw, "records": records, "setup_seconds": setup_s, "modal": modal,
            "warmup_steps": warmup, "steps": steps, "record_steps": done, "seconds": dt,
            "record_steps_per_s": done / dt, "newton_iters": iters, "lockstep_iters": lockstep,
            "ms_per_lockstep_iter": 1e3 * dt / max(lockstep, 1), "status": [int(v) for v in r["status"]],
            "element_state_ms_avg": elem_ms / max(elem_n, 1), "band_solve_ms_avg": solve_ms / max(solve_n, 1),
            "band_solver": "blocked row-swap LU (sixteen-column panels), all records in each launch",
            "cpu_baseline": None,
            "cpu_baseline_note": "not run: the oracle's driver is dense (K of 33,480^2 doubles, 9 GB, and a "
                                 "dense collapse per Newton iteration)"}


def frame3d_cpu_baseline(steps: int) -> dict:
    from oracle import oracle
    from strut_amd import cases
    from strut_amd.frame3d import load_frame3d_case

    fc = load_frame3d_case(cases.config1_case(steps=steps))
    t0 = time.perf_counter()
    r = oracle.run_transient3d(fc, record_history=False)
    dt = time.perf_counter() - t0
    return {"value": r["steps_done"] / dt, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"the first {r['steps_done']} steps ({int(r['iters'][:r['steps_done']].sum())} Newton "
                      f"iterations, {r['dense_fallbacks']} dense-LU fallbacks), {dt:.1f} s, oracle C restatement "
                      "(dense K, collapse, band with the reference's fallback), 1 thread"}


def _motion_case(frame: dict, seed: int, max_steps: int):
    from strut_amd import cases
    from strut_amd.frame2d import load_frame2d_case

    m = cases.config2_motion(seed, max_steps=max_steps)
    c = dict(frame)
    c["time_series"] = [{"type": "Path", "tag": 1, "dt": m["dt"], "factor": 1.0,
                         "values": [float(v) for v in m["values"]]}]
    c["analysis"] = dict(frame["analysis"], steps=m["steps"])
    return load_frame2d_case(c)


def _run_batch(fcs_by_frame, record=False, predictor=None, record_all=False, schedule=None):
    """One FrameBatch over {frame index: [cases]}; returns (batch, ids, seconds)."""
    from strut_amd import _native
    from strut_amd.transient import FrameBatch

    first = next(iter(fcs_by_frame.values()))[0]
    fb = FrameBatch(first.materials)
    ids, fids = [], {}
    for k, fcs in fcs_by_frame.items():
        fids[k] = fb.add_frame(fcs[0], record_all=record_all)
        ids += [(k, fb.add_case_motion(fids[k], fc)) for fc in fcs]
    if record:
        fb.record_samples(True)
    fb.finalize()
    if predictor is not None:
        for k, net in predictor.items():
            fb.set_predictor(fids[k], net)
    fb.set_schedule(_native.SCHEDULE_AUTO if schedule is None else schedule)
    t0 = time.perf_counter()
    fb.run(rounds=16)
    return fb, fids, ids, time.perf_counter() - t0


def ensemble_leg(n_motions: int, max_steps: int, n_single: int = 4, cpu_workers: int = 16,
                 cpu_steps: int = 300, check_records: int = 3, cpu: bool = True, frame_ids=(0, 1, 2),
                 schedule=None) -> dict:
    """The uneven three-frame ensemble batched vs its records one by one on
    the same GPU, its GM-steps/s roofline fraction, the oracle in P pinned
    processes as the CPU throughput baseline, and high-PGA records' stops
    checked against the oracle."""
    from strut_amd import _native, cases

    frames = {k: cases.config2_frame_case(*cases.CONFIG2_FRAMES[k], steps=100) for k in frame_ids}
    seeds = {k: [100000 + 1000 * k + j for j in range(n_motions)] for k in frames}
    by_frame = {k: [_motion_case(f, sd, max_steps) for sd in seeds[k]] for k, f in frames.items()}
    fb, fids, ids, sec = _run_batch(by_frame, schedule=schedule)
    res = [fb.result(i) for _, i in ids]
    evals = fb.section_evals()
    infos = {k: fb.info(fid) for k, fid in fids.items()}
    sched = fb.schedule_info()
    fb.close()
    done = [int(r["steps_done"]) for r in res]
    iters = [int(r["iters"].sum()) for r in res]
    # roofline: bytes of every instance over the batched wall time
    total_bytes = 0.0
    per_frame_bytes = {k: frame_bytes(by_frame[k][0]) for k in by_frame}
    for (k, i), d, it in zip(ids, done, iters):
        rl = gm_roofline(per_frame_bytes[k], infos[k]["free"], infos[k]["hbw"], int(evals[i]), d, it, 1.0)
        total_bytes += rl["algorithmic_bytes"]
    achieved = total_bytes / sec / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": total_bytes,
                "section_evaluations": int(evals.sum())}
    # plain path, records one by one: n_single records of each frame alone
    # (every n_motions / n_single-th motion); each frame's one-by-one time is
    # its batched Newton iterations times the sample's seconds per iteration
    single, single_it, sample = {}, {}, {}
    for k in frames:
        sec1 = steps1 = it1 = 0.0
        picks = list(range(0, n_motions, max(1, n_motions // n_single)))[:n_single]
        sample[k] = [seeds[k][j] for j in picks]
        for j in picks:
            fb1, _, ids1, s1 = _run_batch({k: [by_frame[k][j]]}, schedule=_native.SCHEDULE_ROUNDS)
            r1 = fb1.result(ids1[0][1])
            steps1 += r1["steps_done"]
            it1 += int(r1["iters"].sum())
            sec1 += s1
            fb1.close()
        single[k] = steps1 / sec1
        single_it[k] = sec1 / it1
    by_steps = sum(d / single[k] for (k, _), d in zip(ids, done))
    by_iters = sum(it * single_it[k] for (k, _), it in zip(ids, iters))
    names = [f"{s}x{b}" for s, b in cases.CONFIG2_FRAMES]
    out = {"records": len(ids), "frames": [names[k] for k in frames], "max_steps": max_steps,
           "record_steps": {"min": int(min(r.steps for fl in by_frame.values() for r in fl)),
                            "max": int(max(r.steps for fl in by_frame.values() for r in fl))},
           "gm_steps": int(sum(done)), "newton_iters": int(sum(iters)), "seconds": sec,
           "gm_steps_per_s": sum(done) / sec, "roofline": roofline,
           "stopped_unconverged": int(sum(1 for r in res if r["status"] != 0)),
           "plain_single_record_gm_steps_per_s": {names[k]: single[k] for k in single},
           "plain_single_record_ms_per_newton_iter": {names[k]: 1e3 * single_it[k] for k in single_it},
           "single_records_timed_per_frame": n_single,
           "one_by_one_seconds_est": by_iters, "batched_vs_one_by_one": by_iters / sec,
           "one_by_one_seconds_est_by_steps": by_steps, "batched_vs_one_by_one_by_steps": by_steps / sec,
           "schedule": sched}
    if not cpu:
        return out
    # CPU: the oracle in P pinned processes -- per-frame seconds per Newton
    # iteration from the timed sample records' first cpu_steps steps, and
    # full oracle runs of high-PGA records that the device stopped early
    status, pos = {}, 0
    for k in by_frame:  # ids hold each frame's records in seed order, frame after frame
        for sd in seeds[k]:
            assert ids[pos][0] == k
            status[(k, sd)] = (res[pos], int(evals[ids[pos][1]]))
            pos += 1
    checks = []
    for k in by_frame:  # per frame, the high-PGA records the device stopped earliest (bounded oracle time)
        cand = [(status[(k, sd)][0]["steps_done"], sd) for sd in seeds[k]
                if cases.config2_motion_params(sd)["pga_g"] >= 0.9 and status[(k, sd)][0]["status"] != 0]
        checks += [(k, sd) for _, sd in sorted(cand)[:check_records]]
    pool, P = cpu_pool(cpu_workers)
    t0 = time.perf_counter()
    with pool:
        timing = [pool.submit(_cpu_task, "ensemble_record", k, sd, max_steps, cpu_steps)
                  for k in by_frame for sd in sample[k]]
        checked = [pool.submit(_cpu_task, "ensemble_record", k, sd, max_steps, 0) for k, sd in checks]
        timing = [f.result() for f in timing]
        checked = [f.result() for f in checked]
    wall = time.perf_counter() - t0
    cpu_it = {k: sum(t["seconds"] for t in timing if t["frame"] == k) /
              max(sum(t["newton_iters"] for t in timing if t["frame"] == k), 1) for k in by_frame}
    cpu_serial = sum(it * cpu_it[k] for (k, _), it in zip(ids, iters))
    out["cpu_baseline"] = {
        "value": sum(done) / (cpu_serial / P), "unit": "GM-steps/s", "cores": P, "kind": "port",
        "sample": f"oracle C restatement in {P} processes, one pinned per CPU: the first {cpu_steps} steps of the "
                  f"{n_single} timed records of each frame give each frame's seconds per Newton iteration "
                  f"({', '.join(f'{names[k]} {1e3 * cpu_it[k]:.3g} ms' for k in cpu_it)}); the ensemble's CPU time "
                  f"is its Newton iterations at those rates over {P} processes ({cpu_serial:.0f} core-seconds)",
        "sample_wall_seconds": wall, **host_cpu()}
    out["batched_vs_cpu_processes"] = out["gm_steps_per_s"] / out["cpu_baseline"]["value"]
    rows = []
    for c in checked:
        r, ev = status[(c["frame"], c["seed"])]
        n = c["steps_done"]
        rows.append({"frame": names[c["frame"]], "seed": c["seed"],
                     "pga_g": cases.config2_motion_params(c["seed"])["pga_g"],
                     "status_gpu": int(r["status"]), "status_oracle": c["status"],
                     "steps_done_gpu": int(r["steps_done"]), "steps_done_oracle": n,
                     "same_iterations": [int(v) for v in r["iters"][:n]] == c["iters"],
                     "section_evaluations_gpu": ev, "section_evaluations_oracle": c["section_evals"]})
    out["high_pga_stop_checks"] = rows
    return out


def learned_leg(n_train: int, n_test: int, max_steps: int, epochs: int, sens_records: int = 16,
                cpu_workers: int = 16) -> dict:
    """Learned predictor on the 4 x 3 frame: train on one seed range, compare
    with the plain predictor on held-out records."""
    from strut_amd import _native, cases
    from strut_amd.learned import train_predictor

    frame = cases.config2_frame_case(4, 3, steps=100)
    train = {0: [_motion_case(frame, 200000 + j, max_steps) for j in range(n_train)]}
    fb, fids, _, gen_s = _run_batch(train, record=True)
    x, y = fb.samples(fids[0])
    fb.close()
    t0 = time.perf_counter()
    net, losses = train_predictor(x, y, seed=7, epochs=epochs, batch=512, lr=1e-3)
    train_s = time.perf_counter() - t0
    held = {0: [_motion_case(frame, 300000 + j, max_steps) for j in range(n_test)]}
    out = {"frame": "4x3", "training_pairs": int(x.shape[0]), "training_records": n_train,
           "pair_generation_seconds": gen_s, "epochs": epochs, "training_seconds": train_s,
           "loss_first_last": [float(losses[0]), float(losses[-1])], "held_out_records": n_test}
    del x, y
    runs = {}
    # both runs on the work queue, so that they differ only in the predictor
    roof = {}
    for label, pred in (("plain", None), ("learned", {0: net})):
        fb, hf, ids, sec = _run_batch(held, predictor=pred, record_all=True, schedule=_native.SCHEDULE_QUEUE)
        runs[label] = ([fb.result(i) for _, i in ids], [fb.inst_stats(i) for _, i in ids], sec)
        info, ev = fb.info(hf[0]), int(fb.section_evals().sum())
        fb.close()
        rr = runs[label][0]
        roof[label] = gm_roofline(frame_bytes(held[0][0]), info["free"], info["hbw"], ev,
                                  sum(int(r["steps_done"]) for r in rr), sum(int(r["iters"].sum()) for r in rr), sec)
    net.close()
    out["roofline"] = roof
    (rp, _, sp), (rl, stl, sl) = runs["plain"], runs["learned"]
    same = [a["status"] == b["status"] and a["steps_done"] == b["steps_done"] for a, b in zip(rp, rl)]
    it_p = sum(int(a["iters"][:a["steps_done"]].sum()) for a in rp)
    it_l = sum(int(b["iters"][:b["steps_done"]].sum()) for b in rl)
    steps_p = sum(a["steps_done"] for a in rp)
    steps_l = sum(b["steps_done"] for b in rl)
    # every record: within the bar (same stop, peak drift within 1e-6), or
    # checked for the reference's own sensitivity to where Newton starts
    errs, flagged = [], []
    for j, (a, b, ok) in enumerate(zip(rp, rl, same)):
        n = min(a["steps_done"], b["steps_done"])
        e = 0.0
        if n:
            dp = cases.peak_story_drift(a["u"][:n], 4, 3)
            dl = cases.peak_story_drift(b["u"][:n], 4, 3)
            e = float(np.max(np.abs(dl - dp) / dp))
        errs.append(e)
        if not ok or e > 1e-6:
            flagged.append({"seed": 300000 + j, "stop_plain": int(a["status"]), "stop_learned": int(b["status"]),
                            "max_iters": int(a["iters"][:max(n, 1)].max()), "drift_rel_err": e})
    checked = flagged[:sens_records]
    if checked and sens_records > 0:
        pool, _ = cpu_pool(cpu_workers)
        with pool:
            futs = [pool.submit(_cpu_task, "start_sensitivity", f["seed"], max_steps) for f in checked]
            for f, fu in zip(checked, futs):
                f["oracle_start_sensitivity"] = fu.result()["evidence"]
    within = sum(1 for e, ok in zip(errs, same) if ok and e <= 1e-6)
    out.update({
        "schedule": "queue",
        "newton_iters_per_step_plain": it_p / steps_p, "newton_iters_per_step_learned": it_l / steps_l,
        "newton_iter_reduction": 1.0 - (it_l / steps_l) / (it_p / steps_p),
        "gm_steps_per_s_plain": steps_p / sp, "gm_steps_per_s_learned": steps_l / sl,
        "proposals": int(sum(s["proposals"] for s in stl)), "fallbacks": int(sum(s["fallbacks"] for s in stl)),
        "records_same_stop": int(sum(same)),
        "records_within_bar": within, "records_outside_bar_count": len(flagged),
        "outside_bar_checked_on_oracle": len([f for f in checked if "oracle_start_sensitivity" in f]),
        "outside_bar_oracle_also_sensitive": len([f for f in checked if f.get("oracle_start_sensitivity")]),
        "records_outside_bar": flagged,
    })
    return out


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--gm-batch", type=int, default=256, help="records in the batched GM-steps leg (0 = skip)")
    ap.add_argument("--gm-steps", type=int, default=1000)
    ap.add_argument("--ens-motions", type=int, default=256,
                    help="motions per frame in the ensemble leg (0 = skip; 256 is the ensemble configuration's count)")
    ap.add_argument("--ens-steps", type=int, default=3000,
                    help="step cap of the ensemble records (3000: none; their lengths are 1000-3000 steps)")
    ap.add_argument("--ens-frames", default="0,1,2", help="frame variants of the ensemble (0 4x3, 1 8x4, 2 12x6)")
    ap.add_argument("--ens-schedule", choices=("auto", "rounds", "queue"), default="auto")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="pinned oracle processes of the ensemble's CPU baseline (at most the allowed CPUs)")
    ap.add_argument("--learned-train", type=int, default=256, help="training records of the learned leg (0 = skip)")
    ap.add_argument("--learned-test", type=int, default=64)
    ap.add_argument("--learned-steps", type=int, default=800)
    ap.add_argument("--learned-epochs", type=int, default=30)
    ap.add_argument("--f3-steps", type=int, default=10, help="timed steps of the 3D frame transient leg (0 = skip)")
    ap.add_argument("--f3-warmup", type=int, default=2)
    ap.add_argument("--f3-cpu-steps", type=int, default=6, help="steps of the 3D frame leg's CPU baseline sample")
    ap.add_argument("--f3-records", type=int, default=32, help="record pairs run together in the 3D frame leg (1 = skip)")
    ap.add_argument("--f4-steps", type=int, default=2, help="timed steps of the tower leg (0 = skip)")
    ap.add_argument("--f4-warmup", type=int, default=1)
    ap.add_argument("--f4-records", type=int, default=16, help="records run together in the tower leg")
    ap.add_argument("--f4-modal", type=int, default=100, help="Jacobi rotations of the tower's modal analysis (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(ensembles shard by instance and one large model by floors across GPUs; that path is not built)",
              file=sys.stderr)
        return 2

    sec = section_leg(args)
    _progress(f"section leg: {sec['value']:.4g} fiber updates/s")
    config = {
        "workload": "section population of the 3D RC frame (20 stories x 5x5 bays): "
                    f"{sec['pop']['columns']} columns + {sec['pop']['beams']} beams x 5 IPs = {sec['n_pts']} "
                    f"3D fiber sections, {sec['fibers']} fibers; step = trial evaluation of every section at a "
                    "new seeded deformation from a committed preload history (Strut's fiber micro-benchmark "
                    "unit, no commit in the timed steps)",
        "sections": sec["n_pts"],
        "fibers": sec["fibers"],
        "fibers_per_section": [len(c) for c in sec["cells"]],
        "aux_pass": sec["aux"],
        "commit": sec["commit"],
    }
    if args.gm_batch > 0:
        gm = gm_leg(args.gm_batch, args.gm_steps)
        _progress(f"GM-steps leg: batched/plain {gm['batched_vs_plain']:.3g}")
        config["gm_steps"] = {"workload": f"2D RC frame 3 stories x 1 bay, forceBeamColumn 5 IPs, {args.gm_steps} "
                                          "steps dt 0.01 s, PGA 0.4 g seeded ground motions (seeds 0..N-1)",
                              **gm}
    if args.f3_steps > 0:
        f3 = frame3d_leg(args.f3_steps, args.f3_warmup, records=args.f3_records)
        _progress(f"3D frame leg: {f3['steps_per_s']:.3g} steps/s, {f3['ms_per_newton_iter']:.3g} ms per Newton iteration")
        if f3["records_together"]:
            _progress(f"3D frame leg, {args.f3_records} records together: "
                      f"{f3['records_together']['record_steps_per_s']:.3g} record-steps/s "
                      f"({f3['records_together']['vs_one_record']:.3g}x one record)")
        config["frame3d"] = {"workload": "3D RC frame 20 stories x 5 x 5 bays, forceBeamColumn3d columns (PDelta) and "
                                         "dispBeamColumn3d beams, 5 IPs, 268-fiber sections, rigid diaphragms "
                                         "(Transformation), two seeded records 0.5 g / 0.35 g, dt 0.005 s", **f3}
    if args.f4_steps > 0:
        f4 = tower_leg(args.f4_steps, args.f4_warmup, args.f4_records, args.f4_modal)
        _progress(f"tower leg: {f4['record_steps_per_s']:.3g} record-steps/s over {f4['records']} records, "
                  f"{f4['ms_per_lockstep_iter']:.4g} ms per lockstep Newton iteration (band solve "
                  f"{f4['band_solve_ms_avg']:.4g} ms, element state {f4['element_state_ms_avg']:.4g} ms)")
        config["tower"] = {"workload": "configuration 4: 3D RC tower 60 stories x 6 x 6 bays (forceBeamColumn3d "
                                       "columns, dispBeamColumn3d beams, 5 IPs, 152-fiber sections) around a 2 x "
                                       "2-bay elastic ShellMITC4 core wall (0.4 m, E 30 GPa, nu 0.2), rigid "
                                       "diaphragms, bidirectional seeded records, dt 0.005 s", **f4}
    if args.ens_motions > 0:
        from strut_amd import _native
        sched = {"auto": _native.SCHEDULE_AUTO, "rounds": _native.SCHEDULE_ROUNDS,
                 "queue": _native.SCHEDULE_QUEUE}[args.ens_schedule]
        config["ensemble"] = ensemble_leg(args.ens_motions, args.ens_steps, cpu_workers=args.cpu_workers,
                                          cpu=not args.no_cpu_baseline,
                                          frame_ids=tuple(int(v) for v in args.ens_frames.split(",")),
                                          schedule=sched)
        e = config["ensemble"]
        _progress(f"ensemble leg: {e['gm_steps_per_s']:.4g} GM-steps/s, roofline {e['roofline']['frac']:.3g}, "
                  f"batched/one-by-one {e['batched_vs_one_by_one']:.3g} (by steps "
                  f"{e['batched_vs_one_by_one_by_steps']:.3g})")
        if "cpu_baseline" in e:
            _progress(f"ensemble CPU baseline: {e['cpu_baseline']['value']:.4g} GM-steps/s on "
                      f"{e['cpu_baseline']['cores']} processes; high-PGA checks {e['high_pga_stop_checks']}")
    if args.learned_train > 0:
        config["learned"] = learned_leg(args.learned_train, args.learned_test, args.learned_steps,
                                        args.learned_epochs, sens_records=0 if args.no_cpu_baseline else 16,
                                        cpu_workers=args.cpu_workers)
        ll = config["learned"]
        _progress(f"learned leg: iteration reduction {ll['newton_iter_reduction']:.3g}, GM-steps/s "
                  f"{ll['gm_steps_per_s_learned']:.4g} vs {ll['gm_steps_per_s_plain']:.4g}, within bar "
                  f"{ll['records_within_bar']}, outside {ll['records_outside_bar_count']} (oracle as sensitive: "
                  f"{ll['outside_bar_oracle_also_sensitive']} of {ll['outside_bar_checked_on_oracle']} checked)")
    cpu_main = None
    if not args.no_cpu_baseline:  # the single-core baselines, one after another in one pinned process
        pool, _ = cpu_pool(1)
        with pool:
            cpu_main = pool.submit(_cpu_task, "sections").result()
            _progress(f"section CPU baseline: {cpu_main['value']:.4g} fiber updates/s")
            if "gm_steps" in config:
                config["gm_steps"]["cpu_baseline"] = pool.submit(_cpu_task, "gm", args.gm_steps).result()
            if "frame3d" in config:
                config["frame3d"]["cpu_baseline"] = pool.submit(_cpu_task, "frame3d", args.f3_cpu_steps).result()
                _progress(f"3D frame leg CPU baseline: {config['frame3d']['cpu_baseline']['value']:.3g} steps/s")
        cpu_main.update(host_cpu())
    out = {
        "metric": "fiber updates/s",
        "value": sec["value"],
        "unit": "fiber updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "roofline": sec["roofline"],
    }
    if cpu_main is not None:
        out["cpu_baseline"] = cpu_main
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Host-side driver of the device transient for 2D fiber frames.

``FrameBatch`` registers frames (strut_amd.frame2d.Frame2dCase) and any
number of ground-motion instances per frame, then runs them all on one GPU
through the C ABI (sgpu_frame2d_*).  It mirrors what Strut's
run_transient_nonlinear (transient_nonlinear.mojo:830) does for one case,
for a whole ensemble at once: every instance advances its own Newmark /
Newton loop, one Newton iteration per round, with its own step count and
convergence.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p
from typing import List, Optional

import numpy as np

from . import _native
from ._native import StrutGpuError, dptr, iptr
from .frame2d import MAX_NIP, Frame2dCase
from .sections import SectionRuntime


class FrameDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", c_int32), ("x", POINTER(c_double)), ("y", POINTER(c_double)),
        ("constrained", POINTER(c_int32)), ("mass", POINTER(c_double)),
        ("n_elem", c_int32), ("elem_nodes", POINTER(c_int32)), ("elem_def", POINTER(c_int32)),
        ("elem_nip", POINTER(c_int32)), ("elem_xi", POINTER(c_double)), ("elem_w", POINTER(c_double)),
        ("dt", c_double), ("gamma", c_double), ("beta", c_double), ("tol", c_double),
        ("max_iters", c_int32),
        ("alpha_m", c_double), ("beta_k", c_double), ("beta_k_init", c_double),
        ("exc_dir", c_int32),
        ("n_record", c_int32), ("record_dofs", POINTER(c_int32)), ("elem_type", POINTER(c_int32)),
    ]


class PathSeries(ctypes.Structure):
    _fields_ = [("n", c_int32), ("values", POINTER(c_double)), ("dt", c_double), ("factor", c_double),
                ("start_time", c_double), ("use_last", c_int32)]


def _declare(lib):
    if getattr(lib, "_frame2d_declared", False):
        return
    sig = {
        "sgpu_frame2d_create": (ctypes.c_int, [c_void_p, POINTER(FrameDesc), POINTER(c_int32)]),
        "sgpu_frame2d_add_instance": (ctypes.c_int, [c_void_p, c_int32, c_int32, POINTER(PathSeries),
                                                     POINTER(c_int32)]),
        "sgpu_frame2d_finalize": (ctypes.c_int, [c_void_p]),
        "sgpu_frame2d_launch": (ctypes.c_int, [c_void_p, c_int32]),
        "sgpu_frame2d_run": (ctypes.c_int, [c_void_p, c_int32]),
        "sgpu_frame2d_result": (ctypes.c_int, [c_void_p, c_int32, POINTER(c_double), POINTER(c_int32),
                                               POINTER(c_int32), POINTER(c_int32)]),
        "sgpu_frame2d_info": (ctypes.c_int, [c_void_p, c_int32, POINTER(c_int32)]),
        "sgpu_frame2d_launch_persistent": (ctypes.c_int, [c_void_p, c_int32]),
        "sgpu_frame2d_set_schedule": (ctypes.c_int, [c_void_p, ctypes.c_int]),
        "sgpu_frame2d_schedule_info": (ctypes.c_int, [c_void_p, POINTER(c_double)]),
        "sgpu_frame2d_record_samples": (ctypes.c_int, [c_void_p, c_int32]),
        "sgpu_frame2d_samples": (ctypes.c_int, [c_void_p, c_int32, c_void_p, c_void_p, ctypes.c_int64,
                                                POINTER(ctypes.c_int64)]),
        "sgpu_frame2d_set_predictor": (ctypes.c_int, [c_void_p, c_int32, c_void_p]),
        "sgpu_frame2d_inst_stats": (ctypes.c_int, [c_void_p, c_int32, POINTER(c_int32)]),
        "sgpu_frame2d_section_evals": (ctypes.c_int, [c_void_p, c_int32, c_int32, POINTER(c_int64)]),
        "sgpu_frame2d_launch_queue": (ctypes.c_int, [c_void_p, c_int32]),
        "sgpu_frame2d_envelopes": (ctypes.c_int, [c_void_p, c_int32, ctypes.c_int, POINTER(c_int32), POINTER(c_int32),
                                                  POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._frame2d_declared = True


class FrameBatch:
    """Frames x ground motions on one device context."""

    def __init__(self, materials, device: int = 0):
        self.rt = SectionRuntime(materials, device=device)
        self.lib = self.rt.lib
        _declare(self.lib)
        self.frames: List[Frame2dCase] = []
        self.frame_defs: List[dict] = []
        self.n_record: List[int] = []
        self.insts: List[tuple] = []
        self._keep = []
        self.finalized = False

    def _check(self, rc: int) -> int:
        return self.rt._check(rc)

    def add_frame(self, fc: Frame2dCase, record_all: bool = False) -> int:
        defs = {sid: self.rt.add_definition(2, fc.section_cells[sid]) for sid in fc.section_ids}
        rec = np.arange(fc.total_dofs, dtype=np.int32) if record_all else fc.record_dofs.astype(np.int32)
        k = dict(x=np.ascontiguousarray(fc.x), y=np.ascontiguousarray(fc.y),
                 con=np.ascontiguousarray(fc.constrained, dtype=np.int32), mass=np.ascontiguousarray(fc.mass),
                 en=np.ascontiguousarray(fc.elem_nodes.reshape(-1), dtype=np.int32),
                 ed=np.array([defs[int(s)] for s in fc.elem_section], dtype=np.int32),
                 nip=np.ascontiguousarray(fc.elem_nip, dtype=np.int32),
                 xi=np.ascontiguousarray(fc.elem_xi.reshape(-1)), w=np.ascontiguousarray(fc.elem_w.reshape(-1)),
                 rec=np.ascontiguousarray(rec),
                 et=np.ascontiguousarray(fc.elem_type if len(fc.elem_type) else np.zeros(len(fc.elem_nip)),
                                         dtype=np.int32))
        assert fc.elem_xi.shape[1] == MAX_NIP
        d = FrameDesc()
        d.n_nodes = fc.n_nodes
        d.x, d.y = dptr(k["x"]), dptr(k["y"])
        d.constrained = iptr(k["con"])
        d.mass = dptr(k["mass"])
        d.n_elem = len(fc.elem_nip)
        d.elem_nodes, d.elem_def, d.elem_nip = iptr(k["en"]), iptr(k["ed"]), iptr(k["nip"])
        d.elem_xi, d.elem_w = dptr(k["xi"]), dptr(k["w"])
        d.dt, d.gamma, d.beta, d.tol, d.max_iters = fc.dt, fc.gamma, fc.beta, fc.tol, fc.max_iters
        d.alpha_m, d.beta_k, d.beta_k_init, d.exc_dir = fc.alpha_m, fc.beta_k, fc.beta_k_init, fc.exc_dir
        d.n_record = len(rec)
        d.record_dofs = iptr(k["rec"])
        d.elem_type = iptr(k["et"])
        fid = c_int32()
        self._check(self.lib.sgpu_frame2d_create(self.rt.ctx, ctypes.byref(d), ctypes.byref(fid)))
        self.frames.append(fc)
        self.n_record.append(len(rec))
        self._keep.append(k)
        return fid.value

    def add_motion(self, frame_id: int, values, dt: float, steps: int, factor: float = 1.0,
                   start_time: float = 0.0, use_last: bool = False) -> int:
        v = np.ascontiguousarray(values, dtype=np.float64)
        ps = PathSeries(len(v), dptr(v), dt, factor, start_time, 1 if use_last else 0)
        iid = c_int32()
        self._check(self.lib.sgpu_frame2d_add_instance(self.rt.ctx, frame_id, steps, ctypes.byref(ps),
                                                       ctypes.byref(iid)))
        self.insts.append((frame_id, steps))
        return iid.value

    def add_case_motion(self, frame_id: int, fc: Frame2dCase, steps: Optional[int] = None) -> int:
        return self.add_motion(frame_id, fc.ts_values, fc.ts_dt, fc.steps if steps is None else steps,
                               fc.ts_factor, fc.ts_start, bool(fc.ts_use_last))

    def finalize(self) -> None:
        self._check(self.lib.sgpu_frame2d_finalize(self.rt.ctx))
        self.finalized = True

    def launch(self, rounds: int) -> None:
        """Enqueue `rounds` rounds (one Newton iteration of every active
        instance each), replayed from a captured HIP graph."""
        self._check(self.lib.sgpu_frame2d_launch(self.rt.ctx, rounds))

    def launch_persistent(self, max_steps: int) -> None:
        """Enqueue the persistent schedule: every unfinished instance advances
        up to max_steps steps on its own."""
        self._check(self.lib.sgpu_frame2d_launch_persistent(self.rt.ctx, max_steps))

    def launch_queue(self, max_steps: int) -> None:
        """Enqueue the work-queue schedule: every unfinished instance advances
        up to max_steps steps, its elements taken as tasks by persistent
        worker blocks."""
        self._check(self.lib.sgpu_frame2d_launch_queue(self.rt.ctx, max_steps))

    def set_schedule(self, schedule: int) -> None:
        """_native.SCHEDULE_AUTO (timed choice), SCHEDULE_ROUNDS, SCHEDULE_PERSISTENT or SCHEDULE_QUEUE."""
        self._check(self.lib.sgpu_frame2d_set_schedule(self.rt.ctx, schedule))

    def schedule_info(self) -> dict:
        out = np.zeros(5)
        self._check(self.lib.sgpu_frame2d_schedule_info(self.rt.ctx, dptr(out)))
        names = {0: "auto", 1: "rounds", 2: "persistent", 3: "queue"}
        return {"schedule": names.get(int(out[0]), "?"), "rounds_gm_steps_per_s": out[1],
                "persistent_gm_steps_per_s": out[2], "queue_gm_steps_per_s": out[4], "tunings": int(out[3])}

    def run(self, rounds: int = 64) -> None:
        """Run until every instance has finished or stopped, checking after
        every `rounds` rounds."""
        self._check(self.lib.sgpu_frame2d_run(self.rt.ctx, rounds))

    def synchronize(self) -> None:
        self.rt.synchronize()

    def result(self, inst: int) -> dict:
        frame_id, steps = self.insts[inst]
        nrec = self.n_record[frame_id]
        u = np.zeros((steps, nrec))
        it = np.zeros(steps, dtype=np.int32)
        st, done = c_int32(), c_int32()
        self._check(self.lib.sgpu_frame2d_result(self.rt.ctx, inst, dptr(u), iptr(it), ctypes.byref(st),
                                                 ctypes.byref(done)))
        return {"u": u, "iters": it, "status": st.value, "steps_done": done.value}

    def envelopes(self, frame_id: int, col_i=(), col_j=(), dx=()):
        """Envelopes of every instance of a frame over its completed steps,
        computed on the device from the recorded history: (disp_env
        [instances, n_record, 3], drift_env [instances, n_drift, 3]) with
        min, max and largest magnitude; drift channel d is
        (u[col_j[d]] - u[col_i[d]]) / dx[d] over recorded columns."""
        n_inst = sum(1 for f, _ in self.insts if f == frame_id)
        ci = np.ascontiguousarray(col_i, dtype=np.int32)
        cj = np.ascontiguousarray(col_j, dtype=np.int32)
        d = np.ascontiguousarray(dx, dtype=np.float64)
        disp = np.zeros((n_inst, self.n_record[frame_id], 3))
        drift = np.zeros((n_inst, max(len(ci), 1), 3))
        self._check(self.lib.sgpu_frame2d_envelopes(self.rt.ctx, frame_id, len(ci), iptr(ci), iptr(cj), dptr(d),
                                                    dptr(disp), dptr(drift)))
        return disp, drift[:, :len(ci)]

    def info(self, frame_id: int) -> dict:
        out = np.zeros(4, dtype=np.int32)
        self._check(self.lib.sgpu_frame2d_info(self.rt.ctx, frame_id, iptr(out)))
        return {"free": int(out[0]), "hbw": int(out[1]), "record": int(out[2]), "total_dofs": int(out[3])}

    def record_samples(self, on: bool = True) -> None:
        """Record every converged step's (features, target) pair on the
        device, as training data of the learned predictor (before finalize)."""
        self._check(self.lib.sgpu_frame2d_record_samples(self.rt.ctx, 1 if on else 0))

    def samples(self, frame_id: int):
        """The recorded pairs of a frame's instances as device tensors
        x [count, 5 n + 2] and y [count, n] (float32)."""
        import torch
        cnt = ctypes.c_int64()
        self._check(self.lib.sgpu_frame2d_samples(self.rt.ctx, frame_id, None, None, 0, ctypes.byref(cnt)))
        n = self.info(frame_id)["free"]
        dev = torch.device("cuda", self.rt.device)
        x = torch.empty((max(cnt.value, 1), 5 * n + 2), dtype=torch.float32, device=dev)
        y = torch.empty((max(cnt.value, 1), n), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        self._check(self.lib.sgpu_frame2d_samples(self.rt.ctx, frame_id, c_void_p(x.data_ptr()),
                                                  c_void_p(y.data_ptr()), cnt.value, ctypes.byref(cnt)))
        return x[:cnt.value], y[:cnt.value]

    def set_predictor(self, frame_id: int, mlp) -> None:
        """Start every step of this frame's instances from the network's
        proposal (strut_amd.learned.Mlp; None: the plain Newmark predictor)."""
        self._check(self.lib.sgpu_frame2d_set_predictor(self.rt.ctx, frame_id, None if mlp is None else mlp.h))

    def inst_stats(self, inst: int) -> dict:
        out = np.zeros(4, dtype=np.int32)
        self._check(self.lib.sgpu_frame2d_inst_stats(self.rt.ctx, inst, iptr(out)))
        return {"iters": int(out[0]), "proposals": int(out[1]), "fallbacks": int(out[2]),
                "dense_fallbacks": int(out[3])}

    def section_evals(self) -> np.ndarray:
        """Section point evaluations of every instance so far (int64)."""
        out = np.zeros(max(len(self.insts), 1), dtype=np.int64)
        self._check(self.lib.sgpu_frame2d_section_evals(self.rt.ctx, 0, len(self.insts),
                                                         out.ctypes.data_as(POINTER(c_int64))))
        return out[:len(self.insts)]

    def close(self) -> None:
        self.rt.close()

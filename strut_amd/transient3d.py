"""Transient nonlinear analysis of one 3D fiber beam-column frame on the
GPU (sgpu_frame3d_*): a loaded Frame3dCase (strut_amd.frame3d) becomes the
context's fiber sections, sgpu_elem3d elements, expansion tables, masses and
excitation, and runs step by step on the device -- the driver
transient_nonlinear.mojo:830-3638 with the Transformation handler and the
bidirectional UniformExcitation extension.

There is no host fallback: without libstrut_gpu.so or a device the
constructor raises.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p

import numpy as np

from ._native import dptr, iptr
from .elements3d import Elements3d
from .frame3d import Frame3dCase
from .sections import SectionRuntime

MAX_PATTERNS = 3
STATUS_ELEMENT_FAILED = -1
STATUS_SINGULAR = -11


class Frame3dDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", c_int), ("sp_constrained", POINTER(c_int32)), ("mass", POINTER(c_double)), ("has_mpc", c_int),
        ("row_offsets", POINTER(c_int32)), ("dof_pool", POINTER(c_int32)), ("coeff_pool", POINTER(c_double)),
        ("slave", POINTER(c_int32)), ("n_elem", c_int), ("elem_ids", POINTER(c_int32)),
        ("elem_nodes", POINTER(c_int32)), ("dt", c_double), ("gamma", c_double), ("beta", c_double),
        ("tol", c_double), ("max_iters", c_int), ("alpha_m", c_double), ("beta_k", c_double),
        ("beta_k_init", c_double), ("n_patterns", c_int), ("exc_dir", c_int * MAX_PATTERNS),
        ("ts_n", c_int * MAX_PATTERNS), ("ts_values", POINTER(c_double) * MAX_PATTERNS),
        ("ts_dt", c_double * MAX_PATTERNS), ("ts_factor", c_double * MAX_PATTERNS), ("hbw_rule", c_int),
        ("n_shell", c_int), ("shell_nodes", POINTER(c_int32)), ("shell_xyz", POINTER(c_double)),
        ("shell_props", POINTER(c_double)),
    ]


def _declare(lib):
    if getattr(lib, "_f3_declared", False):
        return
    lib.sgpu_frame3d_create.restype = c_int
    lib.sgpu_frame3d_create.argtypes = [c_void_p, POINTER(Frame3dDesc), POINTER(c_int32), POINTER(c_int32)]
    lib.sgpu_frame3d_run.restype = c_int
    lib.sgpu_frame3d_run.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_int32), POINTER(c_int32),
                                     POINTER(c_int32), POINTER(c_int32)]
    lib.sgpu_frame3d_effective_band.restype = c_int
    lib.sgpu_frame3d_effective_band.argtypes = [c_void_p, POINTER(c_double)]
    lib.sgpu_band_solve.restype = c_int
    lib.sgpu_band_solve.argtypes = [c_void_p, c_int, c_int, c_int, POINTER(c_double), POINTER(c_double),
                                    POINTER(c_double), POINTER(c_int32)]
    lib.sgpu_frame3d_set_recorder.restype = c_int
    lib.sgpu_frame3d_set_recorder.argtypes = [c_void_p, c_int, POINTER(c_int32), c_int]
    lib.sgpu_frame3d_recorded.restype = c_int
    lib.sgpu_frame3d_recorded.argtypes = [c_void_p, POINTER(c_double), POINTER(c_int32)]
    lib.sgpu_frame3d_envelopes.restype = c_int
    lib.sgpu_frame3d_envelopes.argtypes = [c_void_p, c_int, POINTER(c_int32), POINTER(c_int32), POINTER(c_double),
                                           POINTER(c_double), POINTER(c_double)]
    lib.sgpu_frame3d_add_record.restype = c_int
    lib.sgpu_frame3d_add_record.argtypes = [c_void_p, POINTER(c_int32), POINTER(POINTER(c_double)), POINTER(c_int32),
                                            POINTER(c_double), POINTER(c_double), POINTER(c_int32)]
    lib.sgpu_frame3d_run_records.restype = c_int
    lib.sgpu_frame3d_run_records.argtypes = [c_void_p, c_int, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                             POINTER(c_int32)]
    lib.sgpu_frame3d_recorded_record.restype = c_int
    lib.sgpu_frame3d_recorded_record.argtypes = [c_void_p, c_int32, POINTER(c_double), POINTER(c_int32)]
    lib.sgpu_frame3d_envelopes_record.restype = c_int
    lib.sgpu_frame3d_envelopes_record.argtypes = [c_void_p, c_int32, c_int, POINTER(c_int32), POINTER(c_int32),
                                                  POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    lib.sgpu_frame3d_set_band_solver.restype = c_int
    lib.sgpu_frame3d_set_band_solver.argtypes = [c_void_p, c_int]
    lib.sgpu_frame3d_modal.restype = c_int
    lib.sgpu_frame3d_modal.argtypes = [c_void_p, c_int, c_int, c_double, POINTER(c_double), POINTER(c_double),
                                       POINTER(c_int32)]
    lib._f3_declared = True


class Frame3dRun:
    """One 3D frame on the device.  hbw_rule 0 takes the band from the element
    free map as the reference does; 1 keeps every coupling of the collapsed
    system."""

    def __init__(self, fc: Frame3dCase, device: int = 0, hbw_rule: int = 0, records=None, band_solver: int = 0):
        """records: further ground-motion records of the same frame, each a
        list of Pattern (one per pattern of the case, same directions); the
        case's own patterns are record 0.  band_solver 0 makes the
        reference's row swaps (its band LU and dense partial-pivot
        fallback); 1 eliminates without swaps (for comparison)."""
        self.fc = fc
        self.rt = SectionRuntime(fc.materials, device=device)
        _declare(self.rt.lib)
        defs = {sid: self.rt.add_definition(3, fc.section_cells[sid]) for sid in fc.section_ids}
        self.elems = Elements3d(self.rt)
        kinds = ("fbc", "dbc")
        extra = list(records or [])
        for rec in extra:
            if len(rec) != len(fc.patterns) or any(a.direction != b.direction for a, b in zip(rec, fc.patterns)):
                raise ValueError("every record gives one series per pattern of the case, same directions")
        all_ids = []
        for _ in range(1 + len(extra)):  # each record its own elements and section states
            ids = []
            for e in range(fc.n_elem):
                sid = int(fc.elem_section[e])
                nip = int(fc.elem_nip[e])
                vz = None if np.isnan(fc.elem_vecxz[e]).any() else fc.elem_vecxz[e]
                G, J = fc.section_gj[sid]
                i, j = fc.elem_nodes[e]
                ids.append(self.elems.add(kinds[int(fc.elem_kind[e])], defs[sid], fc.elem_xi[e, :nip],
                                          fc.elem_w[e, :nip], fc.coords[i], fc.coords[j], vecxz=vz, G=G, J=J,
                                          transf=("Linear", "PDelta")[int(fc.elem_transf[e])]))
            all_ids.append(np.array(ids, dtype=np.int32))
        ids = all_ids[0]
        self.rt.finalize()
        k = dict(sp=np.ascontiguousarray(fc.sp, dtype=np.int32), mass=np.ascontiguousarray(fc.mass),
                 ro=np.ascontiguousarray(fc.row_offsets, dtype=np.int32),
                 dp=np.ascontiguousarray(fc.dof_pool if len(fc.dof_pool) else np.zeros(1), dtype=np.int32),
                 cp=np.ascontiguousarray(fc.coeff_pool if len(fc.coeff_pool) else np.zeros(1), dtype=np.float64),
                 sl=np.ascontiguousarray(fc.slave, dtype=np.int32), ids=np.array(ids, dtype=np.int32),
                 en=np.ascontiguousarray(fc.elem_nodes.reshape(-1), dtype=np.int32),
                 ts=[np.ascontiguousarray(p.values, dtype=np.float64) for p in fc.patterns],
                 sn=np.ascontiguousarray(fc.shell_nodes.reshape(-1), dtype=np.int32),
                 sx=np.ascontiguousarray(fc.coords[fc.shell_nodes].reshape(-1), dtype=np.float64),
                 spp=np.ascontiguousarray(fc.shell_props.reshape(-1), dtype=np.float64))
        self._keep = k
        d = Frame3dDesc()
        d.n_nodes = fc.n_nodes
        d.sp_constrained, d.mass, d.has_mpc = iptr(k["sp"]), dptr(k["mass"]), 1 if fc.has_mpc else 0
        d.row_offsets, d.dof_pool, d.coeff_pool, d.slave = iptr(k["ro"]), iptr(k["dp"]), dptr(k["cp"]), iptr(k["sl"])
        d.n_elem, d.elem_ids, d.elem_nodes = fc.n_elem, iptr(k["ids"]), iptr(k["en"])
        d.dt, d.gamma, d.beta, d.tol, d.max_iters = fc.dt, fc.gamma, fc.beta, fc.tol, fc.max_iters
        d.alpha_m, d.beta_k, d.beta_k_init = fc.alpha_m, fc.beta_k, fc.beta_k_init
        if len(fc.patterns) > MAX_PATTERNS:
            raise ValueError(f"at most {MAX_PATTERNS} excitation patterns")
        d.n_patterns = len(fc.patterns)
        for q, p in enumerate(fc.patterns):
            d.exc_dir[q], d.ts_n[q], d.ts_values[q] = p.direction, len(k["ts"][q]), dptr(k["ts"][q])
            d.ts_dt[q], d.ts_factor[q] = p.dt, p.factor
        d.hbw_rule = hbw_rule
        d.n_shell = len(fc.shell_nodes)
        if d.n_shell:
            d.shell_nodes, d.shell_xyz, d.shell_props = iptr(k["sn"]), dptr(k["sx"]), dptr(k["spp"])
        hbw, nf = c_int32(), c_int32()
        self.rt._check(self.rt.lib.sgpu_frame3d_create(self.rt.ctx, ctypes.byref(d), ctypes.byref(hbw),
                                                       ctypes.byref(nf)))
        self.hbw, self.free_count = hbw.value, nf.value
        self.rt._check(self.rt.lib.sgpu_frame3d_set_band_solver(self.rt.ctx, band_solver))
        for rec_ids, rec in zip(all_ids[1:], extra):
            vals = [np.ascontiguousarray(p.values, dtype=np.float64) for p in rec]
            self._keep.setdefault("rec", []).append((rec_ids, vals))
            tv = (POINTER(c_double) * MAX_PATTERNS)(*[dptr(v) for v in vals])
            tn = (c_int32 * MAX_PATTERNS)(*[len(v) for v in vals])
            tdt = (c_double * MAX_PATTERNS)(*[p.dt for p in rec])
            tf = (c_double * MAX_PATTERNS)(*[p.factor for p in rec])
            out = c_int32()
            self.rt._check(self.rt.lib.sgpu_frame3d_add_record(self.rt.ctx, iptr(rec_ids), tv, tn, tdt, tf,
                                                               ctypes.byref(out)))
        self.n_records = 1 + len(extra)
        self.steps_done = 0
        self._rec = np.zeros(0, dtype=np.int32)
        self._rec_cap = 0

    def run(self, steps: int, record_history: bool = True) -> dict:
        u = np.zeros((steps, self.fc.total_dofs)) if record_history else None
        iters = np.zeros(steps, dtype=np.int32)
        done, status, fallbacks = c_int32(), c_int32(), c_int32()
        self.rt._check(self.rt.lib.sgpu_frame3d_run(self.rt.ctx, steps, dptr(u) if record_history else None,
                                                    iptr(iters), ctypes.byref(done), ctypes.byref(status),
                                                    ctypes.byref(fallbacks)))
        self.steps_done += done.value
        return {"status": status.value, "u": u, "iters": iters, "steps_done": done.value, "hbw": self.hbw,
                "free_count": self.free_count, "dense_fallbacks": fallbacks.value}

    def modal(self, num_modes: int, max_rotations: int = 100, tol: float = 1.0e-12):
        """run_modal_eigen (modal_eigen.mojo:143-362) on the device: the
        frame at rest, (eigenvalues ascending, modes num_modes x 6 n_nodes,
        Jacobi rotations made).  max_rotations 100 and tol 1e-12 are the
        reference's jacobi_symmetric_eigen defaults."""
        eig = np.zeros(num_modes)
        modes = np.zeros((num_modes, self.fc.total_dofs))
        rot = c_int32(0)
        self.rt._check(self.rt.lib.sgpu_frame3d_modal(self.rt.ctx, num_modes, max_rotations, tol, dptr(eig),
                                                      dptr(modes), ctypes.byref(rot)))
        return eig, modes, rot.value

    def effective_band(self) -> np.ndarray:
        """K + a1 C + a0 M in the band at the current displacement."""
        band = np.zeros((self.free_count, 2 * self.hbw + 1))
        self.rt._check(self.rt.lib.sgpu_frame3d_effective_band(self.rt.ctx, dptr(band)))
        return band

    def run_records(self, steps: int) -> dict:
        """`steps` more steps of every record together: iters [records,
        steps], steps_done, status, dense_fallbacks per record."""
        R = self.n_records
        iters = np.zeros((R, steps), dtype=np.int32)
        done, status, fb = (np.zeros(R, dtype=np.int32) for _ in range(3))
        self.rt._check(self.rt.lib.sgpu_frame3d_run_records(self.rt.ctx, steps, iptr(iters), iptr(done), iptr(status),
                                                            iptr(fb)))
        return {"iters": iters, "steps_done": done, "status": status, "dense_fallbacks": fb}

    def set_recorder(self, dofs, max_steps: int) -> None:
        """Record these full-DOF displacements after every converged step,
        on the device."""
        self._rec = np.ascontiguousarray(dofs, dtype=np.int32)
        self._rec_cap = max_steps
        self.rt._check(self.rt.lib.sgpu_frame3d_set_recorder(self.rt.ctx, len(self._rec), iptr(self._rec), max_steps))

    def recorded(self, record: int = 0) -> np.ndarray:
        hist = np.zeros((max(self._rec_cap, 1), max(len(self._rec), 1)))
        steps = c_int32()
        self.rt._check(self.rt.lib.sgpu_frame3d_recorded_record(self.rt.ctx, record, dptr(hist), ctypes.byref(steps)))
        return hist[:steps.value, :len(self._rec)]

    def envelopes(self, col_i=(), col_j=(), dx=(), record: int = 0):
        """(disp_env (n_record, 3), drift_env (n_drift, 3)): min, max and
        largest magnitude over the recorded steps, computed on the device."""
        ci = np.ascontiguousarray(col_i, dtype=np.int32)
        cj = np.ascontiguousarray(col_j, dtype=np.int32)
        d = np.ascontiguousarray(dx, dtype=np.float64)
        disp = np.zeros((max(len(self._rec), 1), 3))
        drift = np.zeros((max(len(ci), 1), 3))
        self.rt._check(self.rt.lib.sgpu_frame3d_envelopes_record(self.rt.ctx, record, len(ci), iptr(ci), iptr(cj),
                                                                 dptr(d), dptr(disp), dptr(drift)))
        return disp[:len(self._rec)], drift[:len(ci)]

    def close(self) -> None:
        self.rt.close()


def band_solve(rt: SectionRuntime, band: np.ndarray, rhs: np.ndarray, variant: int = 0):
    """Factor and solve one BandGeneral system on the device
    (sgpu_band_solve); returns (x, info)."""
    _declare(rt.lib)
    band = np.ascontiguousarray(band, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    n, w = band.shape
    x = np.zeros(n)
    info = c_int32()
    rt._check(rt.lib.sgpu_band_solve(rt.ctx, variant, n, (w - 1) // 2, dptr(band), dptr(rhs), dptr(x),
                                     ctypes.byref(info)))
    return x, info.value

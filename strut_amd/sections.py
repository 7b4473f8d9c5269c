"""Host-side mirror of Strut's fiber-section interface over libstrut_gpu.

``SectionRuntime`` owns one device context.  Its methods follow the
reference's section API (src/mojo/sections/fiber2d.mojo, fiber3d.mojo):

=====================================  =======================================
reference                              here
=====================================  =======================================
_build_fiber_section2d_def (:3323)     ``add_definition(2, cells)``
_build_fiber_section3d_def (3d:157)    ``add_definition(3, cells)``
fiber_section2d_runtime_alloc_instances ``alloc_instances``
fiber_section2d_batch_eval_same_def    ``eval`` (returns False on a singular
  (:3042)                                section, like the reference)
fiber_section2d_set_trial_from_offset  ``set_trial``
fiber_section2d_commit_runtime_all     ``commit_all`` / ``revert_all``
fiber_section2d_commit_from_offset     ``commit`` / ``revert``
=====================================  =======================================

Every call goes through the C ABI (include/strut_gpu.h); nothing here
computes a fiber response.
"""

from __future__ import annotations

import ctypes
from ctypes import c_int32, c_int64, c_void_p
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _native
from ._native import POINT_DTYPE, RESULT_DTYPE, StrutGpuError, dptr, iptr
from .model import MaterialTable


class SectionRuntime:
    """Fiber-section definitions, instances and their device state."""

    def __init__(self, materials: MaterialTable, device: int = 0):
        self.lib = _native.load()
        self.device = device
        self.ctx = c_void_p()
        rc = self.lib.sgpu_ctx_create(device, ctypes.byref(self.ctx))
        if rc != 0:
            raise StrutGpuError(rc, f"sgpu_ctx_create(device={device}) failed")
        self.materials = materials
        types, params = materials.arrays()
        self._check(self.lib.sgpu_material_defs_upload(self.ctx, len(types), iptr(types), dptr(params)))
        self.dims = []
        self.n_inst = []
        self.n_fibers = []
        self.finalized = False

    # -- plumbing -------------------------------------------------------
    def _check(self, rc: int) -> int:
        if rc < 0:
            msg = self.lib.sgpu_last_error(self.ctx)
            raise StrutGpuError(rc, msg.decode() if msg else "")
        return rc

    def close(self) -> None:
        if self.ctx:
            self.lib.sgpu_ctx_destroy(self.ctx)
            self.ctx = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(self.lib.sgpu_stream(self.ctx) or 0)

    def synchronize(self) -> None:
        self._check(self.lib.sgpu_synchronize(self.ctx))

    # -- definitions and instances ------------------------------------------
    def add_definition(self, dim: int, cells: np.ndarray) -> int:
        y = np.ascontiguousarray(cells["y"], dtype=np.float64)
        z = np.ascontiguousarray(cells["z"], dtype=np.float64)
        a = np.ascontiguousarray(cells["area"], dtype=np.float64)
        m = np.ascontiguousarray(cells["mat"], dtype=np.int32)
        out = c_int32()
        self._check(self.lib.sgpu_section_def_create(self.ctx, dim, len(y), dptr(y), dptr(z), dptr(a), iptr(m),
                                                     ctypes.byref(out)))
        self.dims.append(dim)
        self.n_inst.append(0)
        self.n_fibers.append(len(y))
        return out.value

    def alloc_instances(self, def_id: int, n: int) -> int:
        out = c_int32()
        self._check(self.lib.sgpu_section_instances_alloc(self.ctx, def_id, n, ctypes.byref(out)))
        self.n_inst[def_id] += n
        return out.value

    def finalize(self) -> None:
        self._check(self.lib.sgpu_sections_finalize(self.ctx))
        self.finalized = True

    def def_info(self, def_id: int) -> dict:
        out = np.zeros(8)
        self._check(self.lib.sgpu_section_def_info(self.ctx, def_id, dptr(out)))
        return {"fibers": int(out[0]), "slots": int(out[1]), "y_bar": out[2], "z_bar": out[3],
                "initial_flex_valid": bool(out[4]), "initial_flex": tuple(out[5:8])}

    def profile_enable(self, on: bool = True) -> None:
        """Start (and reset) or stop per-kernel event timing on this context."""
        self._check(self.lib.sgpu_profile_enable(self.ctx, 1 if on else 0))

    def profile_read(self, kernel: int) -> Tuple[float, int]:
        """(total milliseconds, launches) of one kernel since profile_enable;
        kernel ids are strut_amd._native.KERNEL_*."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.lib.sgpu_profile_read(self.ctx, kernel, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def totals(self) -> Tuple[int, int]:
        f, s = c_int64(), c_int64()
        self._check(self.lib.sgpu_fiber_totals(self.ctx, ctypes.byref(f), ctypes.byref(s)))
        return f.value, s.value

    # -- evaluation -------------------------------------------------------
    @staticmethod
    def points(def_ids, instances, eps0, kappa_z, kappa_y=None) -> np.ndarray:
        n = len(def_ids)
        pts = np.zeros(n, dtype=POINT_DTYPE)
        pts["def_id"] = def_ids
        pts["instance"] = instances
        pts["eps0"] = eps0
        pts["kappa_z"] = kappa_z
        pts["kappa_y"] = 0.0 if kappa_y is None else kappa_y
        return pts

    def eval(self, pts: np.ndarray) -> Tuple[bool, np.ndarray]:
        """Evaluate section points; returns (all flexibilities valid, results)."""
        pts = np.ascontiguousarray(pts, dtype=POINT_DTYPE)
        res = np.zeros(len(pts), dtype=RESULT_DTYPE)
        rc = self._check(self.lib.sgpu_section_eval(self.ctx, len(pts), pts.ctypes.data, res.ctypes.data))
        return rc == 0, res

    def eval_device(self, n: int, d_points: int, d_results: int) -> None:
        self._check(self.lib.sgpu_section_eval_device(self.ctx, n, c_void_p(d_points), c_void_p(d_results)))

    def set_trial(self, def_id: int, instance: int, eps0: float, kappa_z: float,
                  kappa_y: float = 0.0) -> np.ndarray:
        """One section: returns (N, Mz, k11, k12, k22) in 2D and
        (N, My, Mz, k11, k12, k13, k22, k23, k33) in 3D."""
        _, r = self.eval(self.points([def_id], [instance], [eps0], [kappa_z], [kappa_y]))
        if self.dims[def_id] == 2:
            return np.array([r["s"][0, 0], r["s"][0, 1], r["k"][0, 0], r["k"][0, 1], r["k"][0, 2]])
        return np.concatenate([r["s"][0], r["k"][0]])

    # -- commit / revert ----------------------------------------------------
    def commit_all(self) -> None:
        self._check(self.lib.sgpu_sections_commit_all(self.ctx))

    def revert_all(self) -> None:
        self._check(self.lib.sgpu_sections_revert_all(self.ctx))

    def _listed(self, fn, def_ids: Sequence[int], instances: Sequence[int]) -> None:
        d = np.ascontiguousarray(def_ids, dtype=np.int32)
        i = np.ascontiguousarray(instances, dtype=np.int32)
        self._check(fn(self.ctx, len(d), iptr(d), iptr(i)))

    def commit(self, def_ids: Sequence[int], instances: Sequence[int]) -> None:
        self._listed(self.lib.sgpu_sections_commit, def_ids, instances)

    def revert(self, def_ids: Sequence[int], instances: Sequence[int]) -> None:
        self._listed(self.lib.sgpu_sections_revert, def_ids, instances)

    def fiber_state(self, def_id: int, instance: int, committed: bool = False) -> np.ndarray:
        """(n_fibers, 3) array of (eps, sig, tangent) in input fiber order."""
        out = np.zeros((self.n_fibers[def_id], 3))
        self._check(self.lib.sgpu_fiber_state_read(self.ctx, def_id, instance, 1 if committed else 0, dptr(out)))
        return out

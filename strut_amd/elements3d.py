"""Host-side mirror of Strut's 3D fiber beam-column element calls over
libstrut_gpu.

=========================================================  ===================
reference                                                  here
=========================================================  ===================
force_beam_column3d_fiber_global_tangent_and_internal      ``Elements3d.state``
  (elements/force_beam_column3d.mojo:2549-2990)              (kind "fbc")
beam_column3d_fiber_global_tangent_and_internal            ``Elements3d.state``
  (elements/beam_column3d_nonlinear.mojo:357-575)            (kind "dbc")
loader's element state allocation (loader.mojo:1821-1850)  ``Elements3d.add``
section commit after a converged step                      ``Elements3d.commit``
  (transient_nonlinear.mojo:3081-3101)
=========================================================  ===================

Many elements are evaluated in one device launch (one workgroup each); the
force-based elements keep their trial basic state on the device between
calls, as the reference keeps force_basic_q across Newton iterations.
Every call goes through the C ABI; nothing here computes an element state.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p
from typing import Optional, Sequence

import numpy as np

from ._native import StrutGpuError, dptr, iptr
from .sections import SectionRuntime

MAX_NIP3D = 5
KINDS = {"fbc": 0, "dbc": 1}
TRANSF = {"Linear": 0, "PDelta": 1}


class Elem3dDesc(ctypes.Structure):
    _fields_ = [("type", c_int32), ("def_id", c_int32), ("nip", c_int32), ("transf", c_int32),
                ("xyz1", c_double * 3), ("xyz2", c_double * 3), ("has_vecxz", c_int32), ("pad_", c_int32),
                ("vecxz", c_double * 3), ("G", c_double), ("J", c_double), ("xi", c_double * MAX_NIP3D),
                ("w", c_double * MAX_NIP3D)]


def _declare(lib):
    if getattr(lib, "_elem3d_declared", False):
        return
    ip = POINTER(c_int32)
    dp = POINTER(c_double)
    sig = {
        "sgpu_elem3d_add": (c_int, [c_void_p, POINTER(Elem3dDesc), ip]),
        "sgpu_elem3d_state": (c_int, [c_void_p, c_int, ip, dp, dp, dp, ip]),
        "sgpu_elem3d_state_device": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
        "sgpu_elem3d_commit": (c_int, [c_void_p, c_int, ip]),
        "sgpu_elem3d_basic_state": (c_int, [c_void_p, c_int32, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._elem3d_declared = True


class Elements3d:
    """forceBeamColumn3d / dispBeamColumn3d elements on a SectionRuntime's
    3D fiber-section definitions (added before the runtime is finalized)."""

    def __init__(self, rt: SectionRuntime):
        self.rt = rt
        self.lib = rt.lib
        _declare(self.lib)
        self.nip = []
        self.kind = []

    def add(self, kind: str, def_id: int, xi: Sequence[float], w: Sequence[float], xyz1, xyz2, vecxz=None,
            G: float = 0.0, J: float = 0.0, transf: str = "Linear") -> int:
        d = Elem3dDesc()
        d.type = KINDS[kind]
        d.def_id = def_id
        d.nip = len(xi)
        d.transf = TRANSF[transf]
        for k in range(3):
            d.xyz1[k] = float(xyz1[k])
            d.xyz2[k] = float(xyz2[k])
        d.has_vecxz = 0 if vecxz is None else 1
        if vecxz is not None:
            for k in range(3):
                d.vecxz[k] = float(vecxz[k])
        d.G, d.J = float(G), float(J)
        if len(xi) > MAX_NIP3D:
            raise ValueError(f"at most {MAX_NIP3D} integration points per 3D element on the device")
        for k, (a, b) in enumerate(zip(xi, w)):
            d.xi[k] = a
            d.w[k] = b
        eid = c_int32()
        self.rt._check(self.lib.sgpu_elem3d_add(self.rt.ctx, ctypes.byref(d), ctypes.byref(eid)))
        self.rt.n_inst[def_id] += len(xi)  # the element's section instances
        self.nip.append(len(xi))
        self.kind.append(kind)
        return eid.value

    def state(self, ids, u):
        """Global tangents [n, 12, 12] and resisting forces [n, 12] of the
        listed elements at their end displacements u [n, 12]."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(len(ids), 12))
        k = np.zeros((len(ids), 144))
        f = np.zeros((len(ids), 12))
        st = np.zeros(len(ids), dtype=np.int32)
        rc = self.lib.sgpu_elem3d_state(self.rt.ctx, len(ids), iptr(ids), dptr(u), dptr(k), dptr(f), iptr(st))
        if rc != 0:
            msg = self.lib.sgpu_last_error(self.rt.ctx)
            raise StrutGpuError(rc, msg.decode() if msg else "")
        return k.reshape(len(ids), 12, 12), f

    def commit(self, ids) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self.rt._check(self.lib.sgpu_elem3d_commit(self.rt.ctx, len(ids), iptr(ids)))

    def basic_state(self, elem_id: int) -> np.ndarray:
        q = np.zeros(5 + 3 * self.nip[elem_id])
        self.rt._check(self.lib.sgpu_elem3d_basic_state(self.rt.ctx, elem_id, dptr(q)))
        return q

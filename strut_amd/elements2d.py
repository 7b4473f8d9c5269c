"""Host-side mirror of Strut's 2D fiber beam-column element calls over
libstrut_gpu, for a host that assembles its own system.

=========================================================  ===================
reference                                                  here
=========================================================  ===================
force_beam_column2d_fiber_global_tangent_and_internal      ``Elements2d.state``
  (elements/force_beam_column2d.mojo:1418-1919)              (kind "fbc")
dispBeamColumn2d state (disp_beam_column2d.mojo:114-254)    ``Elements2d.state``
                                                             (kind "dbc")
loader's element state allocation (loader.mojo:1760-1869)  ``Elements2d.add``
section commit after a converged step                      ``Elements2d.commit``
  (transient_nonlinear.mojo:3081-3101) and the committed
  basic-deformation sync (helpers.mojo:1832-1883)
=========================================================  ===================

Many elements are evaluated in one device launch (one workgroup each) by
the 2D frame path's own element routines; the force-based elements keep
their trial basic state on the device between calls, as the reference
keeps force_basic_q across Newton iterations.  Every call goes through the
C ABI; nothing here computes an element state.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p
from typing import Optional, Sequence

import numpy as np

from ._native import StrutGpuError, dptr, iptr
from .sections import SectionRuntime

MAX_NIP2D = 5
KINDS = {"fbc": 0, "dbc": 1}


class Elem2dDesc(ctypes.Structure):
    _fields_ = [("type", c_int32), ("def_id", c_int32), ("nip", c_int32), ("pad_", c_int32),
                ("xy1", c_double * 2), ("xy2", c_double * 2), ("xi", c_double * MAX_NIP2D),
                ("w", c_double * MAX_NIP2D)]


def _declare(lib):
    if getattr(lib, "_elem2d_declared", False):
        return
    ip = POINTER(c_int32)
    dp = POINTER(c_double)
    sig = {
        "sgpu_elem2d_add": (c_int, [c_void_p, POINTER(Elem2dDesc), ip]),
        "sgpu_elem2d_state": (c_int, [c_void_p, c_int, ip, dp, dp, dp, ip]),
        "sgpu_elem2d_commit": (c_int, [c_void_p, c_int, ip, dp]),
        "sgpu_elem2d_basic_state": (c_int, [c_void_p, c_int32, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._elem2d_declared = True


class Elements2d:
    """forceBeamColumn2d / dispBeamColumn2d elements on a SectionRuntime's
    2D fiber-section definitions (added before the runtime is finalized)."""

    def __init__(self, rt: SectionRuntime):
        self.rt = rt
        self.lib = rt.lib
        _declare(self.lib)
        self.nip = []
        self.kind = []

    def add(self, kind: str, def_id: int, xi: Sequence[float], w: Sequence[float], xy1, xy2) -> int:
        if len(xi) > MAX_NIP2D:
            raise ValueError(f"at most {MAX_NIP2D} integration points per 2D element on the device")
        d = Elem2dDesc()
        d.type = KINDS[kind]
        d.def_id = def_id
        d.nip = len(xi)
        for k in range(2):
            d.xy1[k] = float(xy1[k])
            d.xy2[k] = float(xy2[k])
        for k, (a, b) in enumerate(zip(xi, w)):
            d.xi[k] = a
            d.w[k] = b
        eid = c_int32()
        self.rt._check(self.lib.sgpu_elem2d_add(self.rt.ctx, ctypes.byref(d), ctypes.byref(eid)))
        self.rt.n_inst[def_id] += len(xi)  # the element's section instances
        self.nip.append(len(xi))
        self.kind.append(kind)
        return eid.value

    def state(self, ids, u):
        """Global tangents [n, 6, 6] and resisting forces [n, 6] of the
        listed (distinct) elements at their end displacements u [n, 6]."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(len(ids), 6))
        k = np.zeros((len(ids), 36))
        f = np.zeros((len(ids), 6))
        st = np.zeros(len(ids), dtype=np.int32)
        rc = self.lib.sgpu_elem2d_state(self.rt.ctx, len(ids), iptr(ids), dptr(u), dptr(k), dptr(f), iptr(st))
        if rc != 0:
            msg = self.lib.sgpu_last_error(self.rt.ctx)
            raise StrutGpuError(rc, msg.decode() if msg else "")
        return k.reshape(len(ids), 6, 6), f

    def commit(self, ids, u: Optional[np.ndarray] = None) -> None:
        """Commit the listed elements' sections; with their converged end
        displacements u [n, 6], re-sync the force-based basic deformations."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        ud = None
        if u is not None:
            u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(len(ids), 6))
            ud = dptr(u)
        self.rt._check(self.lib.sgpu_elem2d_commit(self.rt.ctx, len(ids), iptr(ids), ud))

    def basic_state(self, elem_id: int) -> np.ndarray:
        q = np.zeros(3 + 2 * self.nip[elem_id] + 3)
        self.rt._check(self.lib.sgpu_elem2d_basic_state(self.rt.ctx, elem_id, dptr(q)))
        return q

"""Host-side mirror of Strut's elastic shell element over libstrut_gpu.

=========================================================  ====================
reference                                                  here
=========================================================  ====================
shell4_mindlin_stiffness (elements/shell4.mojo:10-330)      ``shell4_stiffness``
=========================================================  ====================

The stiffness is formed on the device (one wavefront per shell) and read
back; nothing here computes it.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

from . import _native
from ._native import StrutGpuError, dptr


def _declare(lib):
    if getattr(lib, "_shell_declared", False):
        return
    fn = lib.sgpu_shell4_stiffness
    fn.restype = c_int
    fn.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    lib._shell_declared = True


def shell4_stiffness(xyz, props, device: int = 0) -> np.ndarray:
    """Global 24 x 24 stiffness of each elastic MITC4 shell.

    xyz: (n, 4, 3) node coordinates; props: (n, 3) rows (E, nu, h) of the
    shells' ElasticMembranePlateSection.  Returns (n, 24, 24), DOFs ux uy uz
    rx ry rz per node in the given node order."""
    lib = _native.load()
    _declare(lib)
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 12))
    props = np.ascontiguousarray(np.asarray(props, dtype=np.float64).reshape(-1, 3))
    if len(xyz) != len(props):
        raise ValueError("one (E, nu, h) row per shell")
    n = len(xyz)
    out = np.zeros((n, 24, 24))
    ctx = c_void_p()
    rc = lib.sgpu_ctx_create(device, ctypes.byref(ctx))
    if rc != 0:
        raise StrutGpuError(rc, f"sgpu_ctx_create(device={device}) failed")
    try:
        rc = lib.sgpu_shell4_stiffness(ctx, n, dptr(xyz), dptr(props), dptr(out))
        if rc < 0:
            msg = lib.sgpu_last_error(ctx)
            raise StrutGpuError(rc, msg.decode() if msg else "")
    finally:
        lib.sgpu_ctx_destroy(ctx)
    return out

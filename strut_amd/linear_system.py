"""BandGeneral linear system on the device, with the reference backend's
calls split (src/mojo/solver/run_case/linear_solver_backend.mojo):

=========================================  ================================
reference                                  here
=========================================  ================================
initialize_structure / initialize_symbolic ``BandSystem(rt, n, hbw)``
  (:1664-1685)
load_reduced_matrix (:1877-1913)           ``load(A)`` (dense n x n; entries
                                           outside the band are dropped, as
                                           :413-426 drops them) or
                                           ``load(band, band_layout=True)``
add_reduced_matrix(A, scale) (:1916-1981)  ``load(A, scale, add=True)``
factorize_loaded (:2174 -> :631-679 and    ``factor()`` -> info: 0 band LU,
  the dense fallback :652-657, :175-214)   1 the dense fallback's row swaps
                                           were taken, 2 singular (the
                                           reference aborts)
solve(rhs, x) (:2277-2306 -> :735-769)     ``solve(rhs)``
=========================================  ================================

Every call goes through the C ABI (sgpu_band_*); the factorization runs on
the GPU and stays there between solves.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p

import numpy as np

from ._native import StrutGpuError, dptr
from .sections import SectionRuntime


def _declare(lib):
    if getattr(lib, "_band_declared", False):
        return
    lib.sgpu_band_create.restype = c_int
    lib.sgpu_band_create.argtypes = [c_void_p, c_int, c_int, POINTER(c_void_p)]
    lib.sgpu_band_destroy.restype = c_int
    lib.sgpu_band_destroy.argtypes = [c_void_p]
    lib.sgpu_band_load.restype = c_int
    lib.sgpu_band_load.argtypes = [c_void_p, c_int, POINTER(c_double), c_double, c_int]
    lib.sgpu_band_load_device.restype = c_int
    lib.sgpu_band_load_device.argtypes = [c_void_p, c_void_p, c_double, c_int]
    lib.sgpu_band_factor.restype = c_int
    lib.sgpu_band_factor.argtypes = [c_void_p, POINTER(c_int32)]
    lib.sgpu_band_solve_factored.restype = c_int
    lib.sgpu_band_solve_factored.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double)]
    lib.sgpu_band_factors.restype = c_int
    lib.sgpu_band_factors.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_int32)]
    lib._band_declared = True


class BandSystem:
    """One n x n BandGeneral system of half bandwidth hbw on the context of
    a SectionRuntime."""

    def __init__(self, rt: SectionRuntime, n: int, hbw: int):
        self.rt = rt
        _declare(rt.lib)
        self.n, self.hbw = n, hbw
        self.h = c_void_p()
        rt._check(rt.lib.sgpu_band_create(rt.ctx, n, hbw, ctypes.byref(self.h)))

    def _check(self, rc: int) -> None:
        self.rt._check(rc)

    def load(self, values: np.ndarray, scale: float = 1.0, add: bool = False, band_layout: bool = False) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        want = (self.n, 2 * self.hbw + 1) if band_layout else (self.n, self.n)
        if v.shape != want:
            raise ValueError(f"expected {want}, got {v.shape}")
        self._check(self.rt.lib.sgpu_band_load(self.h, 0 if band_layout else 1, dptr(v), scale, 1 if add else 0))

    def load_device(self, d_values: int, scale: float = 1.0, add: bool = False) -> None:
        self._check(self.rt.lib.sgpu_band_load_device(self.h, c_void_p(d_values), scale, 1 if add else 0))

    def factor(self) -> int:
        info = c_int32()
        self._check(self.rt.lib.sgpu_band_factor(self.h, ctypes.byref(info)))
        return info.value

    def solve(self, rhs: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        if r.shape != (self.n,):
            raise ValueError(f"rhs of {self.n} values expected")
        x = np.zeros(self.n)
        self._check(self.rt.lib.sgpu_band_solve_factored(self.h, dptr(r), dptr(x)))
        return x

    def factors(self):
        """(U rows n x (2 hbw + 1) diagonal first, L n x hbw, piv n) of the
        last factorization."""
        u = np.zeros((self.n, 2 * self.hbw + 1))
        l = np.zeros((self.n, max(self.hbw, 1)))
        piv = np.zeros(self.n, dtype=np.int32)
        self._check(self.rt.lib.sgpu_band_factors(self.h, dptr(u), dptr(l), piv.ctypes.data_as(POINTER(c_int32))))
        return u, l[:, :self.hbw], piv

    def close(self) -> None:
        if self.h:
            if self.rt.ctx:  # a band never outlives its context (its stream and device are the context's)
                self.rt.lib.sgpu_band_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["BandSystem", "StrutGpuError"]

"""ctypes binding of the oracle (liboracle.so): TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker or the timed CPU baseline.  The
product library never calls into it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    dp, ip = POINTER(c_double), POINTER(c_int)
    sig = {
        "or_material_path": (None, [c_int, dp, c_int, dp, c_int, dp, dp]),
        "or_sec2d_create": (c_void_p, [c_int, dp, dp, ip, c_int, ip, dp]),
        "or_sec2d_free": (None, [c_void_p]),
        "or_sec2d_alloc_instances": (c_int, [c_void_p, c_int]),
        "or_sec2d_ybar": (c_double, [c_void_p]),
        "or_sec2d_set_trial": (c_int, [c_void_p, c_int, c_double, c_double, dp]),
        "or_sec2d_batch_eval": (c_int, [c_void_p, c_int, ip, dp, dp, dp, ip]),
        "or_sec2d_commit": (None, [c_void_p, c_int]),
        "or_sec2d_revert": (None, [c_void_p, c_int]),
        "or_sec2d_commit_all": (None, [c_void_p]),
        "or_sec2d_revert_all": (None, [c_void_p]),
        "or_sec2d_initial_flex": (c_int, [c_void_p, dp]),
        "or_sec2d_fiber_trial": (c_int, [c_void_p, c_int, c_int, dp]),
        "or_sec3d_create": (c_void_p, [c_int, dp, dp, dp, ip, c_int, ip, dp]),
        "or_sec3d_free": (None, [c_void_p]),
        "or_sec3d_alloc_instances": (c_int, [c_void_p, c_int]),
        "or_sec3d_set_trial": (c_int, [c_void_p, c_int, c_double, c_double, c_double, dp]),
        "or_sec3d_commit": (None, [c_void_p, c_int]),
        "or_sec3d_revert": (None, [c_void_p, c_int]),
        "or_sec3d_fiber_trial": (c_int, [c_void_p, c_int, c_int, dp]),
        "or_sec3d_centroid": (None, [c_void_p, dp]),
        "or_sec3d_flexibility": (c_int, [dp, dp]),
        "or_sec3d_eval_commit": (None, [c_void_p, c_int, ip, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(POINTER(c_double))


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(POINTER(c_int))


def material_path(mat_type: int, params, strains, commit_each: bool = True):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    s = np.ascontiguousarray(strains, dtype=np.float64)
    sig = np.zeros(len(s))
    tan = np.zeros(len(s))
    lib.or_material_path(mat_type, _d(p), len(s), _d(s), 1 if commit_each else 0, _d(sig), _d(tan))
    return sig, tan


class _Section:
    def __init__(self, cells, mat_types, mat_params, dim):
        self.lib = load()
        self._keep = [np.ascontiguousarray(cells[k]) for k in ("y", "z", "area")]
        y, z, a = (np.ascontiguousarray(cells[k], dtype=np.float64) for k in ("y", "z", "area"))
        m = np.ascontiguousarray(cells["mat"], dtype=np.int32)
        t = np.ascontiguousarray(mat_types, dtype=np.int32)
        p = np.ascontiguousarray(mat_params, dtype=np.float64).reshape(-1)
        self.n_fibers = len(y)
        if dim == 2:
            self.h = self.lib.or_sec2d_create(len(y), _d(y), _d(a), _i(m), len(t), _i(t), _d(p))
        else:
            self.h = self.lib.or_sec3d_create(len(y), _d(y), _d(z), _d(a), _i(m), len(t), _i(t), _d(p))
        if not self.h:
            raise ValueError("oracle section creation failed")


class Section2d(_Section):
    """Oracle FiberSection2d (fiber2d.mojo restatement)."""

    def __init__(self, cells, mat_types, mat_params):
        super().__init__(cells, mat_types, mat_params, 2)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_sec2d_free(self.h)
            self.h = None

    def alloc(self, n: int) -> int:
        return self.lib.or_sec2d_alloc_instances(self.h, n)

    @property
    def y_bar(self) -> float:
        return self.lib.or_sec2d_ybar(self.h)

    def initial_flex(self):
        out = np.zeros(3)
        ok = self.lib.or_sec2d_initial_flex(self.h, _d_out(out))
        return bool(ok), out

    def set_trial(self, inst: int, eps0: float, kappa: float) -> np.ndarray:
        out = np.zeros(5)
        rc = self.lib.or_sec2d_set_trial(self.h, inst, eps0, kappa, _d_out(out))
        if rc != 0:
            raise IndexError("instance out of range")
        return out

    def batch_eval(self, insts, eps0, kappa):
        n = len(insts)
        out = np.zeros((n, 8))
        valid = np.zeros(n, dtype=np.int32)
        rc = self.lib.or_sec2d_batch_eval(self.h, n, _i(insts), _d(eps0), _d(kappa), _d_out(out),
                                          valid.ctypes.data_as(POINTER(c_int)))
        return rc, out, valid

    def commit(self, inst: int) -> None:
        self.lib.or_sec2d_commit(self.h, inst)

    def revert(self, inst: int) -> None:
        self.lib.or_sec2d_revert(self.h, inst)

    def commit_all(self) -> None:
        self.lib.or_sec2d_commit_all(self.h)

    def revert_all(self) -> None:
        self.lib.or_sec2d_revert_all(self.h)

    def fiber_trial(self, inst: int) -> np.ndarray:
        out = np.zeros((self.n_fibers, 3))
        tmp = np.zeros(3)
        for f in range(self.n_fibers):
            rc = self.lib.or_sec2d_fiber_trial(self.h, inst, f, _d_out(tmp))
            out[f] = tmp if rc == 0 else 0.0
        return out


class Section3d(_Section):
    """Oracle FiberSection3d (fiber3d.mojo restatement)."""

    def __init__(self, cells, mat_types, mat_params):
        super().__init__(cells, mat_types, mat_params, 3)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_sec3d_free(self.h)
            self.h = None

    def alloc(self, n: int) -> int:
        return self.lib.or_sec3d_alloc_instances(self.h, n)

    def centroid(self):
        out = np.zeros(2)
        self.lib.or_sec3d_centroid(self.h, _d_out(out))
        return out

    def set_trial(self, inst: int, eps0: float, kappa_y: float, kappa_z: float) -> np.ndarray:
        out = np.zeros(9)
        rc = self.lib.or_sec3d_set_trial(self.h, inst, eps0, kappa_y, kappa_z, _d_out(out))
        if rc != 0:
            raise IndexError("instance out of range")
        return out

    def commit(self, inst: int) -> None:
        self.lib.or_sec3d_commit(self.h, inst)

    def eval_commit(self, insts, eps0, ky, kz) -> None:
        """Evaluate then commit a batch of instances in one C call (timing)."""
        i = np.ascontiguousarray(insts, dtype=np.int32)
        e, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (eps0, ky, kz))
        self.lib.or_sec3d_eval_commit(self.h, len(i), _i(i), _d(e), _d(y), _d(z))

    def revert(self, inst: int) -> None:
        self.lib.or_sec3d_revert(self.h, inst)

    def fiber_trial(self, inst: int) -> np.ndarray:
        out = np.zeros((self.n_fibers, 3))
        tmp = np.zeros(3)
        for f in range(self.n_fibers):
            rc = self.lib.or_sec3d_fiber_trial(self.h, inst, f, _d_out(tmp))
            out[f] = tmp if rc == 0 else 0.0
        return out


def flexibility3(k6) -> tuple:
    f = np.zeros(6)
    ok = load().or_sec3d_flexibility(_d(np.asarray(k6, dtype=np.float64)), _d_out(f))
    return bool(ok), f


def _d_out(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(POINTER(c_double))

"""ctypes binding of libstrut_gpu.so (include/strut_gpu.h).

The library is built in-tree by ``strut_amd/csrc/Makefile``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in the product
path.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libstrut_gpu.so")

SGPU_OK = 0
SGPU_SINGULAR_SECTION = 1
STATUS_NAMES = {
    0: "SGPU_OK",
    -1: "SGPU_ERR_ARG",
    -2: "SGPU_ERR_HIP",
    -3: "SGPU_ERR_NODEV",
    -4: "SGPU_ERR_STATE",
    1: "SGPU_SINGULAR_SECTION",
    2: "SGPU_ERR_LOCAL_NONCONV",
}

# Every symbol include/strut_gpu.h declares, checked by the CPU test suite.
EXPORTED = (
    "sgpu_abi_version",
    "sgpu_ctx_create",
    "sgpu_ctx_destroy",
    "sgpu_last_error",
    "sgpu_stream",
    "sgpu_synchronize",
    "sgpu_material_defs_upload",
    "sgpu_section_def_create",
    "sgpu_section_instances_alloc",
    "sgpu_sections_finalize",
    "sgpu_section_def_info",
    "sgpu_section_eval",
    "sgpu_section_eval_device",
    "sgpu_section_eval_device_noalloc",
    "sgpu_sections_commit_all",
    "sgpu_sections_revert_all",
    "sgpu_sections_commit",
    "sgpu_sections_revert",
    "sgpu_fiber_state_read",
    "sgpu_fiber_totals",
    "sgpu_frame2d_create",
    "sgpu_frame2d_add_instance",
    "sgpu_frame2d_finalize",
    "sgpu_frame2d_launch",
    "sgpu_frame2d_run",
    "sgpu_frame2d_result",
    "sgpu_frame2d_info",
    "sgpu_frame2d_launch_persistent",
    "sgpu_frame2d_set_schedule",
    "sgpu_frame2d_schedule_info",
    "sgpu_profile_enable",
    "sgpu_profile_read",
    "sgpu_mlp_create",
    "sgpu_mlp_destroy",
    "sgpu_mlp_param_count",
    "sgpu_mlp_layout",
    "sgpu_mlp_get_params",
    "sgpu_mlp_set_params",
    "sgpu_mlp_set_normalization",
    "sgpu_mlp_get_normalization",
    "sgpu_mlp_fit_normalization",
    "sgpu_mlp_gradients",
    "sgpu_mlp_train",
    "sgpu_mlp_predict",
    "sgpu_frame2d_record_samples",
    "sgpu_frame2d_samples",
    "sgpu_frame2d_set_predictor",
    "sgpu_frame2d_inst_stats",
    "sgpu_frame2d_section_evals",
    "sgpu_elem2d_add",
    "sgpu_elem2d_state",
    "sgpu_elem2d_commit",
    "sgpu_elem2d_basic_state",
    "sgpu_frame2d_launch_queue",
    "sgpu_elem3d_add",
    "sgpu_elem3d_state",
    "sgpu_elem3d_state_device",
    "sgpu_elem3d_commit",
    "sgpu_elem3d_basic_state",
    "sgpu_shell4_stiffness",
    "sgpu_frame3d_modal",
    "sgpu_mpc_set",
    "sgpu_mpc_collapse_matrix",
    "sgpu_mpc_collapse_matrix_device",
    "sgpu_mpc_collapse_vector",
    "sgpu_mpc_enforce",
    "sgpu_frame3d_create",
    "sgpu_frame3d_run",
    "sgpu_frame3d_effective_band",
    "sgpu_band_solve",
    "sgpu_frame3d_set_recorder",
    "sgpu_frame3d_recorded",
    "sgpu_frame3d_envelopes",
    "sgpu_frame2d_envelopes",
    "sgpu_frame3d_add_record",
    "sgpu_frame3d_run_records",
    "sgpu_frame3d_recorded_record",
    "sgpu_frame3d_envelopes_record",
    "sgpu_frame3d_set_band_solver",
    "sgpu_band_create",
    "sgpu_band_destroy",
    "sgpu_band_load",
    "sgpu_band_load_device",
    "sgpu_band_factor",
    "sgpu_band_solve_factored",
    "sgpu_band_factors",
)

# Kernel ids of sgpu_profile_read.
KERNEL_AUX_EVAL = 0
KERNEL_SECTION_EVAL = 1
KERNEL_COMMIT = 2
KERNEL_FRAME2D = 3
KERNEL_MPC = 4
KERNEL_F3_ELEM = 5
KERNEL_F3_SOLVE = 6

# Schedules of sgpu_frame2d_run.
SCHEDULE_AUTO = 0
SCHEDULE_ROUNDS = 1
SCHEDULE_PERSISTENT = 2
SCHEDULE_QUEUE = 3


class SectionPoint(ctypes.Structure):
    _fields_ = [
        ("def_id", c_int32),
        ("instance", c_int32),
        ("eps0", c_double),
        ("kappa_z", c_double),
        ("kappa_y", c_double),
    ]


class SectionResult(ctypes.Structure):
    _fields_ = [
        ("s", c_double * 3),
        ("k", c_double * 6),
        ("f", c_double * 6),
        ("valid", c_int32),
        ("pad_", c_int32),
    ]


POINT_DTYPE = np.dtype([("def_id", "<i4"), ("instance", "<i4"), ("eps0", "<f8"), ("kappa_z", "<f8"),
                        ("kappa_y", "<f8")])
RESULT_DTYPE = np.dtype([("s", "<f8", 3), ("k", "<f8", 6), ("f", "<f8", 6), ("valid", "<i4"), ("pad_", "<i4")])
assert POINT_DTYPE.itemsize == ctypes.sizeof(SectionPoint)
assert RESULT_DTYPE.itemsize == ctypes.sizeof(SectionResult)

_lib = None


def load() -> ctypes.CDLL:
    """Load libstrut_gpu.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C strut_amd/csrc` "
            "(or __graft_entry__.build()); the fiber-section path has no CPU fallback")
    # torch ships its own HIP and HSA runtimes under the same sonames as the
    # system ones.  Loading the library first would bring the system runtime
    # in, and torch's later would be a second runtime in the process (its
    # device initialisation then fails).  With torch loaded first, the
    # library binds to the runtime already present.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    dp = POINTER(c_double)
    ip = POINTER(c_int32)
    sig = {
        "sgpu_abi_version": (c_int, []),
        "sgpu_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
        "sgpu_ctx_destroy": (c_int, [c_void_p]),
        "sgpu_last_error": (ctypes.c_char_p, [c_void_p]),
        "sgpu_stream": (c_void_p, [c_void_p]),
        "sgpu_synchronize": (c_int, [c_void_p]),
        "sgpu_material_defs_upload": (c_int, [c_void_p, c_int, ip, dp]),
        "sgpu_section_def_create": (c_int, [c_void_p, c_int, c_int, dp, dp, dp, ip, ip]),
        "sgpu_section_instances_alloc": (c_int, [c_void_p, c_int32, c_int, ip]),
        "sgpu_sections_finalize": (c_int, [c_void_p]),
        "sgpu_section_def_info": (c_int, [c_void_p, c_int32, dp]),
        "sgpu_section_eval": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
        "sgpu_section_eval_device": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
        "sgpu_section_eval_device_noalloc": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
        "sgpu_sections_commit_all": (c_int, [c_void_p]),
        "sgpu_sections_revert_all": (c_int, [c_void_p]),
        "sgpu_sections_commit": (c_int, [c_void_p, c_int, ip, ip]),
        "sgpu_sections_revert": (c_int, [c_void_p, c_int, ip, ip]),
        "sgpu_fiber_state_read": (c_int, [c_void_p, c_int32, c_int32, c_int, dp]),
        "sgpu_fiber_totals": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64)]),
        "sgpu_profile_enable": (c_int, [c_void_p, c_int]),
        "sgpu_profile_read": (c_int, [c_void_p, c_int, POINTER(c_double), POINTER(c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.sgpu_abi_version() != 1:
        raise RuntimeError("libstrut_gpu ABI version mismatch")
    _lib = lib
    return lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_int32))


class StrutGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code

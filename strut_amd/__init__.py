"""MI355X (gfx950) implementation of Strut's nonlinear fiber-section state
determination.  The compute path is libstrut_gpu.so (HIP kernels behind the
C ABI in include/strut_gpu.h); this package is its host-side mirror."""

from .model import MaterialTable, discretize_section, make_uniaxial_def  # noqa: F401

__all__ = ["MaterialTable", "discretize_section", "make_uniaxial_def"]

"""MI355X (gfx950) implementation of plantFEM's static linear-elastic solve path.

FEMDomain / FEMSolver mirror the reference's FEMDomain_ / FEMSolver_ surface; every compute
step runs in libpfgpu.so (include/pfgpu.h).  There is no CPU fallback.
"""
from .fem import CRS, FEMDomain, FEMSolver, bicgstab_crs, cube3d, spmv_crs  # noqa: F401

__all__ = ["CRS", "FEMDomain", "FEMSolver", "bicgstab_crs", "cube3d", "spmv_crs"]

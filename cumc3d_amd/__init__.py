"""MI355X-native implementation of CUMC3D's solver/mhd update step.

The compute path is the HIP library libcumc3d_hip.so (cumc3d_amd/csrc, C-ABI in
include/cumc3d_hip.h); ``cumc3d_amd.solver`` binds it for Python callers and
``cumc3d_amd/fortran`` for the reference's Fortran driver.
"""
from .config import Config  # noqa: F401

__all__ = ["Config"]

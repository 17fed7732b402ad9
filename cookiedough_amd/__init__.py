"""cookiedough_amd -- MI355X (gfx950) renderer for cookiedough's per-pixel
effect loops.

The product is the native library libckd.so (HIP kernels behind the C ABI of
include/ckd.h).  This package holds its sources (csrc/), the build script,
a ctypes binding, a small Python context wrapper and the seeded synthetic
workloads of the benchmark; cookiedough_amd/shim holds the C++ drop-in that
exports the reference's own entry points.
"""
from ._lib import (CKD_EDEVICE, CKD_EINVAL, CKD_ENOMEM, CKD_ESTATE, CKD_FX_SDF, CKD_FX_TUNNEL,
                   CKD_FX_VOXEL, CKD_OK, CKD_SHADE_FP16, CKD_SHADE_FP32, CkdError, LandscapeCam,
                   Request, SdfScene, Tuning, load)

__all__ = ["Context", "CkdError", "LandscapeCam", "Request", "SdfScene", "Tuning", "load",
           "CKD_OK", "CKD_EINVAL", "CKD_ENOMEM", "CKD_EDEVICE", "CKD_ESTATE",
           "CKD_SHADE_FP32", "CKD_SHADE_FP16", "CKD_FX_SDF", "CKD_FX_VOXEL", "CKD_FX_TUNNEL"]


def __getattr__(name):
    if name == "Context":
        from .device import Context
        return Context
    raise AttributeError(name)

"""gpu_cvt_amd -- MI355X (gfx950) surface Lloyd / CVT remesher.

Drop-in for the Lloyd loop of Harper-Yan/GPU_CVT (main.cu:669-1073): a C ABI
(``include/cvt.h``, ``libcvt_hip.so``) over hand-written HIP kernels, and the
ctypes host mirror in :mod:`gpu_cvt_amd.cvt`.
"""
from .cvt import Job, Params, IterStats, CVTError, default_params, lib  # noqa: F401

__all__ = ["Job", "Params", "IterStats", "CVTError", "default_params", "lib"]

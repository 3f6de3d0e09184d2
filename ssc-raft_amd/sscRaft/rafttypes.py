"""ctypes binding of libraft.so (the reference binds its CUDA libraft.so the
same way, sscRaft/rafttypes.py:74-198,310-323) and the array helpers
CNICE / FilterNumber (rafttypes.py:485-519).

Every call checks the library's thread-local error state and raises
RuntimeError (ValueError for invalid parameters) instead of the reference's
process exit.
"""
import ctypes
import logging
import os

import numpy

logger = logging.getLogger("sscRaft")

__all__ = ["libraft", "CNICE", "FilterNumber", "raft_check", "LIB_PATH", "logger", "numpy", "ctypes", "hip_runtime_path"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libraft.so")

# One HIP runtime per process: torch ships its own libamdhip64 (same SONAME,
# libamdhip64.so.7).  Loading torch first makes libraft.so bind to that copy,
# so device pointers from torch tensors and libraft's kernels share one
# runtime.  Without torch, the system ROCm runtime is used.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libraft.so not found at {LIB_PATH}: build it first (python -c 'import __graft_entry__ as g; "
                      f"g.build()' or make -C ssc-raft_amd/csrc)")

libraft = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float

libraft.getFBPMultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp]
libraft.getFBPMultiGPU.restype = None
libraft.get_eEM_RT_MultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp]
libraft.get_eEM_RT_MultiGPU.restype = None
libraft.get_eEM_RT_KL_MultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.get_eEM_RT_KL_MultiGPU.restype = None
libraft.getRadonRTMultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _f]
libraft.getRadonRTMultiGPU.restype = None
libraft.raft_last_error.argtypes = []
libraft.raft_last_error.restype = _i
libraft.raft_last_error_string.argtypes = []
libraft.raft_last_error_string.restype = ctypes.c_char_p
libraft.raft_clear_error.argtypes = []
libraft.raft_clear_error.restype = None
libraft.raft_fbp_device.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _i, _vp]
libraft.raft_fbp_device.restype = _i
libraft.raft_fbp_device_noalloc.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _i, _vp]
libraft.raft_fbp_device_noalloc.restype = _i
libraft.raft_lds_read_probe.argtypes = [_i, _vp, _vp, _vp]
libraft.raft_lds_read_probe.restype = _i
libraft.raft_radon_device.argtypes = [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _f, _vp]
libraft.raft_radon_device.restype = _i
libraft.raft_eem_device.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.raft_eem_device.restype = _i
libraft.raft_eem_device_kl.argtypes = [_i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp]
libraft.raft_eem_device_kl.restype = _i
libraft.getBSTMultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i]
libraft.getBSTMultiGPU.restype = None
libraft.raft_bst_device.argtypes = [_i, _vp, _vp, _vp, _vp, _vp]
libraft.raft_bst_device.restype = _i
libraft.raft_batch_device.argtypes = [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _vp]
libraft.raft_batch_device.restype = _i
libraft.getBackgroundCorrectionMultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i]
libraft.getBackgroundCorrectionMultiGPU.restype = None
libraft.raft_background_device.argtypes = [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]
libraft.raft_background_device.restype = _i
libraft.get_tEM_RT_MultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.get_tEM_RT_MultiGPU.restype = None
libraft.raft_tem_device.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.raft_tem_device.restype = _i
libraft.getFilterLowPassMultiGPU.argtypes = [_vp, _i, _vp, _vp, _vp]
libraft.getFilterLowPassMultiGPU.restype = None
libraft.raft_lowpass_device.argtypes = [_i, _vp, _vp, _vp, _vp]
libraft.raft_lowpass_device.restype = _i
libraft.raft_eem_kl_device.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.raft_eem_kl_device.restype = _i
libraft.raft_fbp_plan_info.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
libraft.raft_fbp_plan_info.restype = _i
libraft.raft_proj_plan_info.argtypes = [_vp, _i, _i, _i, _i, _i, _f, _f, _i, _vp, _vp]
libraft.raft_proj_plan_info.restype = _i
libraft.raft_autotune_dump.argtypes = [ctypes.c_char_p, _i]
libraft.raft_autotune_dump.restype = _i
libraft.raft_autotune_clear.argtypes = []
libraft.raft_autotune_clear.restype = None
libraft.raft_eem_device_sched.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp]
libraft.raft_eem_device_sched.restype = _i
libraft.raft_set_profiling.argtypes = [_i]
libraft.raft_set_profiling.restype = None
libraft.raft_last_kernel_ms.argtypes = [_i]
libraft.raft_last_kernel_ms.restype = ctypes.c_double

RAFT_EINVAL = 1


def raft_check(what):
    """Raise if the last libraft call on this thread failed."""
    code = libraft.raft_last_error()
    if code != 0:
        msg = libraft.raft_last_error_string().decode(errors="replace")
        exc = ValueError if code == RAFT_EINVAL else RuntimeError
        raise exc(f"{what}: {msg} (code {code})")


def CNICE(darray, dtype=numpy.float32):
    """C-contiguous array of `dtype`, copying only when needed (rafttypes.py:485-491)."""
    if darray.dtype != dtype:
        return numpy.ascontiguousarray(darray.astype(dtype))
    if not darray.flags["C_CONTIGUOUS"]:
        return numpy.ascontiguousarray(darray)
    return darray


_FILTERS = {"none": 0, "gaussian": 1, "lorentz": 2, "cosine": 3, "rectangle": 4, "hann": 5, "hamming": 6,
            "ramp": 7}


def FilterNumber(mfilter):
    """Filter name -> number; unknown names map to hamming (6), as rafttypes.py:501-519."""
    return _FILTERS.get(mfilter.lower(), 6)


def hip_runtime_path():
    """Path of the libamdhip64 mapped into this process (diagnostics)."""
    try:
        with open("/proc/self/maps") as fh:
            for line in fh:
                if "libamdhip64" in line:
                    return line.split()[-1]
    except OSError:
        pass
    return None

"""eEMRT_GPU_: emission EM with ray-tracing operators, as
sscRaft/geometries/parallel/em_methods.py:3-70, plus eEMRT_GPU_KL (same call,
returning the per-iteration FP64 KL objective as well)."""
from ...rafttypes import CNICE, libraft, raft_check, numpy, ctypes

__all__ = ["eEMRT_GPU_", "eEMRT_GPU_KL", "tEMRT_GPU_"]


def _prepare(tomo, angles, iterations, gpus, blocksize, obj):
    nslices = 1 if len(tomo.shape) == 2 else tomo.shape[0]
    nangles = tomo.shape[-2]
    nrays = tomo.shape[-1]
    objsize = tomo.shape[-1]
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    if obj is not None:
        obj = CNICE(obj)
    else:
        obj = CNICE(numpy.ones([nslices, objsize, objsize], dtype=numpy.float32))
    tomo = CNICE(tomo)
    angles = CNICE(numpy.array(angles))
    parami = CNICE(numpy.array([nrays, nangles, nslices, objsize, 0, 0, 0, 1, iterations, blocksize]), numpy.int32)
    paramf = CNICE(numpy.array([0]), numpy.float32)
    return gpus, obj, tomo, angles, paramf, parami, nslices


def eEMRT_GPU_(tomo, angles, iterations, gpus, blocksize, obj=None):
    """Emission EM (MLEM) of data [slices, angles, rays]; `obj` (if given) is the
    initial guess and is updated in place when it is float32 C-contiguous."""
    gpus, obj, tomo, angles, paramf, parami, _ = _prepare(tomo, angles, iterations, gpus, blocksize, obj)
    libraft.get_eEM_RT_MultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                                obj.ctypes.data_as(ctypes.c_void_p), tomo.ctypes.data_as(ctypes.c_void_p),
                                angles.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                                parami.ctypes.data_as(ctypes.c_void_p))
    raft_check("get_eEM_RT_MultiGPU")
    return obj


def eEMRT_GPU_KL(tomo, angles, iterations, gpus, blocksize, obj=None):
    """As eEMRT_GPU_; also returns kl[slices, iterations] (FP64 KL(y || A x_k) before
    each update) and the count of bins left out of each sum."""
    gpus, obj, tomo, angles, paramf, parami, nslices = _prepare(tomo, angles, iterations, gpus, blocksize, obj)
    kl = numpy.zeros((nslices, max(iterations, 1)), dtype=numpy.float64)
    ex = numpy.zeros((nslices, max(iterations, 1)), dtype=numpy.int64)
    libraft.get_eEM_RT_KL_MultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                                   obj.ctypes.data_as(ctypes.c_void_p), tomo.ctypes.data_as(ctypes.c_void_p),
                                   angles.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                                   parami.ctypes.data_as(ctypes.c_void_p), kl.ctypes.data_as(ctypes.c_void_p),
                                   ex.ctypes.data_as(ctypes.c_void_p))
    raft_check("get_eEM_RT_KL_MultiGPU")
    return obj, kl[:, :iterations], ex[:, :iterations]


def tEMRT_GPU_(counts, flat, angles, iterations, gpus, blocksize, obj=None):
    """Transmission EM is not part of this build yet."""
    raise NotImplementedError("tEMRT (transmission EM) is not implemented in this build")

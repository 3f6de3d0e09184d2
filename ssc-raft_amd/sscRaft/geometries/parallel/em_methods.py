"""eEMRT_GPU_: emission EM with ray-tracing operators, as
sscRaft/geometries/parallel/em_methods.py:3-70, plus eEMRT_GPU_KL (same call,
returning the per-iteration FP64 KL objective as well), and tEMRT_GPU_
(transmission EM, em_methods.py:72-151)."""
from ...rafttypes import CNICE, libraft, raft_check, numpy, ctypes

__all__ = ["eEMRT_GPU_", "eEMRT_GPU_KL", "tEMRT_GPU_"]


def _check_sizes(obj, angles, nslices, objsize, nangles):
    """The C entry points copy nslices * objsize^2 values to and from `obj` and
    read nangles angles: refuse arrays of any other size instead of letting
    the copies run past their ends."""
    if obj.size != nslices * objsize * objsize:
        raise ValueError(f"obj must hold nslices * objsize * objsize = {nslices} * {objsize}^2 values, "
                         f"got shape {obj.shape}")
    if angles.ndim != 1 or angles.shape[0] != nangles:
        raise ValueError(f"angles must hold {nangles} values (one per projection), got shape {angles.shape}")


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
    _check_sizes(obj, angles, nslices, objsize, nangles)
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
    """Transmission EM of data [slices, angles, rays] with flat [slices, 1, rays]
    (or [slices, rays]), as em_methods.py:72-151: the C function receives
    exp(-counts) and a flat row per slice.  Unlike the reference (em_methods.py:109,
    numpy.exp(-counts, counts)), the caller's `counts` array is not overwritten.
    `obj` (if given) is the initial guess, updated in place when float32 C-contiguous."""
    nslices = 1 if len(counts.shape) == 2 else counts.shape[0]
    nflats = 1 if len(counts.shape) == 2 else numpy.asarray(flat).shape[0]
    nangles, nrays = counts.shape[-2], counts.shape[-1]
    objsize = nrays
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    count = CNICE(numpy.exp(-numpy.asarray(counts, dtype=numpy.float32)))
    flat = CNICE(numpy.asarray(flat, dtype=numpy.float32))
    if flat.size != nslices * nrays:
        raise ValueError(f"flat must hold one row of {nrays} values per slice ({nslices} slices), got {flat.shape}")
    if obj is not None:
        obj = CNICE(obj)
    else:
        obj = CNICE(numpy.ones([nslices, objsize, objsize], dtype=numpy.float32))
    angles = CNICE(numpy.array(angles))
    _check_sizes(obj, angles, nslices, objsize, nangles)
    parami = CNICE(numpy.array([nrays, nangles, nslices, objsize, 0, 0, 0, nflats, iterations, blocksize]),
                   numpy.int32)
    paramf = CNICE(numpy.array([0]), numpy.float32)
    libraft.get_tEM_RT_MultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                                obj.ctypes.data_as(ctypes.c_void_p), count.ctypes.data_as(ctypes.c_void_p),
                                flat.ctypes.data_as(ctypes.c_void_p), angles.ctypes.data_as(ctypes.c_void_p),
                                paramf.ctypes.data_as(ctypes.c_void_p), parami.ctypes.data_as(ctypes.c_void_p))
    raft_check("get_tEM_RT_MultiGPU")
    return obj

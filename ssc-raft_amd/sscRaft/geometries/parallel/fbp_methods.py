"""fbpGPU: parameter packing and call of getFBPMultiGPU, as
sscRaft/geometries/parallel/fbp_methods.py:4-103."""
from ...rafttypes import CNICE, FilterNumber, libraft, logger, raft_check, numpy, ctypes

__all__ = ["fbpGPU", "fbp_params", "bstGPU"]


def fbp_params(nrays, nangles, nslices, dic):
    """(paramf, parami) exactly as fbpGPU packs them (fbp_methods.py:48-95)."""
    filter_type = FilterNumber(dic["filter"])
    beta_delta = dic["beta/delta"]
    regularization = dic["regularization"]
    offset = dic["rotation axis offset"]
    blocksize = dic["blocksize"]
    energy = dic["energy[eV]"]
    z2 = dic["z2[m]"]
    pixelx = pixely = dic["detectorPixel[m]"]
    if beta_delta != 0.0:
        beta_delta = 1.0 / beta_delta
    else:
        beta_delta, z2, energy = 0.0, 0.0, 1.0
    padx, pady, padz = dic["padding"], 0, 0
    objsize = nrays
    parami = CNICE(numpy.array([nrays, nangles, nslices, objsize, padx, pady, padz, filter_type, 0, blocksize]),
                   numpy.int32)
    paramf = CNICE(numpy.array([beta_delta, regularization, energy, z2, pixelx, pixely, offset]), numpy.float32)
    return paramf, parami


def fbpGPU(tomogram, angles, gpus, dic, obj=None):
    """Filtered backprojection (RT) of a parallel-beam tomogram [slices, angles, rays]
    on the GPUs in `gpus`; returns obj [slices, rays, rays] (float32)."""
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    ngpus = len(gpus)
    nrays = tomogram.shape[-1]
    nangles = tomogram.shape[-2]
    nslices = 1 if len(tomogram.shape) == 2 else tomogram.shape[0]
    paramf, parami = fbp_params(nrays, nangles, nslices, dic)
    logger.info(f"Set FBP RT pad value as {dic['padding']} x horizontal dimension = ({dic['padding'] * nrays}).")
    tomogram = CNICE(tomogram)
    if obj is None:
        obj = numpy.zeros([nslices, nrays, nrays], dtype=numpy.float32)
    obj = CNICE(obj)
    angles = CNICE(numpy.array(angles))
    if angles.shape[0] != nangles:
        raise ValueError(f"angles has {angles.shape[0]} entries, tomogram has {nangles} angles")
    if obj.size != nslices * nrays * nrays:
        raise ValueError("obj must hold nslices * nrays * nrays values")
    libraft.getFBPMultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(ngpus),
                           obj.ctypes.data_as(ctypes.c_void_p), tomogram.ctypes.data_as(ctypes.c_void_p),
                           angles.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                           parami.ctypes.data_as(ctypes.c_void_p))
    raft_check("getFBPMultiGPU")
    return obj


def bstGPU(tomogram, angles, gpus, dic, obj=None, nstreams=0):
    """FBP by the Backprojection Slice Theorem, as fbp_methods.py:105-205: same
    parameter packing as fbpGPU except that the axis offset is passed as an
    integer (int(dic['rotation axis offset'])); calls getBSTMultiGPU.  The
    reconstruction assumes angles uniform on [0, pi)."""
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    nrays = tomogram.shape[-1]
    nangles = tomogram.shape[-2]
    nslices = 1 if len(tomogram.shape) == 2 else tomogram.shape[0]
    d = dict(dic)
    d["rotation axis offset"] = int(d["rotation axis offset"])
    paramf, parami = fbp_params(nrays, nangles, nslices, d)
    tomogram = CNICE(tomogram)
    if obj is None:
        obj = numpy.zeros([nslices, nrays, nrays], dtype=numpy.float32)
    obj = CNICE(obj)
    angles = CNICE(numpy.array(angles))
    if obj.size != nslices * nrays * nrays:
        raise ValueError("obj must hold nslices * nrays * nrays values")
    libraft.getBSTMultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                           obj.ctypes.data_as(ctypes.c_void_p), tomogram.ctypes.data_as(ctypes.c_void_p),
                           angles.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                           parami.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(nstreams))
    raft_check("getBSTMultiGPU")
    return obj

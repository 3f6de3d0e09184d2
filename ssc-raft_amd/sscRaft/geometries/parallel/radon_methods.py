"""radon_RT: ray-tracing Radon transform, as
sscRaft/geometries/parallel/radon_methods.py:3-61."""
from ...rafttypes import CNICE, libraft, logger, raft_check, numpy, ctypes

__all__ = ["radon_RT"]


def radon_RT(phantom, angles, gpus, pixel=1.0):
    """Radon transform of a phantom [slices, y, x] (a 2-D phantom is taken as one
    slice) at `angles` (radians); returns [slices, angles, x] scaled by
    nrays * pixel / 2 like the reference wrapper."""
    a = 1.0
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    ngpus = len(gpus)
    if len(phantom.shape) == 2:
        phantom = phantom[None]
    nslices, img_sizey, img_sizex = phantom.shape
    nrays = img_sizex
    phantom = CNICE(phantom)
    angles = CNICE(numpy.array(angles))
    nangles = angles.shape[0]
    tomogram = -numpy.ones((nslices, nangles, nrays), dtype=numpy.float32)
    libraft.getRadonRTMultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(ngpus),
                               tomogram.ctypes.data_as(ctypes.c_void_p), phantom.ctypes.data_as(ctypes.c_void_p),
                               angles.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(img_sizex),
                               ctypes.c_int(img_sizey), ctypes.c_int(nrays), ctypes.c_int(nangles),
                               ctypes.c_int(nslices), ctypes.c_float(a), ctypes.c_float(a))
    raft_check("getRadonRTMultiGPU")
    tomogram = tomogram * nrays * pixel / 2
    logger.info("Finished Radon RT method")
    return tomogram

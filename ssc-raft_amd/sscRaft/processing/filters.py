"""lowpass: the FBP filter alone, as sscRaft/processing/filters.py:5-89.

Same dictionary keys, defaults and parameter packing; calls
getFilterLowPassMultiGPU (libraft.so), which filters the tomogram in place.
Errors raise (ValueError / RuntimeError) instead of the reference's process exit.
"""
from ..io.io_ import SetDictionary
from ..rafttypes import CNICE, FilterNumber, ctypes, libraft, numpy, raft_check

__all__ = ["lowpass", "lowpass_params"]


def lowpass_params(nrays, nangles, nslices, dic):
    """(paramf, parami) exactly as lowpass packs them (filters.py:51-80)."""
    filter_type = FilterNumber(dic["filter"])
    beta_delta = dic["beta/delta"]
    regularization = dic["regularization"]
    offset = 0.0
    blocksize = dic["blocksize"]
    energy = dic["energy[eV]"]
    z2 = dic["z2[m]"]
    pixelx = pixely = dic["detectorPixel[m]"]
    if beta_delta != 0.0:
        beta_delta = 1.0 / beta_delta
    else:
        beta_delta, z2, energy = 0.0, 0.0, 1.0
    padx, pady, padz = dic["padding"], 0, 0
    parami = CNICE(numpy.array([nrays, nangles, nslices, nrays, padx, pady, padz, filter_type, 0, blocksize]),
                   numpy.int32)
    paramf = CNICE(numpy.array([beta_delta, regularization, energy, z2, pixelx, pixely, offset]), numpy.float32)
    return paramf, parami


def lowpass(tomogram, dic=None, **kwargs):
    """Low-pass (FBP) filter of a parallel-beam tomogram [slices, angles, rays]
    on the GPUs in dic['gpu']; returns the filtered tomogram (float32, filtered
    in place when the input already is C-contiguous float32)."""
    required = ("gpu",)
    optional = ("filter", "padding", "regularization", "beta/delta", "blocksize", "energy[eV]", "z2[m]",
                "detectorPixel[m]")
    default = ("ramp", 2, 0.0, 0.0, 0, 1.0, 1.0, 1.0)
    dic = SetDictionary(dic, required, optional, default)
    gpus = numpy.ascontiguousarray(numpy.array(dic["gpu"]).astype(numpy.intc))
    nrays = tomogram.shape[-1]
    nangles = tomogram.shape[-2]
    nslices = 1 if len(tomogram.shape) == 2 else tomogram.shape[0]
    paramf, parami = lowpass_params(nrays, nangles, nslices, dic)
    tomogram = CNICE(tomogram)
    libraft.getFilterLowPassMultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                                     tomogram.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                                     parami.ctypes.data_as(ctypes.c_void_p))
    raft_check("getFilterLowPassMultiGPU")
    return tomogram

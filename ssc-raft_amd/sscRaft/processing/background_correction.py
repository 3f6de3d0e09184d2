"""Flat/dark background correction, as sscRaft/processing/background_correction.py:
_background_correctionGPU, correct_projections and correct_background with the
same arguments, checks and layouts; calls getBackgroundCorrectionMultiGPU
(libraft.so).  Errors raise instead of the reference's process exit."""
from ..io.io_ import SetDictionary
from ..rafttypes import CNICE, ctypes, libraft, logger, numpy, raft_check

__all__ = ["correct_projections", "correct_background"]


def _background_correctionGPU(frames, flat, dark, gpus=[0], is_log=False, blocksize=0):
    """frames [slices, angles, rays], flat [slices, nflats, rays] (or [slices, rays]),
    dark [slices, rays] (or [slices, 1, rays]); returns the corrected frames."""
    gpus = numpy.ascontiguousarray(numpy.array(gpus).astype(numpy.intc))
    nrays = frames.shape[-1]
    nangles = frames.shape[-2]
    is_log = 1 if is_log else 0
    if len(frames.shape) == 2:
        nslices = 1
    elif len(frames.shape) == 3:
        nslices = frames.shape[0]
    else:
        raise ValueError(f"Incorrect data dimension: {frames.shape}! It accepts only 2- or 3-dimension array.")
    if len(flat.shape) == 2:
        nflats = 1
    elif len(flat.shape) == 3:
        nflats = flat.shape[-2]
    else:
        raise ValueError(f"Incorrect flat dimension: {flat.shape}! It accepts only 2- or 3-dimension array.")
    if len(dark.shape) == 3:
        dark = dark[:, 0, :]
    elif len(dark.shape) != 2:
        raise ValueError(f"Incorrect dark dimension: {dark.shape}! It accepts only 2- or 3-dimension array.")
    if flat.shape[-1] != nrays or flat.shape[0] != nslices:
        raise ValueError(f"Flat dimension ({flat.shape[0]},{flat.shape[-1]}) does not match data dimension "
                         f"({frames.shape[0]},{frames.shape[-1]}).")
    if dark.shape[-1] != nrays or dark.shape[0] != nslices:
        raise ValueError(f"Dark dimension ({dark.shape[0]},{dark.shape[-1]}) does not match data dimension "
                         f"({frames.shape[0]},{frames.shape[-1]}).")
    logger.info(f"Number of flats is {nflats}.")
    flat = CNICE(flat)
    dark = CNICE(dark)
    frames = CNICE(frames)
    libraft.getBackgroundCorrectionMultiGPU(gpus.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(gpus)),
                                            frames.ctypes.data_as(ctypes.c_void_p),
                                            flat.ctypes.data_as(ctypes.c_void_p),
                                            dark.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(nrays),
                                            ctypes.c_int(nangles), ctypes.c_int(nslices), ctypes.c_int(nflats),
                                            ctypes.c_int(is_log), ctypes.c_int(blocksize))
    raft_check("getBackgroundCorrectionMultiGPU")
    return frames


def correct_projections(frames, flat, dark, dic, **kwargs):
    """frames [angles, slices, rays], flat [nflats, slices, rays], dark [1, slices, rays]
    (detector order); returns the corrected data as [slices, angles, rays].
    dic: 'gpu' (required), 'uselog' (False), 'blocksize' (0)."""
    dic = SetDictionary(dic, ("gpu",), ("uselog", "blocksize"), (False, 0))
    frames = numpy.swapaxes(frames, 0, 1)
    flat = numpy.swapaxes(flat, 0, 1)
    dark = numpy.swapaxes(dark, 0, 1)
    return _background_correctionGPU(frames, flat, dark, dic["gpu"], dic["uselog"], dic["blocksize"])


def correct_background(frames, flat, dark, gpus=[0], is_log=False, blocksize=0):
    """frames [slices, angles, rays], flat [slices, nflats, rays], dark [slices, rays]."""
    return _background_correctionGPU(frames, flat, dark, gpus, is_log, blocksize)

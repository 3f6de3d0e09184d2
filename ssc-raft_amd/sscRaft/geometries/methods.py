"""fbp() and em(): dictionary front ends, as sscRaft/geometries/methods.py:7-186."""
from ..io import SetDictionary
from ..rafttypes import logger, numpy
from .parallel.em_methods import eEMRT_GPU_, tEMRT_GPU_
from .parallel.fbp_methods import bstGPU, fbpGPU

__all__ = ["fbp", "em"]


def fbp(tomogram, angles=None, obj=None, dic=None, nstreams=0, **kwargs):
    """Filtered backprojection of a parallel-beam tomogram [slices, angles, rays].

    dic keys (defaults): 'gpu' (required), 'angles[rad]', 'filter' ('ramp'),
    'rotation axis offset' (0.0), 'padding' (2), 'regularization' (0.0),
    'beta/delta' (0.0), 'blocksize' (0), 'energy[eV]' (1.0), 'z2[m]' (1.0),
    'method' ('RT' or 'BST'), 'detectorPixel[m]' (1.0)."""
    required = ("gpu",)
    optional = ("filter", "rotation axis offset", "padding", "regularization", "beta/delta", "blocksize",
                "energy[eV]", "z2[m]", "method", "detectorPixel[m]")
    default = ("ramp", 0.0, 2, 0.0, 0.0, 0, 1.0, 1.0, "RT", 1.0)
    dic = SetDictionary(dic, required, optional, default)
    method = dic["method"]
    gpus = dic["gpu"]
    if method == "RT":
        if "angles[rad]" in dic:
            angles = dic["angles[rad]"]
        elif angles is None:
            logger.error("Missing angles list!! Finishing run...")
            raise ValueError("Missing angles list!!")
        return fbpGPU(tomogram, angles, gpus, dic, obj=obj)
    if method == "BST":
        angles = numpy.linspace(0.0, numpy.pi, tomogram.shape[-2], endpoint=False)
        return bstGPU(tomogram, angles, gpus, dic, obj=obj, nstreams=nstreams)
    raise ValueError(f"Invalid FBP method: {method}")


def em(data, flat=None, angles=None, obj=None, dic=None, **kwargs):
    """Expectation maximisation for parallel-beam data [slices, angles, rays].

    dic keys (defaults): 'gpu' (required), 'method' (required: 'eEMRT' or 'tEMRT'),
    'flat' (tEMRT: [slices, rays], ones when absent),
    'angles[rad]', 'iterations' (10), 'detectorPixel[m]' (0.0), 'padding' (2),
    'beamgeometry' ('parallel'), 'interpolation' ('bilinear'), 'blocksize' (0)."""
    required = ("gpu",)
    optional = ("iterations", "detectorPixel[m]", "padding", "beamgeometry", "interpolation", "blocksize")
    default = (10, 0.0, 2, "parallel", "bilinear", 0)
    dic = SetDictionary(dic, required, optional, default)
    blocksize = dic["blocksize"]
    gpus = dic["gpu"]
    method = dic["method"]
    iterations = dic["iterations"]
    if angles is None:
        if "angles[rad]" not in dic:
            logger.error("Missing angles list!! Finishing run...")
            raise ValueError("Missing angles list!!")
        angles = dic["angles[rad]"]
    if method != "eEMRT" and flat is None:
        flat = dic.get("flat", None)
        if flat is None:
            logger.warning("No flat provided.")
            flat = numpy.ones((1 if len(data.shape) == 2 else data.shape[0], data.shape[-1]))
    if method == "eEMRT":
        return eEMRT_GPU_(data, angles, iterations, gpus, blocksize, obj=obj)
    if method == "tEMRT":
        return tEMRT_GPU_(data, flat, angles, iterations, gpus, blocksize, obj=obj)
    if method == "tEMFQ":
        raise NotImplementedError("tEMFQ is not implemented in this build")
    logger.error(f"Invalid EM method:{method}")
    raise ValueError(f"Invalid EM method:{method}")

"""sscRaft-compatible Python layer over libraft.so (MI355X / gfx950).

Same entry points, dictionary keys, defaults and array layouts as the
reference package for the parallel-beam hot path:
``fbp``, ``em`` (method 'eEMRT'), ``fbpGPU``, ``eEMRT_GPU_``, ``radon_RT``,
``lowpass``, plus ``SetDictionary``, ``CNICE`` and ``FilterNumber``.

Unlike the reference (sscRaft/__init__.py:12-13, which only logs when the
shared library is missing), importing this package raises if libraft.so
cannot be loaded: there is no CPU fallback.
"""
from ._version import __version__
from .rafttypes import *
from .io import *
from .geometries import *
from .processing import *

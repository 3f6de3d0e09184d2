"""CPU checks of the C ABI and of the host logic (no GPU compute).

* libraft.so loads and exports every function include/*.h declares;
* without a GPU the reference entry points fail with an error code and a
  message (the reference exit()s), and the library stays usable;
* the planning query (raft_fbp_plan_info) rejects what the reference cannot
  run and returns a staged bin window that covers every bin the
  backprojector can touch (property test against the oracle's FP64 formula);
* the Python layer packs parameters exactly as sscRaft's wrappers do
  (fbp_methods.py:48-95, em_methods.py:52-63, rafttypes.py:485-519).
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import sscRaft

    names = declared_functions()
    assert {"getFBPMultiGPU", "get_eEM_RT_MultiGPU", "getRadonRTMultiGPU", "raft_last_error",
            "raft_nn_conv3x3"} <= set(names)
    lib = ctypes.CDLL(sscRaft.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def _has_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_error_path():
    import sscRaft

    y = np.zeros((1, 8, 16), np.float32)
    ang = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    with pytest.raises(RuntimeError, match="no HIP device"):
        sscRaft.fbp(y, dic={"gpu": [0], "angles[rad]": ang})
    with pytest.raises(RuntimeError):
        sscRaft.radon_RT(np.zeros((1, 16, 16), np.float32), ang, [0])
    with pytest.raises(RuntimeError):
        sscRaft.eEMRT_GPU_(np.ones((1, 8, 16), np.float32), ang, 2, [0], 0)
    with pytest.raises(RuntimeError):
        sscRaft.lowpass(np.ones((1, 8, 16), np.float32), dic={"gpu": [0]})
    with pytest.raises(RuntimeError):
        sscRaft.correct_background(np.ones((1, 8, 16), np.float32), np.ones((1, 1, 16), np.float32),
                                   np.zeros((1, 16), np.float32))
    with pytest.raises(ValueError, match="Flat dimension"):
        sscRaft.correct_background(np.ones((2, 8, 16), np.float32), np.ones((1, 1, 16), np.float32),
                                   np.zeros((2, 16), np.float32))
    with pytest.raises(RuntimeError):
        sscRaft.tEMRT_GPU_(np.ones((1, 8, 16), np.float32), np.ones((1, 1, 16), np.float32), ang, 2, [0], 0)
    sscRaft.libraft.raft_clear_error()
    assert sscRaft.libraft.raft_last_error() == 0


def test_lowpass_param_packing_and_defaults_match_reference():
    """processing/filters.py:38-80: defaults, offset forced to 0, beta/delta inversion."""
    from sscRaft.io.io_ import SetDictionary
    from sscRaft.processing.filters import lowpass_params

    dic = SetDictionary({"gpu": [0]}, ("gpu",),
                        ("filter", "padding", "regularization", "beta/delta", "blocksize", "energy[eV]", "z2[m]",
                         "detectorPixel[m]"), ("ramp", 2, 0.0, 0.0, 0, 1.0, 1.0, 1.0))
    pf, pi = lowpass_params(100, 40, 2, dic)
    np.testing.assert_array_equal(pi, np.array([100, 40, 2, 100, 2, 0, 0, 7, 0, 0], np.int32))
    np.testing.assert_array_equal(pf, np.array([0.0, 0.0, 1.0, 0.0, 1.0, 1.0, 0.0], np.float32))
    dic.update({"beta/delta": 50.0, "filter": "gaussian", "regularization": 0.3, "energy[eV]": 8.0, "z2[m]": 2.0})
    pf, pi = lowpass_params(100, 40, 2, dic)
    np.testing.assert_array_equal(pf, np.array([1 / 50.0, 0.3, 8.0, 2.0, 1.0, 1.0, 0.0], np.float32))
    assert pi[7] == 1


def plan(angles, paramf, parami):
    from sscRaft import device as dv

    return dv.fbp_plan_info(angles, paramf, parami)


def test_plan_rejects_invalid_parameters():
    ang = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    paramf, parami = oracle.fbp_params(16, 8, 1, pad=2)
    for idx, val in [(0, 1), (1, 1), (4, -1), (7, 8), (7, -1)]:
        pi = parami.copy()
        pi[idx] = val
        with pytest.raises(ValueError):
            plan(ang, paramf, pi)
    pi = parami.copy()
    pi[3] = 17  # objsize != nrays: outside what sscRaft.fbpGPU passes
    with pytest.raises(RuntimeError, match="objsize"):
        plan(ang, paramf, pi)


def test_plan_fft_path_selection():
    ang = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    for n, pad, fft in [(2048, 2, 1), (1024, 2, 1), (100, 2, 1), (17, 2, 0), (7000, 2, 0), (33, 0, 0), (45, 0, 1)]:
        paramf, parami = oracle.fbp_params(n, 8, 1, pad=pad)
        assert plan(ang, paramf, parami)[2] == fft, (n, pad)


@settings(max_examples=40, deadline=None)
@given(n=st.integers(2, 300), a=st.integers(2, 64), pad=st.integers(0, 3), seed=st.integers(0, 10 ** 6),
       spread=st.sampled_from([np.pi, 2 * np.pi, 7.0]))
def test_plan_window_covers_every_bin(n, a, pad, seed, spread):
    """Every bin T = (int)u of every pixel of the backprojector's tile grid (64 x 32
    pixel tiles, so ceil(n/64)*64 columns by ceil(n/32)*32 rows) lies in the
    staged window."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(-1.0, spread, a)).astype(np.float32)
    paramf, parami = oracle.fbp_params(n, a, 1, pad=pad)
    tlo, width, _ = plan(ang, paramf, parami)
    npad, L = n * (1 + pad), (pad * n) // 2
    half = npad / 2.0
    I = np.arange(-(-n // 64) * 64) + L - half
    J = np.arange(-(-n // 32) * 32) + L - half
    th = ang.astype(np.float64)
    for c, s in zip(np.cos(th), np.sin(th)):
        u = I[None, :] * c - J[:, None] * s + half
        T = np.trunc(u)
        assert T.min() >= tlo and T.max() < tlo + width


def test_fbp_param_packing_matches_reference():
    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    dic = {"filter": "hann", "beta/delta": 0.0, "regularization": 0.5, "rotation axis offset": 1.5,
           "blocksize": 4, "energy[eV]": 9.0, "z2[m]": 3.0, "detectorPixel[m]": 0.01, "padding": 1}
    pf, pi = fbp_params(64, 32, 3, dic)
    np.testing.assert_array_equal(pi, np.array([64, 32, 3, 64, 1, 0, 0, 5, 0, 4], np.int32))
    # beta/delta == 0: z2 -> 0, energy -> 1 (fbp_methods.py:57-62)
    np.testing.assert_array_equal(pf, np.array([0.0, 0.5, 1.0, 0.0, 0.01, 0.01, 1.5], np.float32))
    dic["beta/delta"] = 200.0
    pf, _ = fbp_params(64, 32, 3, dic)
    np.testing.assert_array_equal(pf, np.array([1 / 200.0, 0.5, 9.0, 3.0, 0.01, 0.01, 1.5], np.float32))
    assert pi.dtype == np.int32 and pf.dtype == np.float32


def test_filter_numbers_and_cnice():
    import sscRaft

    names = ["none", "gaussian", "lorentz", "cosine", "rectangle", "hann", "hamming", "ramp"]
    assert [sscRaft.FilterNumber(n) for n in names] == list(range(8))
    assert sscRaft.FilterNumber("RAMP") == 7
    assert sscRaft.FilterNumber("shepp-logan") == 6  # unknown -> hamming, as the reference
    a = np.arange(12, dtype=np.float64).reshape(3, 4)
    b = sscRaft.CNICE(a)
    assert b.dtype == np.float32 and b.flags["C_CONTIGUOUS"]
    c = np.zeros((3, 4), np.float32)
    assert sscRaft.CNICE(c) is c
    d = sscRaft.CNICE(c.T)
    assert d.flags["C_CONTIGUOUS"] and d is not c


def test_set_dictionary():
    from sscRaft import SetDictionary

    with pytest.raises(ValueError, match="gpu"):
        SetDictionary({}, ("gpu",), ("a",), (1,))
    d = SetDictionary({"gpu": [0], "a": 5}, ("gpu",), ("a", "b"), (1, "x"))
    assert d == {"gpu": [0], "a": 5, "b": "x"}


def test_fbp_front_end_validation():
    import sscRaft

    y = np.zeros((1, 8, 16), np.float32)
    with pytest.raises(ValueError, match="angles"):
        sscRaft.fbp(y, dic={"gpu": [0]})
    with pytest.raises(ValueError, match="gpu"):
        sscRaft.fbp(y, angles=np.zeros(8), dic={})
    with pytest.raises(RuntimeError, match="no HIP device"):
        sscRaft.fbp(y, angles=np.zeros(8), dic={"gpu": [0], "method": "BST"})
    with pytest.raises(ValueError, match="Invalid FBP method"):
        sscRaft.fbp(y, angles=np.zeros(8), dic={"gpu": [0], "method": "XX"})
    with pytest.raises(KeyError):
        sscRaft.em(y, angles=np.zeros(8), dic={"gpu": [0]})  # 'method' is required by the reference too
    with pytest.raises(NotImplementedError):
        sscRaft.em(y, angles=np.zeros(8), dic={"gpu": [0], "method": "tEMFQ"})
    with pytest.raises(ValueError, match="Invalid EM method"):
        sscRaft.em(y, angles=np.zeros(8), dic={"gpu": [0], "method": "xEM"})

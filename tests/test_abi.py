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
from hypothesis import example, given, settings
from hypothesis import strategies as st

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)  # preprocessor lines (macros are not exports)
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


def test_em_and_bst_front_ends_refuse_wrong_sizes():
    """eEMRT_GPU_ / eEMRT_GPU_KL / tEMRT_GPU_ / bstGPU check `obj` and `angles`
    against the data shape before the C call: the C side copies nslices *
    objsize^2 values to and from `obj` and reads nangles angles, so a wrong
    size would run past the ends of the caller's arrays."""
    import sscRaft
    from sscRaft.geometries.parallel.fbp_methods import bstGPU

    y = np.ones((2, 8, 16), np.float32)
    ang = np.linspace(0, np.pi, 8, endpoint=False).astype(np.float32)
    small = np.ones((2, 15, 15), np.float32)
    for fn in (sscRaft.eEMRT_GPU_, sscRaft.eEMRT_GPU_KL):
        with pytest.raises(ValueError, match="obj"):
            fn(y, ang, 2, [0], 0, obj=small)
        with pytest.raises(ValueError, match="obj"):
            fn(y, ang, 2, [0], 0, obj=np.ones((1, 16, 16), np.float32))
        with pytest.raises(ValueError, match="angles"):
            fn(y, ang[:7], 2, [0], 0)
        with pytest.raises(ValueError, match="angles"):
            fn(y, np.zeros((2, 8), np.float32), 2, [0], 0)
    flat = np.ones((2, 1, 16), np.float32)
    with pytest.raises(ValueError, match="obj"):
        sscRaft.tEMRT_GPU_(y, flat, ang, 2, [0], 0, obj=small)
    with pytest.raises(ValueError, match="angles"):
        sscRaft.tEMRT_GPU_(y, flat, ang[:5], 2, [0], 0)
    dic = {"filter": "ramp", "beta/delta": 0.0, "regularization": 1.0, "rotation axis offset": 0.0,
           "blocksize": 0, "energy[eV]": 1.0, "z2[m]": 0.0, "detectorPixel[m]": 1.0, "padding": 2}
    with pytest.raises(ValueError, match="obj"):
        bstGPU(y, ang, [0], dic, obj=small)


REFERENCE_RAFTTYPES = "/root/reference/sscRaft/rafttypes.py"


def _rafttypes_symbols():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "rafttypes_symbols.json")) as fh:
        return json.load(fh)


def test_rafttypes_symbol_fixture_matches_the_reference_loader():
    """The committed list equals a fresh parse of the reference's rafttypes.py
    (where the reference is present, as in the build container)."""
    if not os.path.exists(REFERENCE_RAFTTYPES):
        pytest.skip("reference sources not present on this machine")
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "list_rafttypes_symbols", os.path.join(ROOT, "tests", "golden", "list_rafttypes_symbols.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(REFERENCE_RAFTTYPES) as fh:
        assert mod.parse(fh.read()) == _rafttypes_symbols()


class _LabFDK(ctypes.Structure):  # struct RaftStubLabFDK (include/raft_stubs.h)
    _fields_ = ([(n, ctypes.c_float) for n in ("x", "y", "z", "dx", "dy", "dz")]
                + [(n, ctypes.c_int) for n in ("nx", "ny", "nz")]
                + [(n, ctypes.c_float) for n in ("h", "v", "dh", "dv")]
                + [(n, ctypes.c_int) for n in ("nh", "nv")]
                + [(n, ctypes.c_float) for n in ("D", "Dsd", "beta_max", "dbeta")]
                + [(n, ctypes.c_int) for n in ("nbeta", "fourier", "filter_type")]
                + [("reg", ctypes.c_float)]
                + [(n, ctypes.c_int) for n in ("is_slice", "s0", "s1", "s2", "s3", "nph", "padh")]
                + [("energy", ctypes.c_float)]
                + [(n, ctypes.c_int) for n in ("rotation_axis_offset", "blocksize")])


class _LabRay(ctypes.Structure):  # struct RaftStubLabEM / RaftStubLabCB (include/raft_stubs.h)
    _fields_ = ([(n, ctypes.c_float) for n in ("a", "b", "c", "x0", "y0", "z0")]
                + [(n, ctypes.c_int) for n in ("nx", "ny", "nz")]
                + [(n, ctypes.c_float) for n in ("sx", "sy", "sz")]
                + [(n, ctypes.c_int) for n in ("nbeta", "ndet", "npts")])


def test_every_reference_binding_resolves_and_stubs_report_not_implemented():
    """Every function the reference's loader binds is exported by libraft.so,
    and binding it as the loader does works (argtypes / restype set from the
    fixture).  The ones outside the hot path record RAFT_EUNSUPPORTED and
    touch nothing; no GPU is needed to call them."""
    import sscRaft

    lib = ctypes.CDLL(sscRaft.LIB_PATH)
    structs = {"Lab": _LabFDK, "Lab_EM": _LabRay, "Lab_CB": _LabRay}
    built = {"getFBPMultiGPU", "get_eEM_RT_MultiGPU", "getRadonRTMultiGPU", "get_tEM_RT_MultiGPU",
             "getFilterLowPassMultiGPU", "getBSTMultiGPU", "getBackgroundCorrectionMultiGPU"}
    syms = _rafttypes_symbols()
    assert len(syms) == 23 and built <= set(syms)
    lib.raft_last_error.restype = ctypes.c_int
    lib.raft_last_error_string.restype = ctypes.c_char_p
    for name, sig in syms.items():
        fn = getattr(lib, name)  # AttributeError if not exported
        types = [structs.get(t) or getattr(ctypes, t) for t in sig["argtypes"]]
        fn.argtypes = types
        fn.restype = None if sig["restype"] == "None" else getattr(ctypes, sig["restype"])
        if name in built:
            continue
        buf = np.full(8, 3.0, np.float32)
        args = [t() if t in structs.values() else (buf.ctypes.data_as(ctypes.c_void_p) if t is ctypes.c_void_p
                                                   else t(1)) for t in types]
        lib.raft_clear_error()
        ret = fn(*args)
        assert lib.raft_last_error() == 4, name  # RAFT_EUNSUPPORTED
        assert b"not implemented: " + name.encode() in lib.raft_last_error_string()
        assert np.all(buf == 3.0)
        if sig["restype"] == "c_int":
            assert ret == -1
        elif sig["restype"] == "c_float":
            assert np.isnan(ret)
    lib.raft_clear_error()


def _proj_geom(Nx, Ny, nrays, ax, ay, em):
    """Sampling constants of the forward projector (ray_geom_radon / ray_geom_em)."""
    if em:
        c = (nrays - 1) / 2.0
        return dict(ci=c, ck=c, ucx=c, ucy=c, ry=1.0, rx=1.0, nk=nrays)
    return dict(ci=(Nx - 1) / 2.0, ck=(Ny - 1) / 2.0, ucx=(Nx - 1) / (2.0 * ax), ucy=(Ny - 1) / (2.0 * ay),
                ry=(ay * (Nx - 1)) / (ax * (Ny - 1)), rx=(ax * (Ny - 1)) / (ay * (Nx - 1)), nk=Ny)


@settings(max_examples=30, deadline=None)
@example(em=True, nx=64, ny=64, nrays=64, a=64, kind="uniform", seed=0, ax=1.0, ay=1.0)
@example(em=False, nx=96, ny=80, nrays=100, a=100, kind="uniform", seed=0, ax=1.0, ay=1.0)
@example(em=True, nx=120, ny=120, nrays=120, a=160, kind="uniform", seed=3, ax=1.0, ay=1.0)
@example(em=False, nx=8, ny=8, nrays=8, a=8, kind="uniform", seed=0, ax=1.0, ay=1.0)
@given(em=st.booleans(), nx=st.integers(8, 120), ny=st.integers(8, 120), nrays=st.integers(8, 130),
       a=st.integers(8, 160), kind=st.sampled_from(["uniform", "random", "cluster"]), seed=st.integers(0, 10 ** 6),
       ax=st.sampled_from([1.0, 0.7, 1.6]), ay=st.sampled_from([1.0, 1.3, 0.8]))
def test_projector_plan_keeps_every_sample_in_the_staged_box(em, nx, ny, nrays, a, kind, seed, ax, ay):
    """The staged projector gathers from LDS with no clamp: it relies on every
    sample's cell lying in its wave's patch box, and on the union of the 8
    boxes of a workgroup fitting the tile whenever raft_proj_plan_info admits
    the staged form.  Brute force over every chunk of 8 angles, every 32-ray
    block and a spread of step offsets (the kernel's steps start anywhere in
    [0, nk)): the patch box (FP64 extremes over the patch corners, margins -1
    and +3) contains each of the 32 x 32 samples, and the union of all 8 boxes
    (a superset of the kernel's union, which skips waves with no samples in
    the step) fits PL_CAP entries at the kernel's narrowest pitch (1 mod 4
    entries; the kernel widens it to a multiple of 16 only where that fits)."""
    from sscRaft import device as dv

    rng = np.random.default_rng(seed)
    if em:
        ny, ax, ay = nx, 1.0, 1.0
        nrays = min(nrays, nx + nx // 16)  # what the eEM entry point admits (backprojection window)
    if kind == "uniform":
        ang = (np.arange(a) * (np.pi / a)).astype(np.float32)
    elif kind == "random":
        ang = np.sort(rng.uniform(-np.pi, 2 * np.pi, a)).astype(np.float32)
    else:  # dense clusters straddling the axes and the diagonal
        ang = np.sort(rng.choice([0.0, np.pi / 4, np.pi / 2, np.pi], a) + rng.normal(0, 0.05, a)).astype(np.float32)
    rc, (PR, PK, PA, CAP) = dv.proj_plan_info(ang, nx, ny, nrays, nslices=4, ax=ax, ay=ay, em=em)
    if (em, nx, a, kind) in ((True, 64, 64, "uniform"), (False, 96, 100, "uniform"), (True, 120, 160, "uniform")):
        assert rc > 0  # dense sets take the staged form
    if rc == 0:
        return
    g = _proj_geom(nx, ny, nrays, ax, ay, em)
    nk = g["nk"]
    th = ang.astype(np.float64)
    cs_all, sn_all = np.cos(th), np.sin(th)
    r = np.arange(PR, dtype=np.float64)
    m = np.arange(PK, dtype=np.float64)
    ks_all = np.unique(np.concatenate([np.arange(0, nk, 7), [0, nk - 1], np.arange(0, nk, PK)]))
    for a0 in range(0, a, PA):
        am = min(a0 + PA // 2, a - 1)
        steep = abs(sn_all[am]) > abs(cs_all[am])
        cs, sn = cs_all[a0:a0 + PA], sn_all[a0:a0 + PA]
        bxk, byi = -sn * g["ry"], sn * g["rx"]
        rI, rK = (cs, bxk) if steep else (byi, cs)
        cI, cK = (byi, cs) if steep else (cs, bxk)
        rU, cU = (g["ucx"], g["ucy"]) if steep else (g["ucy"], g["ucx"])
        W = len(cs)
        for ib in range(0, nrays, PR):
            ir0 = ib - g["ci"]
            kr0 = ks_all - g["ck"]  # [K]
            ir1, kr1 = ir0 + (PR - 1), kr0 + (PK - 1)

            def ext(U, I, K, f):
                return U + f(ir0 * I, ir1 * I)[:, None] + f(kr0[None, :] * K[:, None], kr1[None, :] * K[:, None])

            lo_r = np.floor(ext(rU, rI, rK, np.minimum)) - 1  # [W, K]
            hi_r = np.floor(ext(rU, rI, rK, np.maximum)) + 3
            lo_c = np.floor(ext(cU, cI, cK, np.minimum)) - 1
            hi_c = np.floor(ext(cU, cI, cK, np.maximum)) + 3
            # samples: ray ib + r, sample ks + m (m < nk - ks)
            ir = ir0 + r  # [R]
            kr = kr0[:, None] + m[None, :]  # [K, M]
            row = np.trunc(kr[None, :, None, :] * rK[:, None, None, None]
                           + (ir[None, None, :, None] * rI[:, None, None, None] + rU))  # [W, K, R, M]
            col = np.trunc(kr[None, :, None, :] * cK[:, None, None, None]
                           + (ir[None, None, :, None] * cI[:, None, None, None] + cU))
            valid = (ks_all[:, None] + m[None, :] < nk)[None, :, None, :]
            inside = ((row >= lo_r[:, :, None, None]) & (row < hi_r[:, :, None, None])
                      & (col >= lo_c[:, :, None, None]) & (col < hi_c[:, :, None, None]))
            assert np.all(inside | ~valid), (a0, ib)
            rows = hi_r.max(axis=0) - lo_r.min(axis=0)
            w = (hi_c.max(axis=0) - lo_c.min(axis=0)).astype(np.int64)
            pitch = w + ((1 - w) & 3)  # the narrowest pitch the kernel stages at
            assert np.all(rows * pitch <= CAP), (a0, ib, int((rows * pitch).max()), CAP)
    assert W >= 1


def test_autotune_cache_file_round_trip(tmp_path, monkeypatch):
    """RAFT_AUTOTUNE_CACHE: choices written by an earlier process are read back
    (no GPU needed: reading and dumping the cache is host work)."""
    from sscRaft import device as dv

    f = tmp_path / "tune.json"
    f.write_text('{\n "bp/n=64/A=48/W=99/store=0/mode=0/groups=1": {"choice": 16, "ms": [[32, 0.5], [16, 0.4]]},\n'
                 ' "proj/Nx=64/Ny=64/nrays=64/A=64/em=1/groups=1/ry=1.000000": {"choice": 4, "ms": []}\n}\n')
    monkeypatch.setenv("RAFT_AUTOTUNE_CACHE", str(f))
    dv.autotune_clear()
    c = dv.autotune_choices()
    assert c["bp/n=64/A=48/W=99/store=0/mode=0/groups=1"]["choice"] == 16
    assert c["proj/Nx=64/Ny=64/nrays=64/A=64/em=1/groups=1/ry=1.000000"]["choice"] == 4
    monkeypatch.delenv("RAFT_AUTOTUNE_CACHE")
    dv.autotune_clear()
    assert dv.autotune_choices() == {}

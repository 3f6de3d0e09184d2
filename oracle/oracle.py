"""ctypes front end of the CPU oracle (oracle/raft_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.

Every function takes the same arguments as the reference entry point it
restates (see raft_oracle.c's header for file:line citations) and returns
numpy arrays.  ``prec='f64'`` selects the FP64 oracle, ``prec='f32'`` the FP32
restatement that is timed as the CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def default_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for prec, ctype in (("f64", np.float64), ("f32", np.float32)):
            rp = np.ctypeslib.ndpointer(dtype=ctype, flags="C_CONTIGUOUS")
            f = getattr(L, "oracle_fbp_" + prec)
            f.argtypes = [rp, _f32p, _f32p, _f32p, _i32p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_fbp_padded_" + prec)
            f.argtypes = [rp, _f32p, _f32p, _f32p, _i32p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_fbp_filter_" + prec)
            f.argtypes = [rp, _f32p, _f32p, _i32p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_radon_" + prec)
            f.argtypes = [rp, _f32p, _f32p] + [ctypes.c_int] * 5 + [ctypes.c_float, ctypes.c_float, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_kl_" + prec)
            f.argtypes = [rp, rp, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
            f.restype = ctypes.c_double
            f = getattr(L, "oracle_eem_" + prec)
            f.argtypes = [rp, _f32p, _f32p, _i32p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_em_backproject_" + prec)
            f.argtypes = [rp, rp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(L, "oracle_em_forward_" + prec)
            f.argtypes = [rp, rp, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _dt(prec):
    return np.float64 if prec == "f64" else np.float32


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with code {rc}")


def fbp_params(nrays, nangles, nslices, pad=2, filter_type=7, pixel=1.0, offset=0.0, reg=0.0,
               beta_delta=0.0, energy=1.0, z2=1.0):
    """parami/paramf exactly as fbpGPU packs them (fbp_methods.py:57-95)."""
    if beta_delta != 0.0:
        beta_delta = 1.0 / beta_delta
    else:
        beta_delta, z2, energy = 0.0, 0.0, 1.0
    parami = np.array([nrays, nangles, nslices, nrays, pad, 0, 0, filter_type, 0, 0], dtype=np.int32)
    paramf = np.array([beta_delta, reg, energy, z2, pixel, pixel, offset], dtype=np.float32)
    return paramf, parami


def fbp(tomo, angles, paramf, parami, prec="f64", padded=False, nthreads=None):
    tomo = np.ascontiguousarray(tomo, dtype=np.float32)
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    n, S = int(parami[0]), int(parami[2])
    out = np.zeros((S, n, n), dtype=_dt(prec))
    fn = getattr(lib(), ("oracle_fbp_padded_" if padded else "oracle_fbp_") + prec)
    _check(fn(out, tomo, angles, np.ascontiguousarray(paramf, np.float32),
              np.ascontiguousarray(parami, np.int32), nthreads or default_threads()), "fbp")
    return out


def fbp_filter(tomo, paramf, parami, prec="f64", nthreads=None):
    """Padded + filtered rows q[s][a][0..npad) (filter and 1/(npad*pixel) applied)."""
    tomo = np.ascontiguousarray(tomo, dtype=np.float32)
    n, A, S, P = int(parami[0]), int(parami[1]), int(parami[2]), int(parami[4])
    out = np.zeros((S, A, n * (1 + P)), dtype=_dt(prec))
    fn = getattr(lib(), "oracle_fbp_filter_" + prec)
    _check(fn(out, tomo, np.ascontiguousarray(paramf, np.float32), np.ascontiguousarray(parami, np.int32),
              nthreads or default_threads()), "fbp_filter")
    return out


def lowpass(tomo, paramf, parami, prec="f64", nthreads=None):
    """getFilterLowPassMultiGPU restated (low_pass.cu:9-30, filters.cu:120-161):
    the FBP filter stage with the axis offset forced to 0, cropped back to the
    N original bins [L, L + N); filter 'none' returns the data unchanged."""
    tomo = np.ascontiguousarray(tomo, dtype=np.float32)
    if tomo.ndim == 2:
        tomo = tomo[None]
    pf = np.array(paramf, dtype=np.float32).copy()
    pf[6] = 0.0
    if int(parami[7]) == 0:
        return tomo.astype(_dt(prec))
    n, P = int(parami[0]), int(parami[4])
    L = (P * n) // 2
    return np.ascontiguousarray(fbp_filter(tomo, pf, parami, prec, nthreads)[:, :, L:L + n])


def radon(phantom, angles, nrays=None, ax=1.0, ay=1.0, prec="f64", nthreads=None):
    """getRadonRTMultiGPU on a [S, Ny, Nx] phantom; no Python-side rescale."""
    phantom = np.ascontiguousarray(phantom, dtype=np.float32)
    if phantom.ndim == 2:
        phantom = phantom[None]
    S, Ny, Nx = phantom.shape
    nrays = Nx if nrays is None else nrays
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((S, angles.shape[0], nrays), dtype=_dt(prec))
    fn = getattr(lib(), "oracle_radon_" + prec)
    _check(fn(out, phantom, angles, Nx, Ny, nrays, angles.shape[0], S, ax, ay, nthreads or default_threads()),
           "radon")
    return out


def kl(y, v, prec="f64"):
    y = np.ascontiguousarray(y, dtype=_dt(prec)).ravel()
    v = np.ascontiguousarray(v, dtype=_dt(prec)).ravel()
    ex = ctypes.c_long(0)
    val = getattr(lib(), "oracle_kl_" + prec)(y, v, y.size, ctypes.byref(ex))
    return val, ex.value


def eem(tomo, angles, iterations, obj=None, objsize=None, prec="f64", with_kl=True, nthreads=None):
    """get_eEM_RT_MultiGPU restated; returns (obj, kl[S, K] or None, excluded[S, K] or None)."""
    tomo = np.ascontiguousarray(tomo, dtype=np.float32)
    if tomo.ndim == 2:
        tomo = tomo[None]
    S, A, R = tomo.shape
    n = R if objsize is None else objsize
    x = np.ones((S, n, n), dtype=_dt(prec)) if obj is None else np.array(obj, dtype=_dt(prec), copy=True)
    x = np.ascontiguousarray(x)
    parami = np.array([R, A, S, n, 0, 0, 0, 1, iterations, 0], dtype=np.int32)
    klv = np.zeros((S, max(iterations, 1)), dtype=np.float64) if with_kl else None
    exv = np.zeros((S, max(iterations, 1)), dtype=np.int64) if with_kl else None
    fn = getattr(lib(), "oracle_eem_" + prec)
    _check(fn(x, tomo, np.ascontiguousarray(angles, dtype=np.float32), parami,
              klv.ctypes.data if with_kl else None, exv.ctypes.data if with_kl else None,
              nthreads or default_threads()), "eem")
    if with_kl:
        return x, klv[:, :iterations], exv[:, :iterations]
    return x, None, None


def em_backproject(q, angles, n, ones=False, prec="f64"):
    q = np.ascontiguousarray(q, dtype=_dt(prec))
    A, R = q.shape
    out = np.zeros((n, n), dtype=_dt(prec))
    _check(getattr(lib(), "oracle_em_backproject_" + prec)(out, q, np.ascontiguousarray(angles, np.float32), R, A,
                                                           n, 1 if ones else 0), "em_backproject")
    return out


def background(frames, flat, dark, is_log=False):
    """getBackgroundCorrectionMultiGPU restated (background_correction.cu:13-45,
    opt.cu:204-214) in float64 from the float32 inputs: frames [S, A, R],
    flat [S, nflats, R], dark [S, R]."""
    D = np.asarray(frames, np.float32).astype(np.float64)
    S, A, R = D.shape
    F = np.asarray(flat, np.float32).astype(np.float64).reshape(S, -1, R)
    dk = np.asarray(dark, np.float32).astype(np.float64).reshape(S, 1, R)
    if F.shape[1] > 1:
        interp = (np.arange(A, dtype=np.float32) / np.float32(A)).astype(np.float64)[None, :, None]
        ft = F[:, :1, :] * (1.0 - interp) + interp * F[:, 1:2, :]
    else:
        ft = F[:, :1, :]
    v = np.maximum(D - dk, 0.5) / np.maximum(ft - dk, 0.5)
    return -np.log(v) if is_log else v


def tem(count, flat, angles, iterations, obj=None, prec="f64"):
    """get_tEM_RT_MultiGPU restated (em_rt.cu:155-192, em_RT_kernel.cu:20-31,
    52-65,92-184) from the pinned eEM operators:
        backcounts = B(count);  K x { x <- x * B(flat[s] * exp(-A x)) / backcounts }
    count [S, A, R] (already exp(-data)), flat [S, R]: slice s uses its own
    flat row (the reference indexes the flat buffer by angle, em_RT_kernel.cu:29).
    obj: initial guess [S, n, n] (ones by default, n = R)."""
    count = np.asarray(count, dtype=np.float32)
    if count.ndim == 2:
        count = count[None]
    S, A, R = count.shape
    flat = np.asarray(flat, dtype=np.float32).reshape(S, R)
    if obj is None:
        x = np.ones((S, R, R))
    else:
        x = np.array(obj, dtype=np.float64, copy=True)
        x = x.reshape(S, x.shape[-2], x.shape[-1])  # n x n, n = objsize (may differ from R)
    n = x.shape[-1]
    for s in range(S):
        bc = em_backproject(count[s], angles, n, prec=prec)
        for _ in range(iterations):
            v = em_forward(x[s], angles, R, prec=prec)
            t = flat[s][None, :].astype(np.float64) * np.exp(-v)
            x[s] = x[s] * em_backproject(t, angles, n, prec=prec) / bc
    return x


def em_forward(x, angles, nrays, prec="f64"):
    x = np.ascontiguousarray(x, dtype=_dt(prec))
    n = x.shape[0]
    A = len(angles)
    out = np.zeros((A, nrays), dtype=_dt(prec))
    _check(getattr(lib(), "oracle_em_forward_" + prec)(out, x, np.ascontiguousarray(angles, np.float32), nrays, A,
                                                       n), "em_forward")
    return out

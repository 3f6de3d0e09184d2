"""Device-resident entry points for torch tensors already in HBM
(raft_fbp_device / raft_radon_device / raft_eem_device).  torch is only the
allocator and stream provider here; every computation runs in libraft."""
import numpy

from .rafttypes import CNICE, libraft, raft_check, ctypes

__all__ = ["fbp_device", "lds_read_probe", "radon_device", "eem_device", "eem_kl_device", "batch_device", "lowpass_device",
           "bst_device", "tem_device", "background_device", "set_profiling", "last_kernel_ms", "fbp_plan_info", "proj_plan_info", "autotune_choices",
           "autotune_clear", "eem_device_sched",
           "KF_FILTER", "KF_BP", "KF_PROJ", "KF_EMBP"]

KF_FILTER, KF_BP, KF_PROJ, KF_EMBP = 0, 1, 2, 3


def _stream(t, stream):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(t.device)
    return ctypes.c_void_p(stream.cuda_stream)


def _check_tensor(t, name):
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")


def fbp_device(tomo, angles, paramf, parami, out=None, store="f32", stream=None, noalloc=False):
    """FBP of tomo [S, A, N] (CUDA float32) -> out [S, N, N].
    noalloc: run raft_fbp_device_noalloc (no device allocation by the
    library: its intermediates live in a fixed per-device scratch array;
    out must then be given, since allocating it here would defeat the point)."""
    import torch

    _check_tensor(tomo, "tomo")
    N, A, S = int(parami[0]), int(parami[1]), int(parami[2])
    if out is None:
        if noalloc:
            raise ValueError("noalloc=True needs a preallocated out tensor")
        out = torch.empty((S, N, N), dtype=torch.float32, device=tomo.device)
    _check_tensor(out, "out")
    angles = CNICE(numpy.asarray(angles))
    paramf = CNICE(numpy.asarray(paramf), numpy.float32)
    parami = CNICE(numpy.asarray(parami), numpy.int32)
    if out.numel() < S * N * N or tomo.numel() < S * A * N or angles.shape[0] < A:
        raise ValueError("tomo, out or angles smaller than parami says")
    fn = libraft.raft_fbp_device_noalloc if noalloc else libraft.raft_fbp_device
    rc = fn(tomo.device.index or 0, ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(tomo.data_ptr()), angles.ctypes.data_as(ctypes.c_void_p),
                                 paramf.ctypes.data_as(ctypes.c_void_p), parami.ctypes.data_as(ctypes.c_void_p),
                                 1 if store == "f16" else 0, _stream(tomo, stream))
    if rc:
        raft_check("raft_fbp_device_noalloc" if noalloc else "raft_fbp_device")
    return out


def lds_read_probe(device=0):
    """LDS read rate of the device (raft_lds_read_probe): conflict-free
    16-byte reads by every CU.  Returns dict(bytes_per_s, cus, clock_khz,
    bytes_per_clk_per_cu)."""
    r = ctypes.c_double()
    cus = ctypes.c_int()
    clk = ctypes.c_int()
    if libraft.raft_lds_read_probe(int(device), ctypes.byref(r), ctypes.byref(cus), ctypes.byref(clk)):
        raft_check("raft_lds_read_probe")
    per = r.value / max(cus.value, 1) / (max(clk.value, 1) * 1e3)
    return dict(bytes_per_s=r.value, cus=cus.value, clock_khz=clk.value, bytes_per_clk_per_cu=per)


def radon_device(img, angles, nrays=None, ax=1.0, ay=1.0, out=None, stream=None):
    """Radon (no Python-side rescale) of img [S, Ny, Nx] -> out [S, A, nrays]."""
    import torch

    _check_tensor(img, "img")
    S, Ny, Nx = img.shape
    nrays = Nx if nrays is None else nrays
    angles = CNICE(numpy.asarray(angles))
    A = angles.shape[0]
    if out is None:
        out = torch.empty((S, A, nrays), dtype=torch.float32, device=img.device)
    _check_tensor(out, "out")
    rc = libraft.raft_radon_device(img.device.index or 0, ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(img.data_ptr()), angles.ctypes.data_as(ctypes.c_void_p), Nx, Ny,
                                   nrays, A, S, ctypes.c_float(ax), ctypes.c_float(ay), _stream(img, stream))
    if rc:
        raft_check("raft_radon_device")
    return out


def eem_device(tomo, angles, iterations, obj=None, with_kl=True, stream=None, kl_iterations=None):
    """eEM of tomo [S, A, R]; obj [S, R, R] is the in/out initial guess (ones if None).
    Returns (obj, kl [S, K] float64 or None, excluded [S, K] or None).
    kl_iterations: record the KL of the first kl_iterations iterations only
    (kl [S, min(kl_iterations, K)], raft_eem_device_kl); the others run
    without the FP64 KL terms."""
    import torch

    _check_tensor(tomo, "tomo")
    S, A, R = tomo.shape
    if obj is None:
        obj = torch.ones((S, R, R), dtype=torch.float32, device=tomo.device)
    _check_tensor(obj, "obj")
    angles = CNICE(numpy.asarray(angles))
    parami = CNICE(numpy.array([R, A, S, obj.shape[-1], 0, 0, 0, 1, iterations, 0]), numpy.int32)
    K = iterations if kl_iterations is None else min(kl_iterations, iterations)
    with_kl = with_kl and K > 0
    kl = numpy.zeros((S, max(K, 1)), dtype=numpy.float64) if with_kl else None
    ex = numpy.zeros((S, max(K, 1)), dtype=numpy.int64) if with_kl else None
    args = (ctypes.c_void_p(obj.data_ptr()), ctypes.c_void_p(tomo.data_ptr()), angles.ctypes.data_as(ctypes.c_void_p),
            parami.ctypes.data_as(ctypes.c_void_p))
    outs = (kl.ctypes.data_as(ctypes.c_void_p) if with_kl else None,
            ex.ctypes.data_as(ctypes.c_void_p) if with_kl else None, _stream(tomo, stream))
    if kl_iterations is None:
        rc = libraft.raft_eem_device(tomo.device.index or 0, *args, *outs)
    else:
        rc = libraft.raft_eem_device_kl(tomo.device.index or 0, *args, int(kl_iterations), *outs)
    if rc:
        raft_check("raft_eem_device")
    if with_kl:
        return obj, kl[:, :K], ex[:, :K]
    return obj, None, None


def batch_device(tomos, angles, paramf, pad=2, filter_type=7, store="f16", fbp_out=None, ys=None, em=None,
                 iterations=0, stream=None):
    """Batched uneven scans in one graph launch (raft_batch_device).

    tomos: list of CUDA float32 [A_i, N_i] sinograms (FBP input); angles: list of
    float32 arrays; paramf: [n, 7] FBP float parameters per scan; fbp_out: list of
    [N_i, N_i] outputs (None: skip FBP); ys / em: eEM data and in/out iterates
    (None: skip eEM).  Returns the device time (ms) of the graph launch."""
    n = len(tomos) if tomos is not None else len(ys)
    nr = numpy.array([(tomos[i] if tomos is not None else ys[i]).shape[-1] for i in range(n)], numpy.int32)
    na = numpy.array([(tomos[i] if tomos is not None else ys[i]).shape[-2] for i in range(n)], numpy.int32)
    angs = [CNICE(numpy.asarray(a)) for a in angles]
    pa = (ctypes.c_void_p * n)(*[a.ctypes.data for a in angs])

    def ptrs(lst):
        if lst is None:
            return None
        for t in lst:
            _check_tensor(t, "scan buffer")
        return (ctypes.c_void_p * n)(*[t.data_ptr() for t in lst])

    pf = CNICE(numpy.asarray(paramf, numpy.float32).reshape(n, 7))
    ms = ctypes.c_float(0.0)
    ref = tomos[0] if tomos is not None else ys[0]
    rc = libraft.raft_batch_device(ref.device.index or 0, n, nr.ctypes.data_as(ctypes.c_void_p),
                                   na.ctypes.data_as(ctypes.c_void_p), pa, ptrs(tomos), ptrs(ys), ptrs(fbp_out),
                                   ptrs(em), pf.ctypes.data_as(ctypes.c_void_p), pad, filter_type,
                                   1 if store == "f16" else 0, iterations, _stream(ref, stream), ctypes.byref(ms))
    if rc:
        raft_check("raft_batch_device")
    return ms.value


def lowpass_device(tomo, paramf, parami, stream=None):
    """Filter-only export on device data, in place on tomo [S, A, N]; returns tomo."""
    _check_tensor(tomo, "tomo")
    paramf = CNICE(numpy.asarray(paramf), numpy.float32)
    parami = CNICE(numpy.asarray(parami), numpy.int32)
    rc = libraft.raft_lowpass_device(tomo.device.index or 0, ctypes.c_void_p(tomo.data_ptr()),
                                     paramf.ctypes.data_as(ctypes.c_void_p), parami.ctypes.data_as(ctypes.c_void_p),
                                     _stream(tomo, stream))
    if rc:
        raft_check("raft_lowpass_device")
    return tomo


def bst_device(tomo, paramf, parami, out=None, stream=None):
    """FBP by the Backprojection Slice Theorem (raft_bst_device) on tomo [S, A, N]
    (angles uniform on [0, pi)); returns out [S, N, N]."""
    import torch

    _check_tensor(tomo, "tomo")
    S, A, N = tomo.shape
    if out is None:
        out = torch.empty((S, N, N), dtype=torch.float32, device=tomo.device)
    _check_tensor(out, "out")
    paramf = CNICE(numpy.asarray(paramf), numpy.float32)
    parami = CNICE(numpy.asarray(parami), numpy.int32)
    rc = libraft.raft_bst_device(tomo.device.index or 0, ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(tomo.data_ptr()), paramf.ctypes.data_as(ctypes.c_void_p),
                                 parami.ctypes.data_as(ctypes.c_void_p), _stream(tomo, stream))
    if rc:
        raft_check("raft_bst_device")
    return out


def tem_device(count, flat, angles, iterations, obj=None, stream=None):
    """Transmission EM (raft_tem_device): count [S, A, R] (= exp(-data)), flat
    [S, R], obj [S, n, n] in/out (ones when None); returns obj."""
    import torch

    _check_tensor(count, "count")
    _check_tensor(flat, "flat")
    S, A, R = count.shape
    if obj is None:
        obj = torch.ones((S, R, R), dtype=torch.float32, device=count.device)
    _check_tensor(obj, "obj")
    angles = CNICE(numpy.asarray(angles))
    parami = CNICE(numpy.array([R, A, S, obj.shape[-1], 0, 0, 0, 1, iterations, 0]), numpy.int32)
    rc = libraft.raft_tem_device(count.device.index or 0, ctypes.c_void_p(obj.data_ptr()),
                                 ctypes.c_void_p(count.data_ptr()), ctypes.c_void_p(flat.data_ptr()),
                                 angles.ctypes.data_as(ctypes.c_void_p), parami.ctypes.data_as(ctypes.c_void_p),
                                 _stream(count, stream))
    if rc:
        raft_check("raft_tem_device")
    return obj


def background_device(frames, flat, dark, is_log=False, stream=None):
    """Flat/dark correction in place on frames [S, A, R] (raft_background_device),
    flat [S, F, R], dark [S, R]; returns frames."""
    for t, name in ((frames, "frames"), (flat, "flat"), (dark, "dark")):
        _check_tensor(t, name)
    S, A, R = frames.shape
    rc = libraft.raft_background_device(frames.device.index or 0, ctypes.c_void_p(frames.data_ptr()),
                                        ctypes.c_void_p(flat.data_ptr()), ctypes.c_void_p(dark.data_ptr()), R, A, S,
                                        flat.shape[1], int(bool(is_log)), _stream(frames, stream))
    if rc:
        raft_check("raft_background_device")
    return frames


def eem_kl_device(tomo, angles, obj, stream=None):
    """KL(y || A x) per slice for images obj [S, R, R] (not modified); returns (kl [S], excluded [S])."""
    _check_tensor(tomo, "tomo")
    _check_tensor(obj, "obj")
    S, A, R = tomo.shape
    angles = CNICE(numpy.asarray(angles))
    parami = CNICE(numpy.array([R, A, S, obj.shape[-1], 0, 0, 0, 1, 1, 0]), numpy.int32)
    kl = numpy.zeros(S, dtype=numpy.float64)
    ex = numpy.zeros(S, dtype=numpy.int64)
    rc = libraft.raft_eem_kl_device(tomo.device.index or 0, ctypes.c_void_p(obj.data_ptr()),
                                    ctypes.c_void_p(tomo.data_ptr()), angles.ctypes.data_as(ctypes.c_void_p),
                                    parami.ctypes.data_as(ctypes.c_void_p), kl.ctypes.data_as(ctypes.c_void_p),
                                    ex.ctypes.data_as(ctypes.c_void_p), _stream(tomo, stream))
    if rc:
        raft_check("raft_eem_kl_device")
    return kl, ex


def set_profiling(on):
    libraft.raft_set_profiling(1 if on else 0)


def last_kernel_ms(which):
    return libraft.raft_last_kernel_ms(which)


def fbp_plan_info(angles, paramf, parami):
    """(tlo, width, fft_path) of the FBP plan for these parameters (host only)."""
    angles = CNICE(numpy.asarray(angles))
    paramf = CNICE(numpy.asarray(paramf), numpy.float32)
    parami = CNICE(numpy.asarray(parami), numpy.int32)
    tlo, width, fft = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    rc = libraft.raft_fbp_plan_info(angles.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
                                    parami.ctypes.data_as(ctypes.c_void_p), ctypes.byref(tlo), ctypes.byref(width),
                                    ctypes.byref(fft))
    if rc:
        raft_check("raft_fbp_plan_info")
    return tlo.value, width.value, fft.value


def proj_plan_info(angles, Nx, Ny=None, nrays=None, nslices=4, ax=1.0, ay=1.0, em=False):
    """(ray_chunks, (rays per patch, samples per step, angles per workgroup,
    staged entries)) of the forward projector's plan (host only): ray_chunks
    > 0 selects the staged (LDS) form, 0 the direct form.  em=True plans the
    eEM projector of an Nx x Nx image from nrays-bin rows."""
    angles = CNICE(numpy.asarray(angles))
    Ny = Nx if Ny is None else Ny
    nrays = Nx if nrays is None else nrays
    rc = ctypes.c_int(0)
    consts = numpy.zeros(4, numpy.int32)
    err = libraft.raft_proj_plan_info(angles.ctypes.data_as(ctypes.c_void_p), angles.shape[0], Nx, Ny, nrays,
                                      nslices, ax, ay, 1 if em else 0, ctypes.byref(rc),
                                      consts.ctypes.data_as(ctypes.c_void_p))
    if err:
        raft_check("raft_proj_plan_info")
    return rc.value, tuple(int(c) for c in consts)


def autotune_choices():
    """The autotuner's choices held by the library: {key: {"choice": c, "ms":
    [[candidate, ms], ...]}} (raft_autotune_dump)."""
    import json

    n = libraft.raft_autotune_dump(None, 0)
    buf = ctypes.create_string_buffer(n)
    libraft.raft_autotune_dump(buf, n)
    return json.loads(buf.value.decode())


def autotune_clear():
    """Forget the choices held in memory (raft_autotune_clear)."""
    libraft.raft_autotune_clear()


def eem_device_sched(tomo, angles, iterations, obj=None, stream=None):
    """eEM with an iteration count per slice (raft_eem_device_sched):
    iterations [S] must be non-increasing; slice s stops after iterations[s].
    obj [S, n, n] is the in/out initial guess (ones if None).  Returns obj."""
    import torch

    _check_tensor(tomo, "tomo")
    S, A, R = tomo.shape
    if obj is None:
        obj = torch.ones((S, R, R), dtype=torch.float32, device=tomo.device)
    _check_tensor(obj, "obj")
    it = CNICE(numpy.asarray(iterations), numpy.int32)
    if it.shape != (S,):
        raise ValueError(f"iterations must hold one count per slice ({S}), got shape {it.shape}")
    angles = CNICE(numpy.asarray(angles))
    parami = CNICE(numpy.array([R, A, S, obj.shape[-1], 0, 0, 0, 1, int(it.max()) if S else 0, 0]), numpy.int32)
    rc = libraft.raft_eem_device_sched(tomo.device.index or 0, ctypes.c_void_p(obj.data_ptr()),
                                       ctypes.c_void_p(tomo.data_ptr()), angles.ctypes.data_as(ctypes.c_void_p),
                                       parami.ctypes.data_as(ctypes.c_void_p), it.ctypes.data_as(ctypes.c_void_p),
                                       _stream(tomo, stream))
    if rc:
        raft_check("raft_eem_device_sched")
    return obj

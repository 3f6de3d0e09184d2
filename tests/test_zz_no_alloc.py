"""raft_fbp_device_noalloc: FBP whose intermediates live in the library's fixed
per-device scratch array (no device allocation by the library).

Checked through the C ABI against raft_fbp_device on the same device inputs,
bit for bit (same kernels, slice groups independent), on two calls in a row,
the second on a smaller input, for FP32 and FP16 rows; the smaller call is
also checked against the FP64 CPU oracle.  The first call is the 2048^2,
2048-angle geometry at 20 slices: its filtered rows (95 MB per slice group)
do not all fit the scratch, so the call runs in several blocks.
"""
import ctypes

import numpy as np
import pytest

import oracle
from sscRaft import phantom as ph

pytestmark = pytest.mark.gpu

RAFT_EUNSUPPORTED = 4


def _call(fn, out, tomo, ang, paramf, parami, store):
    import torch

    ang = np.ascontiguousarray(ang, dtype=np.float32)
    paramf = np.ascontiguousarray(paramf, dtype=np.float32)
    parami = np.ascontiguousarray(parami, dtype=np.int32)
    stream = torch.cuda.current_stream(tomo.device).cuda_stream
    return fn(0, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(tomo.data_ptr()),
              ang.ctypes.data_as(ctypes.c_void_p), paramf.ctypes.data_as(ctypes.c_void_p),
              parami.ctypes.data_as(ctypes.c_void_p), store, ctypes.c_void_p(stream))


def _sino(n, a, s, seed):
    ang = ph.uniform_angles(a)
    rng = np.random.default_rng(seed)
    # a smooth sinogram-shaped input with noise: exact values do not matter for
    # the bit-for-bit comparison, only that every bin is exercised
    t = np.linspace(-1.0, 1.0, n, dtype=np.float64)
    base = np.maximum(0.0, 1.0 - t * t)[None, None, :] * (1.0 + 0.3 * np.sin(np.arange(a))[None, :, None])
    y = (base * (1.0 + np.arange(s))[:, None, None] + 0.05 * rng.standard_normal((s, a, n))).astype(np.float32)
    return y, ang


@pytest.mark.parametrize("store", [0, 1])
def test_noalloc_matches_fbp_device_on_two_calls(gpu, store):
    import torch

    from sscRaft.rafttypes import libraft

    cases = [
        # (n, a, s, pad, filter, offset): the 2048^2 geometry at 20 slices (several scratch blocks),
        # then a smaller input with another filter and an axis offset
        (2048, 2048, 20, 2, 7, 0.0),
        (200, 150, 5, 1, 5, 1.5),
    ]
    for i, (n, a, s, pad, filt, off) in enumerate(cases):
        print(f"noalloc case {i}: n={n} a={a} s={s} pad={pad} filter={filt} offset={off} store={store}", flush=True)
        y, ang = _sino(n, a, s, seed=40 + i)
        paramf = np.array([0.0, 0.0, 0.0, 0.0, 2.0 / n, 2.0 / n, off], np.float32)
        parami = np.array([n, a, s, n, pad, 0, 0, filt, 0, 0], np.int32)
        dy = torch.from_numpy(y).cuda()
        ref = torch.full((s, n, n), float("nan"), dtype=torch.float32, device="cuda")
        out = torch.full((s, n, n), float("nan"), dtype=torch.float32, device="cuda")
        assert _call(libraft.raft_fbp_device, ref, dy, ang, paramf, parami, store) == 0
        assert _call(libraft.raft_fbp_device_noalloc, out, dy, ang, paramf, parami, store) == 0, \
            libraft.raft_last_error_string()
        torch.cuda.synchronize()
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert np.isfinite(o).all()
        assert np.array_equal(o, r), f"case {i}: max |diff| {np.abs(o - r).max():.3e}"
        if n <= 256:
            want = oracle.fbp(y, ang, paramf, parami)
            e = float(np.linalg.norm(o - want) / np.linalg.norm(want))
            print(f"noalloc case {i}: rel-L2 vs FP64 oracle {e:.3e}", flush=True)
            assert e < (1e-5 if store == 0 else 5e-4), e


def test_noalloc_user_blocksize_and_repeat(gpu):
    """A caller-set blocksize below the scratch capacity, and the same call
    repeated (the scratch keeps the previous call's contents)."""
    import torch

    from sscRaft.rafttypes import libraft

    n, a, s = 96, 80, 9
    y, ang = _sino(n, a, s, seed=3)
    paramf = np.array([0.0, 0.0, 0.0, 0.0, 2.0 / n, 2.0 / n, 0.0], np.float32)
    parami = np.array([n, a, s, n, 2, 0, 0, 7, 0, 4], np.int32)
    dy = torch.from_numpy(y).cuda()
    ref = torch.empty((s, n, n), dtype=torch.float32, device="cuda")
    assert _call(libraft.raft_fbp_device, ref, dy, ang, paramf, parami, 0) == 0
    for _ in range(2):
        out = torch.full((s, n, n), float("nan"), dtype=torch.float32, device="cuda")
        assert _call(libraft.raft_fbp_device_noalloc, out, dy, ang, paramf, parami, 0) == 0
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref.cpu().numpy())


def test_noalloc_refuses_geometry_beyond_scratch(gpu):
    """4096 angles x 4096 rays at pad 2: one slice group of rows is 363 MB,
    more than the scratch holds -> RAFT_EUNSUPPORTED before any launch."""
    import torch

    from sscRaft.rafttypes import libraft

    n = a = 4096
    ang = ph.uniform_angles(a)
    paramf = np.array([0.0, 0.0, 0.0, 0.0, 2.0 / n, 2.0 / n, 0.0], np.float32)
    parami = np.array([n, a, 1, n, 2, 0, 0, 7, 0, 0], np.int32)
    dy = torch.zeros((1, 4), dtype=torch.float32, device="cuda")  # never read: the call is refused first
    out = torch.zeros((1, 4), dtype=torch.float32, device="cuda")
    rc = _call(libraft.raft_fbp_device_noalloc, out, dy, ang, paramf, parami, 0)
    assert rc == RAFT_EUNSUPPORTED
    assert b"scratch" in libraft.raft_last_error_string()
    libraft.raft_clear_error()

"""GPU parity: libraft (through the reference C ABI) against the FP64 CPU oracle.

Bars: FP32 paths within 1e-5 relative L2 of the FP64 oracle (bin indices are
computed with the oracle's own FP64 expressions, so only value rounding
differs); FP16-stored filtered rows within 2e-3 (reported); eEM iterates and
per-iteration KL tracked against the oracle.
"""
import numpy as np
import pytest

import oracle
from sscRaft import phantom as ph

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def noisy_sino(n, a, s, seed, i0=1e4):
    img = ph.phantoms(n, s, config=9, first=seed)
    ang = ph.uniform_angles(a)
    p = oracle.radon(img, ang) * 1.0  # pixel = 2/N: nrays*pixel/2 = 1
    return ph.poisson_log(p, i0, seed=seed), ang


def fbp_dic(ang, pad=2, filt="ramp", pixel=1.0, offset=0.0, reg=0.0, bd=0.0):
    return {"gpu": [0], "angles[rad]": ang, "filter": filt, "padding": pad, "detectorPixel[m]": pixel,
            "rotation axis offset": offset, "regularization": reg, "beta/delta": bd, "energy[eV]": 2.0e4,
            "z2[m]": 0.5}


FILTERS = ["none", "gaussian", "lorentz", "cosine", "rectangle", "hann", "hamming", "ramp"]


@pytest.mark.parametrize("n,a,pad", [(16, 12, 0), (33, 17, 1), (64, 64, 2), (100, 90, 2), (128, 77, 3),
                                     (17, 20, 2), (255, 128, 1), (96, 2, 2)])
def test_fbp_sizes(gpu, n, a, pad):
    import sscRaft

    y, ang = noisy_sino(n, a, 2, seed=n + a)
    dic = fbp_dic(ang, pad=pad, pixel=2.0 / n)
    out = sscRaft.fbp(y, dic=dic)
    paramf, parami = oracle.fbp_params(n, a, 2, pad=pad, filter_type=7, pixel=2.0 / n)
    ref = oracle.fbp(y, ang, paramf, parami)
    e = rel(out, ref)
    print(f"fbp n={n} a={a} pad={pad} rel-L2 {e:.2e}")
    assert e <= FP32_TOL


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("offset", [0.0, 1.5, -0.7])
def test_fbp_filters(gpu, filt, offset):
    import sscRaft

    n, a = 64, 48
    y, ang = noisy_sino(n, a, 1, seed=3)
    dic = fbp_dic(ang, pad=2, filt=filt, pixel=0.37, offset=offset, reg=0.8, bd=150.0)
    out = sscRaft.fbp(y, dic=dic)
    paramf, parami = oracle.fbp_params(n, a, 1, pad=2, filter_type=sscRaft.FilterNumber(filt), pixel=0.37,
                                       offset=offset, reg=0.8, beta_delta=150.0, energy=2.0e4, z2=0.5)
    ref = oracle.fbp(y, ang, paramf, parami)
    e = rel(out, ref)
    print(f"fbp filter={filt} offset={offset} rel-L2 {e:.2e}")
    assert e <= FP32_TOL


def test_fbp_angles_over_pi_and_nonuniform(gpu):
    import sscRaft

    n, a = 48, 40
    rng = np.random.default_rng(5)
    ang = np.sort(rng.uniform(0, 2 * np.pi, a)).astype(np.float32)  # last > pi -> angular scale
    y = rng.standard_normal((2, a, n)).astype(np.float32)
    out = sscRaft.fbp(y, dic=fbp_dic(ang, pad=1, pixel=1.0))
    paramf, parami = oracle.fbp_params(n, a, 2, pad=1, filter_type=7, pixel=1.0)
    assert rel(out, oracle.fbp(y, ang, paramf, parami)) <= FP32_TOL


def test_fbp_2d_input_and_given_obj(gpu):
    import sscRaft

    n, a = 40, 30
    y, ang = noisy_sino(n, a, 1, seed=11)
    obj = np.full((1, n, n), 7.0, dtype=np.float32)
    out = sscRaft.fbp(y[0], obj=obj, dic=fbp_dic(ang, pixel=2.0 / n))
    assert out is obj
    paramf, parami = oracle.fbp_params(n, a, 1, pad=2, filter_type=7, pixel=2.0 / n)
    assert rel(out, oracle.fbp(y, ang, paramf, parami)) <= FP32_TOL


def test_fbp_2048(gpu):
    """The headline size: one 2048^2 slice, 2048 angles, pad 2, ramp, I0 = 1e4."""
    import sscRaft

    n = a = 2048
    img = ph.phantoms(n, 1, config=2)
    ang = ph.uniform_angles(a)
    y = ph.poisson_log(oracle.radon(img, ang), 1e4, seed=2000)
    out = sscRaft.fbp(y, dic=fbp_dic(ang, pixel=2.0 / n))
    paramf, parami = oracle.fbp_params(n, a, 1, pad=2, filter_type=7, pixel=2.0 / n)
    ref = oracle.fbp(y, ang, paramf, parami)
    e = rel(out, ref)
    print(f"fbp 2048^2 rel-L2 vs FP64 oracle {e:.3e}, max abs {np.abs(out - ref).max():.3e}")
    assert e <= FP32_TOL


@pytest.mark.parametrize("n,a,s,filt", [(1024, 48, 8, "ramp"), (768, 40, 6, "hann"), (2048, 16, 4, "ramp"),
                                        (1024, 20, 5, "none")])
def test_filter_two_workgroup_form_is_bit_identical(gpu, monkeypatch, n, a, s, filt):
    """The FP32 filter's default form for padded lengths in (2048, 6144] keeps
    the first slice pair's window in LDS so that two workgroups fit a CU; it
    runs each pair's FFTs with the same arithmetic as the form that holds the
    window in registers (RAFT_FFT_FORM=regs), so FBP agrees bit for bit
    (padded lengths 3072, 2304 and 6144; 5 slices leave a partial group)."""
    import sscRaft

    y, ang = noisy_sino(n, a, s, seed=n + a)
    dic = fbp_dic(ang, pad=2, filt=filt, pixel=2.0 / n)
    monkeypatch.setenv("RAFT_FFT_FORM", "regs")
    o0 = sscRaft.fbp(y, dic=dic)
    monkeypatch.delenv("RAFT_FFT_FORM")
    o1 = sscRaft.fbp(y, dic=dic)
    np.testing.assert_array_equal(o0, o1)


def test_fbp_fp16_storage(gpu):
    import torch
    from sscRaft import device as dv

    n, a, s = 256, 256, 2
    y, ang = noisy_sino(n, a, s, seed=21)
    paramf, parami = oracle.fbp_params(n, a, s, pad=2, filter_type=7, pixel=2.0 / n)
    ref = oracle.fbp(y, ang, paramf, parami)
    t = torch.from_numpy(y).cuda()
    o32 = dv.fbp_device(t, ang, paramf, parami, store="f32")
    o16 = dv.fbp_device(t, ang, paramf, parami, store="f16")
    torch.cuda.synchronize()
    e32 = rel(o32.cpu().numpy(), ref)
    e16 = rel(o16.cpu().numpy(), ref)
    print(f"fbp 256^2 device path: f32 rel-L2 {e32:.2e}, f16-stored rows rel-L2 {e16:.2e}")
    assert e32 <= FP32_TOL
    assert e16 <= 2e-3


@pytest.mark.parametrize("ny,nx,a", [(16, 16, 12), (33, 33, 17), (64, 64, 90), (20, 32, 19), (100, 100, 64),
                                     (256, 256, 256)])
def test_radon(gpu, ny, nx, a):
    import sscRaft

    rng = np.random.default_rng(ny * nx + a)
    img = rng.random((2, ny, nx)).astype(np.float32)
    ang = ph.uniform_angles(a)
    out = sscRaft.radon_RT(img, ang, [0], pixel=1.0)
    ref = oracle.radon(img, ang) * (nx * 1.0 / 2)
    e = rel(out, ref)
    print(f"radon {ny}x{nx} a={a} rel-L2 {e:.2e}")
    assert e <= FP32_TOL


def test_radon_random_angles(gpu):
    import sscRaft

    rng = np.random.default_rng(1)
    img = ph.phantoms(128, 2, config=4)
    ang = np.sort(rng.uniform(-np.pi, 2 * np.pi, 50)).astype(np.float32)
    out = sscRaft.radon_RT(img, ang, [0], pixel=2.0 / 128)
    ref = oracle.radon(img, ang)
    assert rel(out, ref) <= FP32_TOL


@pytest.mark.parametrize("form", ["default", "direct"])
@pytest.mark.parametrize("ny,nx,a", [(64, 64, 64), (80, 96, 100), (128, 128, 256), (200, 200, 200), (31, 31, 40)])
def test_radon_projector_forms(gpu, monkeypatch, form, ny, nx, a):
    """Dense angle sets (all the cases here) take the staged (LDS) projector by
    default, and RAFT_PROJ_LDS=0 forces the direct one: both against the FP64
    oracle.  Sparse sets (test_radon's (100, 100, 64)) take the direct one."""
    import sscRaft

    if form == "direct":
        monkeypatch.setenv("RAFT_PROJ_LDS", "0")
    rng = np.random.default_rng(ny + nx + a)
    img = rng.random((5, ny, nx)).astype(np.float32)
    ang = ph.uniform_angles(a)
    out = sscRaft.radon_RT(img, ang, [0], pixel=2.0 / nx)
    ref = oracle.radon(img, ang)
    e = rel(out, ref)
    print(f"radon {form} {ny}x{nx} a={a}: rel-L2 {e:.2e}")
    assert e <= FP32_TOL


@pytest.mark.parametrize("ny,nx,a,s", [(64, 64, 64, 5), (200, 200, 200, 4), (31, 31, 40, 3), (256, 256, 240, 4),
                                       (1024, 1024, 960, 4)])
def test_staged_projector_prefetch_is_bit_identical(gpu, monkeypatch, ny, nx, a, s):
    """The staged projector's prefetch form (the default: the next step's
    first rows loaded into registers during the current step's gathers)
    stages the same values as the plain form (RAFT_PROJ_PREFETCH=0), so Radon
    sums, eEM iterates, KL and excluded counts, and tEM iterates agree bit for
    bit.
    Most union boxes here are taller than the rows prefetched per step (56 in
    the Radon form, 40 in the eEM form), so the plain staging of the rest
    runs too."""
    import sscRaft

    rng = np.random.default_rng(ny * a + s)
    img = rng.random((s, ny, nx)).astype(np.float32)
    ang = ph.uniform_angles(a)
    monkeypatch.setenv("RAFT_PROJ_PREFETCH", "0")
    r0 = sscRaft.radon_RT(img, ang, [0], pixel=2.0 / nx)
    monkeypatch.delenv("RAFT_PROJ_PREFETCH")
    r1 = sscRaft.radon_RT(img, ang, [0], pixel=2.0 / nx)
    np.testing.assert_array_equal(r0, r1)
    if ny != nx or ny > 256:
        return
    y = np.maximum(ph.poisson_log(r0 / ny, 1e4, seed=a), 0)
    monkeypatch.setenv("RAFT_PROJ_PREFETCH", "0")
    e0 = sscRaft.eEMRT_GPU_KL(y, ang, 4, [0], 0)
    t0 = sscRaft.tEMRT_GPU_(y * 0.01, np.ones((s, 1, nx), np.float32), ang, 3, [0], 0)
    monkeypatch.delenv("RAFT_PROJ_PREFETCH")
    e1 = sscRaft.eEMRT_GPU_KL(y, ang, 4, [0], 0)
    t1 = sscRaft.tEMRT_GPU_(y * 0.01, np.ones((s, 1, nx), np.float32), ang, 3, [0], 0)
    for u, v in zip(e0, e1):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(t0, t1)


@pytest.mark.parametrize("n,a,k", [(64, 64, 5), (40, 48, 3)])
def test_eem_projector_forms_agree(gpu, monkeypatch, n, a, k):
    """eEM with the staged projector (default for these dense sets) and with
    the direct one (forced) agree to FP32 rounding, and both track the oracle."""
    import sscRaft

    img = ph.phantoms(n, 3, config=3, first=n + 7)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n + 7), 0)
    x1, kl1, ex1 = sscRaft.eEMRT_GPU_KL(y, ang, k, [0], 0)
    monkeypatch.setenv("RAFT_PROJ_LDS", "0")
    x2, kl2, ex2 = sscRaft.eEMRT_GPU_KL(y, ang, k, [0], 0)
    xr, klr, exr = oracle.eem(y, ang, k)
    print(f"eem forms n={n} a={a}: staged {rel(x1, xr):.2e}, direct {rel(x2, xr):.2e}, between {rel(x1, x2):.2e}")
    assert rel(x1, xr) <= FP32_TOL and rel(x2, xr) <= FP32_TOL
    assert rel(x1, x2) <= FP32_TOL
    assert float(np.max(np.abs(kl1 - klr) / np.abs(klr))) <= 1e-4
    np.testing.assert_array_equal(ex1, exr)
    np.testing.assert_array_equal(ex2, exr)


@pytest.mark.parametrize("nrays,n", [(48, 64), (64, 60), (33, 40)])
@pytest.mark.parametrize("form", ["staged", "direct"])
def test_eem_objsize_differs_from_nrays(gpu, monkeypatch, nrays, n, form):
    """eEM of an n x n image from nrays-bin rows (objsize != nrays), both projector
    forms, against the oracle; the image side comes from the initial guess."""
    import torch
    from sscRaft import device as dv

    a, k = 40, 4
    if form == "direct":
        monkeypatch.setenv("RAFT_PROJ_LDS", "0")
    img = ph.phantoms(nrays, 3, config=3, first=n)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n), 0).astype(np.float32)
    x0 = np.random.default_rng(n).uniform(0.5, 1.5, (3, n, n)).astype(np.float32)
    xt = torch.from_numpy(x0).cuda()
    x, kl, ex = dv.eem_device(torch.from_numpy(y).cuda(), ang, k, obj=xt)
    x = x.cpu().numpy()
    xr, klr, exr = oracle.eem(y, ang, k, obj=x0.astype(np.float64), objsize=n)
    e = rel(x, xr)
    ek = float(np.max(np.abs(kl - klr) / np.abs(klr)))
    print(f"eem objsize {n} nrays {nrays} {form}: x rel-L2 {e:.2e}, KL max rel {ek:.2e}")
    assert e <= FP32_TOL and ek <= 1e-5
    np.testing.assert_array_equal(ex, exr)


def test_eem_objsize_outside_tile_window_is_refused(gpu):
    """nrays much larger than objsize spreads a pixel tile over more bins than the
    backprojection window holds: refused with an error, nothing written."""
    import torch
    from sscRaft import device as dv

    y = torch.ones((1, 8, 64), dtype=torch.float32, device="cuda")
    x = torch.ones((1, 40, 40), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="objsize"):
        dv.eem_device(y, ph.uniform_angles(8), 2, obj=x)
    assert bool(torch.all(x == 1))


@pytest.mark.parametrize("n,a,k", [(16, 16, 1), (33, 17, 5), (64, 64, 20), (47, 60, 5)])
def test_eem(gpu, n, a, k):
    import sscRaft

    img = ph.phantoms(n, 2, config=3, first=n)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n), 0)
    x, kl, ex = sscRaft.eEMRT_GPU_KL(y, ang, k, [0], 0)
    xr, klr, exr = oracle.eem(y, ang, k)
    e = rel(x, xr)
    ek = float(np.max(np.abs(kl - klr) / np.abs(klr)))
    print(f"eem n={n} a={a} K={k}: x rel-L2 {e:.2e}, KL max rel {ek:.2e}, KL {kl[0, 0]:.4e}->{kl[0, -1]:.4e}")
    assert e <= (FP32_TOL if k <= 5 else 1e-4)
    assert ek <= 1e-4
    np.testing.assert_array_equal(ex, exr)


def test_eem_plain_abi_matches_kl_variant(gpu):
    import sscRaft

    n, a = 40, 36
    img = ph.phantoms(n, 1, config=3)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=1), 0)
    x1 = sscRaft.em(y, angles=ang, dic={"gpu": [0], "method": "eEMRT", "iterations": 4})
    x2, _, _ = sscRaft.eEMRT_GPU_KL(y, ang, 4, [0], 0)
    np.testing.assert_array_equal(x1, x2)


def test_eem_initial_guess_in_place(gpu):
    import sscRaft

    n, a = 32, 30
    img = ph.phantoms(n, 1, config=3, first=5)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=2), 0)
    x0 = np.full((1, n, n), 0.5, dtype=np.float32)
    out = sscRaft.eEMRT_GPU_(y, ang, 3, [0], 0, obj=x0)
    assert out is x0
    xr, _, _ = oracle.eem(y, ang, 3, obj=np.full((1, n, n), 0.5))
    assert rel(out, xr) <= FP32_TOL


def test_errors(gpu):
    import sscRaft

    y = np.zeros((1, 1, 16), np.float32)
    with pytest.raises(ValueError):
        sscRaft.fbp(y, dic={"gpu": [0], "angles[rad]": np.zeros(1, np.float32)})
    with pytest.raises(RuntimeError):
        sscRaft.fbp(np.zeros((1, 8, 16), np.float32), dic={"gpu": [99], "angles[rad]": ph.uniform_angles(8)})
    # library still usable afterwards
    out = sscRaft.fbp(np.zeros((1, 8, 16), np.float32), dic={"gpu": [0], "angles[rad]": ph.uniform_angles(8)})
    assert np.all(out == 0)


def test_eem_graph_replay_matches_direct_launches(gpu, monkeypatch):
    """hipGraph-replayed iterations (default) and direct launches give identical results."""
    import sscRaft

    n, a, k = 64, 60, 7
    img = ph.phantoms(n, 5, config=3, first=9)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=4), 0)
    monkeypatch.setenv("RAFT_EM_GRAPH", "1")
    x1, kl1, ex1 = sscRaft.eEMRT_GPU_KL(y, ang, k, [0], 0)
    monkeypatch.setenv("RAFT_EM_GRAPH", "0")
    x2, kl2, ex2 = sscRaft.eEMRT_GPU_KL(y, ang, k, [0], 0)
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(kl1, kl2)
    np.testing.assert_array_equal(ex1, ex2)
    xr, klr, _ = oracle.eem(y, ang, k)
    assert rel(x1, xr) <= 1e-5


@pytest.mark.parametrize("kl_iterations", [0, 1, 2, 7])
def test_eem_kl_of_first_iterations_only(gpu, kl_iterations):
    """raft_eem_device_kl: the same image as a full run, and the KL of the
    first kl_iterations iterations equal to the full run's first columns."""
    import torch
    from sscRaft import device as dv

    n, a, k = 48, 40, 5
    img = ph.phantoms(n, 3, config=3, first=11)
    ang = ph.uniform_angles(a)
    y = torch.from_numpy(np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=12), 0).astype(np.float32)).cuda()
    x1, kl1, ex1 = dv.eem_device(y, ang, k)
    x2, kl2, ex2 = dv.eem_device(y, ang, k, kl_iterations=kl_iterations)
    assert torch.equal(x1, x2)
    m = min(kl_iterations, k)
    if m == 0:
        assert kl2 is None and ex2 is None
    else:
        assert kl2.shape == (3, m)
        np.testing.assert_array_equal(kl2, kl1[:, :m])
        np.testing.assert_array_equal(ex2, ex1[:, :m])


def test_eem_kl_only_matches_first_iteration(gpu):
    """raft_eem_kl_device: KL of the given image, image unchanged, equal to KL[0] of an eEM run."""
    import torch
    from sscRaft import device as dv

    n, a = 48, 40
    img = ph.phantoms(n, 3, config=3, first=2)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=6), 0).astype(np.float32)
    x0 = np.random.default_rng(1).uniform(0.2, 1.0, (3, n, n)).astype(np.float32)
    yt, xt = torch.from_numpy(y).cuda(), torch.from_numpy(x0).cuda()
    kl, ex = dv.eem_kl_device(yt, ang, xt)
    np.testing.assert_array_equal(xt.cpu().numpy(), x0)
    _, kl_ref, ex_ref = oracle.eem(y, ang, 1, obj=x0.astype(np.float64))
    assert np.max(np.abs(kl - kl_ref[:, 0]) / kl_ref[:, 0]) <= 1e-5
    np.testing.assert_array_equal(ex, ex_ref[:, 0])
    _, kl1, _ = dv.eem_device(yt, ang, 1, obj=xt.clone())
    np.testing.assert_array_equal(kl, kl1[:, 0])


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,a,pad", [(64, 48, 2), (37, 20, 1), (256, 64, 3)])
def test_lowpass(gpu, filt, n, a, pad):
    """getFilterLowPassMultiGPU (in place) vs the oracle: FBP filter stage, offset 0, cropped to N bins."""
    import sscRaft

    rng = np.random.default_rng(n + a + pad)
    tomo = rng.standard_normal((3, a, n)).astype(np.float32)
    dic = {"gpu": [0], "filter": filt, "padding": pad, "regularization": 0.6, "beta/delta": 90.0,
           "energy[eV]": 2.0e4, "z2[m]": 0.4, "detectorPixel[m]": 0.3, "blocksize": 0}
    from sscRaft.processing.filters import lowpass_params

    pf, pi = lowpass_params(n, a, 3, sscRaft.SetDictionary(dict(dic), ("gpu",), (), ()))
    ref = oracle.lowpass(tomo, pf, pi)
    out = sscRaft.lowpass(tomo.copy(), dic=dic)
    e = rel(out, ref)
    print(f"lowpass {filt} n={n} a={a} pad={pad}: rel-L2 {e:.2e}")
    assert e <= FP32_TOL


@pytest.mark.parametrize("n,pad", [(4096, 2), (2730, 5)])
def test_filter_largest_fft_sizes(gpu, n, pad):
    """Padded rows above 8192 points (12288 = 2^12 * 3, 16380 = 2^2 * 3^2 * 5 * 7 * 13
    -> the direct-convolution path, and the 16-elements-per-thread LDS FFT)."""
    import sscRaft

    rng = np.random.default_rng(n)
    a = 6
    tomo = rng.standard_normal((2, a, n)).astype(np.float32)
    dic = {"gpu": [0], "filter": "hann", "padding": pad, "regularization": 0.0, "beta/delta": 0.0,
           "energy[eV]": 2.0e4, "z2[m]": 0.4, "detectorPixel[m]": 0.3, "blocksize": 0}
    from sscRaft.processing.filters import lowpass_params

    pf, pi = lowpass_params(n, a, 2, sscRaft.SetDictionary(dict(dic), ("gpu",), (), ()))
    ref = oracle.lowpass(tomo, pf, pi)
    out = sscRaft.lowpass(tomo.copy(), dic=dic)
    e = rel(out, ref)
    print(f"lowpass n={n} pad={pad} (npad {n * (1 + pad)}): rel-L2 {e:.2e}")
    assert e <= FP32_TOL


@pytest.mark.parametrize("n,a,s,filt", [(1024, 16, 5, "hann"), (2048, 8, 3, "ramp")])
def test_lowpass_rows_written_in_place(gpu, monkeypatch, n, a, s, filt):
    """Padded sizes 3072 and 6144 take the FP32 filter form that writes the
    filtered rows straight back into the data (no staged rows, no copy back):
    against the oracle, and bit for bit against the form that goes through
    staged rows (RAFT_FFT_FORM=regs).  Odd slice counts leave the last group
    of four partial."""
    import sscRaft

    rng = np.random.default_rng(n + s)
    tomo = rng.standard_normal((s, a, n)).astype(np.float32)
    dic = {"gpu": [0], "filter": filt, "padding": 2, "regularization": 0.0, "beta/delta": 0.0,
           "energy[eV]": 2.0e4, "z2[m]": 0.4, "detectorPixel[m]": 0.3, "blocksize": 0}
    from sscRaft.processing.filters import lowpass_params

    pf, pi = lowpass_params(n, a, s, sscRaft.SetDictionary(dict(dic), ("gpu",), (), ()))
    ref = oracle.lowpass(tomo, pf, pi)
    out = sscRaft.lowpass(tomo.copy(), dic=dict(dic))
    e = rel(out, ref)
    print(f"lowpass n={n} pad=2 (npad {3 * n}) slices={s}: rel-L2 {e:.2e}")
    assert e <= FP32_TOL
    monkeypatch.setenv("RAFT_FFT_FORM", "regs")
    staged = sscRaft.lowpass(tomo.copy(), dic=dict(dic))
    np.testing.assert_array_equal(out, staged)


def test_lowpass_device_in_place_and_offset_ignored(gpu):
    import torch
    from sscRaft import device as dv

    n, a = 96, 30
    tomo = np.random.default_rng(3).standard_normal((5, a, n)).astype(np.float32)
    pi = np.array([n, a, 5, n, 2, 0, 0, 7, 0, 0], np.int32)
    pf = np.array([0, 0, 1, 0, 0.5, 0.5, 3.0], np.float32)
    t = torch.from_numpy(tomo).cuda()
    dv.lowpass_device(t, pf, pi)
    assert rel(t.cpu().numpy(), oracle.lowpass(tomo, pf, pi)) <= FP32_TOL


@pytest.mark.parametrize("n,a,k", [(32, 30, 1), (48, 40, 4), (33, 20, 3), (40, 48, 3)])
def test_tem(gpu, n, a, k):
    """get_tEM_RT_MultiGPU vs the oracle (slice-own flat rows)."""
    import sscRaft

    rng = np.random.default_rng(n + k)
    img = ph.phantoms(n, 3, config=3, first=n + 1) * 0.5
    ang = ph.uniform_angles(a)
    data = ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n)  # -ln(counts / I0)
    flat = rng.uniform(0.9, 1.1, (3, 1, n)).astype(np.float32)
    keep = data.copy()
    x = sscRaft.tEMRT_GPU_(data, flat, ang, k, [0], 0)
    np.testing.assert_array_equal(data, keep)  # the caller's array is not overwritten
    xr = oracle.tem(np.exp(-data.astype(np.float64)).astype(np.float32), flat.reshape(3, n), ang, k)
    e = rel(x, xr)
    print(f"tem n={n} a={a} K={k}: rel-L2 {e:.2e}")
    assert e <= 1e-5 * max(1, k)


@pytest.mark.parametrize("nrays,n", [(48, 64), (64, 60)])
def test_tem_objsize_differs_from_nrays(gpu, nrays, n):
    """Transmission EM of an n x n image from nrays-bin rows (objsize != nrays)
    against the oracle; the image side comes from the initial guess."""
    import torch
    from sscRaft import device as dv

    a, k = 36, 3
    rng = np.random.default_rng(nrays + n)
    img = ph.phantoms(nrays, 2, config=3, first=n) * 0.5
    ang = ph.uniform_angles(a)
    count = np.exp(-ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n)).astype(np.float32)
    flat = rng.uniform(0.9, 1.1, (2, nrays)).astype(np.float32)
    x0 = rng.uniform(0.5, 1.5, (2, n, n)).astype(np.float32)
    x = dv.tem_device(torch.from_numpy(count).cuda(), torch.from_numpy(flat).cuda(), ang, k,
                      obj=torch.from_numpy(x0).cuda()).cpu().numpy()
    xr = oracle.tem(count, flat, ang, k, obj=x0)
    e = rel(x, xr)
    print(f"tem objsize {n} nrays {nrays}: rel-L2 {e:.2e}")
    assert x.shape == (2, n, n) and e <= 1e-5 * k


def test_tem_dispatch_and_device_entry(gpu):
    import torch
    import sscRaft
    from sscRaft.rafttypes import libraft, ctypes

    n, a, k = 40, 36, 3
    img = ph.phantoms(n, 2, config=3, first=7) * 0.5
    ang = ph.uniform_angles(a)
    data = ph.poisson_log(oracle.radon(img, ang), 1e4, seed=3)
    x1 = sscRaft.em(data, angles=ang, dic={"gpu": [0], "method": "tEMRT", "iterations": k})  # flat = ones
    count = np.exp(-data.astype(np.float32))  # what tEMRT_GPU_ passes to the C function
    xr = oracle.tem(count, np.ones((2, n), np.float32), ang, k)
    assert rel(x1, xr) <= 1e-5 * k
    xd = torch.ones((2, n, n), device="cuda")
    cd = torch.from_numpy(count).cuda()
    fd = torch.ones((2, n), device="cuda")
    pi = np.array([n, a, 2, n, 0, 0, 0, 2, k, 0], np.int32)
    rc = libraft.raft_tem_device(0, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(cd.data_ptr()),
                                 ctypes.c_void_p(fd.data_ptr()), ang.ctypes.data_as(ctypes.c_void_p),
                                 pi.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    np.testing.assert_array_equal(xd.cpu().numpy(), x1)


@pytest.mark.parametrize("nflats,is_log", [(1, False), (2, True), (3, True)])
def test_background_correction(gpu, nflats, is_log):
    """getBackgroundCorrectionMultiGPU vs the oracle (0.5 floors, two-flat blend, -log)."""
    import sscRaft

    rng = np.random.default_rng(nflats)
    S, A, R = 3, 50, 70
    flat = rng.uniform(800, 1200, (S, nflats, R)).astype(np.float32)
    dark = rng.uniform(0, 40, (S, R)).astype(np.float32)
    frames = (rng.uniform(0.05, 1.0, (S, A, R)) * 1000).astype(np.float32)
    frames[0, 0, :5] = 0.0  # below dark: the 0.5 floor
    ref = oracle.background(frames, flat, dark, is_log)
    out = sscRaft.correct_background(frames.copy(), flat, dark, [0], is_log)
    # relative error, but absolute near 0 (-log of values close to 1)
    e = float(np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1.0 if is_log else 1e-3)))
    print(f"background nflats={nflats} log={is_log}: max rel err {e:.2e}")
    assert e <= 2e-6
    # detector-order front end: frames [angles, slices, rays], flat [nflats, slices, rays], dark [1, slices, rays]
    out2 = sscRaft.correct_projections(np.swapaxes(frames, 0, 1).copy(), np.swapaxes(flat, 0, 1).copy(),
                                       dark[None].copy(), {"gpu": [0], "uselog": is_log})
    np.testing.assert_array_equal(out2, out)


def test_batch_uneven_scans_match_single_calls(gpu):
    """raft_batch_device (one graph, buckets by size) gives what the
    single-volume entry points give for each scan: eEM bit for bit, FBP (FP16
    rows) to rounding."""
    import torch
    from sscRaft import device as dv

    sizes = [32, 48, 32, 64, 48, 32]
    K = 6
    tomos, ys, angs, pfs, fouts, ems = [], [], [], [], [], []
    for k, n in enumerate(sizes):
        img = ph.phantoms(n, 1, config=6, first=k)
        ang = ph.uniform_angles(n)
        p = ph.poisson_log(oracle.radon(img, ang), 10.0 ** (3 + 2 * k / len(sizes)), seed=k)[0]
        tomos.append(torch.from_numpy(p).cuda())
        ys.append(torch.from_numpy(np.maximum(p, 0)).cuda())
        angs.append(ang)
        pfs.append([0, 0, 1, 0, 2.0 / n, 2.0 / n, 0])
        fouts.append(torch.zeros((n, n), device="cuda"))
        ems.append(torch.ones((n, n), device="cuda"))
    ms = dv.batch_device(tomos, angs, pfs, pad=2, filter_type=7, store="f16", fbp_out=fouts, ys=ys, em=ems,
                         iterations=K)
    print(f"batch of {len(sizes)} scans: graph {ms:.3f} ms")
    for k, n in enumerate(sizes):
        pf = np.array(pfs[k], np.float32)
        pi = np.array([n, n, 1, n, 2, 0, 0, 7, 0, 0], np.int32)
        ref = dv.fbp_device(tomos[k][None].contiguous(), angs[k], pf, pi, store="f16")
        # the filter transforms two rows as one complex FFT; in a batch a row's
        # partner is another scan's row instead of a zero row, so rounding differs
        assert rel(fouts[k].cpu().numpy(), ref[0].cpu().numpy()) <= 1e-4  # FP16 rows: one f16 rounding may flip
        x, _, _ = dv.eem_device(ys[k][None].contiguous(), angs[k], K, with_kl=False)
        np.testing.assert_array_equal(ems[k].cpu().numpy(), x[0].cpu().numpy())


@pytest.mark.parametrize("n,a,pad,filt,offset", [(64, 96, 2, "ramp", 0.0), (96, 64, 2, "hann", 2.0),
                                                 (100, 90, 1, "gaussian", 0.0), (64, 48, 2, "none", 0.0),
                                                 (128, 128, 2, "ramp", -3.0)])
def test_bst(gpu, n, a, pad, filt, offset):
    """getBSTMultiGPU vs the float64 restatement (oracle_np.bst)."""
    import oracle_np
    import sscRaft

    y, ang = noisy_sino(n, a, 2, seed=n + a)
    dic = fbp_dic(ang, pad=pad, filt=filt, pixel=2.0 / n, offset=offset, reg=0.5)
    dic["method"] = "BST"
    out = sscRaft.fbp(y, dic=dic)
    pf, pi = oracle.fbp_params(n, a, 2, pad=pad, filter_type=sscRaft.FilterNumber(filt), pixel=2.0 / n,
                               offset=float(int(offset)), reg=0.5)
    ref = oracle_np.bst(y, pf, pi)
    e = rel(out, ref)
    print(f"bst n={n} a={a} pad={pad} {filt} offset={offset}: rel-L2 {e:.2e}")
    assert e <= 2e-5


@pytest.mark.parametrize("n,a,pad,slices", [(64, 96, 2, 2), (100, 90, 1, 11), (48, 40, 0, 9)])
def test_bst_fused_rows_match_separate_passes(gpu, monkeypatch, n, a, pad, slices):
    """The fused polar-to-Cartesian gather + row inverse FFT (one kernel per
    grid row, RAFT_BST_FUSED default) gives the separate passes' result bit
    for bit; the slice counts leave the last group of slices of a workgroup
    partial."""
    import sscRaft

    y, ang = noisy_sino(n, a, slices, seed=3 * n + a)
    dic = fbp_dic(ang, pad=pad, filt="ramp", pixel=2.0 / n)
    dic["method"] = "BST"
    fused = sscRaft.fbp(y, dic=dict(dic))
    monkeypatch.setenv("RAFT_BST_FUSED", "0")
    separate = sscRaft.fbp(y, dic=dict(dic))
    assert np.isfinite(fused).all()
    np.testing.assert_array_equal(fused, separate)


@pytest.mark.parametrize("n,a,pad,slices,filt", [(64, 96, 2, 2, "ramp"), (100, 90, 1, 11, "hann"),
                                                  (48, 40, 0, 9, "ramp"), (64, 48, 2, 5, "none")])
def test_bst_filter_in_polar_pass_matches_separate_filter(gpu, monkeypatch, n, a, pad, slices, filt):
    """The filter inside the polar pass (k_bst_filter_polar, the default)
    against the filter as a separate pass into staged rows
    (RAFT_BST_FILTER_FUSED=0): the polar FFTs pair two slices' half-rows
    instead of one slice's two half-rows, so the sums round differently;
    within FP32 rounding (rel-L2 1e-6), and both within the oracle's gate.
    Odd slice counts leave the last pair with one slice."""
    import sscRaft

    y, ang = noisy_sino(n, a, slices, seed=5 * n + a)
    dic = fbp_dic(ang, pad=pad, filt=filt, pixel=2.0 / n)
    dic["method"] = "BST"
    one_pass = sscRaft.fbp(y, dic=dict(dic))
    monkeypatch.setenv("RAFT_BST_FILTER_FUSED", "0")
    separate = sscRaft.fbp(y, dic=dict(dic))
    assert np.isfinite(one_pass).all()
    e = rel(one_pass, separate)
    print(f"bst filter pass n={n} a={a} pad={pad} slices={slices} {filt}: rel-L2 {e:.2e}")
    assert e <= 1e-6


def test_bst_filter_in_polar_pass_above_8192_points(gpu, monkeypatch):
    """The one-pass BST filter at a padded size above 8192 points (4096 bins,
    pad 2: 12288, the 16-elements-per-thread FFT) against the separate
    filter pass, on random rows (an odd slice count)."""
    import sscRaft

    n, a, S = 4096, 6, 3
    rng = np.random.default_rng(12288)
    y = rng.standard_normal((S, a, n)).astype(np.float32)
    dic = fbp_dic(ph.uniform_angles(a), pad=2, filt="ramp", pixel=2.0 / n)
    dic["method"] = "BST"
    one_pass = sscRaft.fbp(y, dic=dict(dic))
    monkeypatch.setenv("RAFT_BST_FILTER_FUSED", "0")
    separate = sscRaft.fbp(y, dic=dict(dic))
    assert np.isfinite(one_pass).all()
    e = rel(one_pass, separate)
    print(f"bst filter pass n={n} npad=12288: rel-L2 {e:.2e}")
    assert e <= 1e-6


def test_slab_split_over_repeated_device_ids(gpu):
    """The host-pointer entry points give each entry of gpus[] one contiguous
    slab of slices, each run by its own host thread bound to that device
    (fbp.cu:267-300, em_rt.cu:221-259, radon_rt.cu:259-294).  With the one
    device listed several times, the threads share it.  eEM (+ KL) and Radon
    equal the one-slab results bit for bit (slices are independent).  The
    FBP filter transforms the rows of two slices in one complex FFT, so FBP
    is bit-identical when the slabs keep those pairs (even slab sizes) and
    within FP32 rounding (rel-L2 1e-6) when a slab boundary splits one.  A
    device number past the device count fails the whole call before any
    thread starts."""
    import sscRaft

    n, a, S = 48, 40, 7  # three slabs: 3, 3 and 1 slices
    y, ang = noisy_sino(n, a, S + 1, seed=31)
    d = dict(fbp_dic(ang, pad=2, pixel=2.0 / n), gpu=[0])
    one = sscRaft.fbp(y, dic=dict(d))
    two = sscRaft.fbp(y, dic=dict(d, gpu=[0, 0]))  # slabs of 4 and 4
    np.testing.assert_array_equal(two, one)
    y = y[:S]
    one = one[:S]
    three = sscRaft.fbp(y, dic=dict(d, gpu=[0, 0, 0]))
    assert rel(three, one) < 1e-6
    yem = np.maximum(y, 0).astype(np.float32)
    x1, kl1, ex1 = sscRaft.eEMRT_GPU_KL(yem, ang, 4, [0], 0)
    x3, kl3, ex3 = sscRaft.eEMRT_GPU_KL(yem, ang, 4, [0, 0, 0], 0)
    np.testing.assert_array_equal(x3, x1)
    np.testing.assert_array_equal(kl3, kl1)
    np.testing.assert_array_equal(ex3, ex1)
    img = ph.phantoms(n, S, config=9, first=40)
    p1 = sscRaft.radon_RT(img, ang, [0], pixel=2.0 / n)
    p3 = sscRaft.radon_RT(img, ang, [0, 0, 0], pixel=2.0 / n)
    np.testing.assert_array_equal(p3, p1)
    with pytest.raises(RuntimeError, match="invalid device"):
        sscRaft.fbp(y, dic=dict(fbp_dic(ang, pad=2, pixel=2.0 / n), gpu=[0, 4096]))


def test_blocked_calls_match_one_block(gpu):
    """Volumes split into blocks of slices (parami[9] = blocksize, the
    reference's memory-driven blocking) give what one block gives: FBP and
    the eEM iterate bit for bit, the FP64 KL up to summation order."""
    import sscRaft

    n, a, S = 48, 40, 9
    y, ang = noisy_sino(n, a, S, seed=21)
    d0 = dict(fbp_dic(ang, pad=2, pixel=2.0 / n), blocksize=0)
    d4 = dict(fbp_dic(ang, pad=2, pixel=2.0 / n), blocksize=4)
    np.testing.assert_array_equal(sscRaft.fbp(y, dic=d4), sscRaft.fbp(y, dic=d0))
    ye = np.maximum(y, 0).astype(np.float32)
    x0, kl0, ex0 = sscRaft.eEMRT_GPU_KL(ye, ang, 4, [0], 0)
    x4, kl4, ex4 = sscRaft.eEMRT_GPU_KL(ye, ang, 4, [0], 4)
    np.testing.assert_array_equal(x4, x0)
    np.testing.assert_allclose(kl4, kl0, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(ex4, ex0)
    r0 = sscRaft.radon_RT(x0, ang, [0], pixel=2.0 / n)
    np.testing.assert_array_equal(sscRaft.radon_RT(x0[:5], ang, [0], pixel=2.0 / n), r0[:5])


@pytest.mark.parametrize("depth", ["2", "3"])
def test_host_pointer_fbp_pipeline_depths(gpu, monkeypatch, depth):
    """The host-pointer FBP's chunk pipeline (three chunks in flight by
    default, two with RAFT_COPY_DEPTH=2) over six chunks of 4 slices (the last
    one partial), so every staging buffer is reused: bit for bit what one
    block gives."""
    import sscRaft

    n, a, S = 48, 40, 22
    y, ang = noisy_sino(n, a, S, seed=29)
    one = sscRaft.fbp(y, dic=dict(fbp_dic(ang, pad=2, pixel=2.0 / n), blocksize=0))
    monkeypatch.setenv("RAFT_COPY_DEPTH", depth)
    chunked = sscRaft.fbp(y, dic=dict(fbp_dic(ang, pad=2, pixel=2.0 / n), blocksize=4))
    np.testing.assert_array_equal(chunked, one)


def test_zero_slices_are_a_no_op(gpu):
    """Device entry points accept an empty slice range and touch nothing."""
    import torch
    from sscRaft import device as dv
    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    n = a = 32
    ang = ph.uniform_angles(a)
    empty = torch.zeros((0, a, n), dtype=torch.float32, device="cuda")
    pf, pi = fbp_params(n, a, 0, dict(fbp_dic(ang, pixel=2.0 / n), blocksize=0))
    out = dv.fbp_device(empty, ang, pf, pi, out=torch.zeros((0, n, n), device="cuda"))
    assert out.shape == (0, n, n)
    proj = dv.radon_device(torch.zeros((0, n, n), device="cuda"), ang, out=torch.zeros((0, a, n), device="cuda"))
    assert proj.shape == (0, a, n)
    torch.cuda.synchronize()


def test_batch_work_queue_matches_per_bucket_branches(gpu, monkeypatch):
    """The persistent work-queue launches (one per stage and iteration over all
    buckets, default) and the per-bucket graph branches (RAFT_BATCH_QUEUE=0)
    run the same per-tile arithmetic: FBP (FP32 and FP16 rows) and eEM outputs
    are bit-identical.  Sizes 40..96 (several buckets, ragged slice groups)."""
    import torch
    from sscRaft import device as dv

    sizes = [64, 40, 96, 64, 40, 72, 64, 96, 40]
    K = 4
    data = []
    for k, n in enumerate(sizes):
        img = ph.phantoms(n, 1, config=6, first=40 + k)
        ang = ph.uniform_angles(n)
        p = ph.poisson_log(oracle.radon(img, ang), 1e4, seed=40 + k)[0]
        data.append((torch.from_numpy(p).cuda(), torch.from_numpy(np.maximum(p, 0)).cuda(), ang,
                     [0, 0, 1, 0, 2.0 / n, 2.0 / n, 0]))
    for store in ("f32", "f16"):
        outs = {}
        for mode in ("1", "0"):
            monkeypatch.setenv("RAFT_BATCH_QUEUE", mode)
            fouts = [torch.zeros((n, n), device="cuda") for n in sizes]
            ems = [torch.ones((n, n), device="cuda") for n in sizes]
            dv.batch_device([d[0] for d in data], [d[2] for d in data], [d[3] for d in data], pad=2, filter_type=7,
                            store=store, fbp_out=fouts, ys=[d[1] for d in data], em=ems, iterations=K)
            outs[mode] = ([f.cpu().numpy() for f in fouts], [e.cpu().numpy() for e in ems])
        for a, b in zip(outs["1"][0] + outs["1"][1], outs["0"][0] + outs["0"][1]):
            np.testing.assert_array_equal(a, b)


def test_eem_schedule_matches_per_slice_runs(gpu):
    """raft_eem_device_sched: slice s after iterations[s] iterations equals a
    run of that slice alone for that many iterations, bit for bit, including
    slices that stop inside a slice group and slices that never start; the
    staged and direct projector forms both."""
    import torch
    from sscRaft import device as dv

    counts = [7, 7, 5, 4, 4, 2, 1, 0, 0]
    for n, a in [(64, 64), (96, 24)]:  # staged form, direct form (sparse angles)
        img = ph.phantoms(n, len(counts), config=8, first=n)
        ang = ph.uniform_angles(a)
        y = torch.from_numpy(np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n), 0)
                             .astype(np.float32)).cuda()
        x0 = torch.from_numpy(np.random.default_rng(n).uniform(0.5, 1.5, (len(counts), n, n)).astype(np.float32)).cuda()
        xs = dv.eem_device_sched(y, ang, counts, obj=x0.clone())
        for s, k in enumerate(counts):
            if k == 0:
                assert torch.equal(xs[s], x0[s])
                continue
            x1, _, _ = dv.eem_device(y[s:s + 1].contiguous(), ang, k, obj=x0[s:s + 1].clone(), with_kl=False)
            assert torch.equal(xs[s], x1[0]), (n, a, s, k)
    with pytest.raises(ValueError, match="non-increasing"):
        dv.eem_device_sched(y[:3].contiguous(), ang, [1, 2, 1])

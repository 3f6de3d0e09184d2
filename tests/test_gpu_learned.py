"""The learned eEM warm start on the GPU: grid mapping, lowest-KL targets,
and what the trained network (ssc-raft_amd/sscRaft/weights/) gains over the
FBP start and the uniform start on held-out slices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fbp_to_em_grid_orientation(gpu):
    """The FBP image moved onto the eEM grid lines up with the phantom (the
    raw FBP grid is flipped in y and shifted by one column)."""
    import torch
    from sscRaft import learned, phantom as ph
    from sscRaft.device import radon_device

    cfg = learned.WarmStartConfig
    n, a = 256, 384
    img = ph.phantoms(n, 2, config=5, first=11, device="cuda")
    ang = ph.uniform_angles(a)
    p = radon_device(img.contiguous(), ang)

    class C(cfg):
        pass

    C.n, C.angles = n, a
    fb = learned.fbp_start(p, ang, C)
    raw = torch.empty_like(fb)
    from sscRaft.geometries.parallel.fbp_methods import fbp_params
    from sscRaft.device import fbp_device

    pf, pi = fbp_params(n, a, 2, learned._fbp_dic(n, 2))
    fbp_device(p, ang, pf, pi, out=raw)
    ref = img.cpu().numpy()[:, 4:-4, 4:-4]
    err_mapped = np.linalg.norm(fb.cpu().numpy()[:, 4:-4, 4:-4] - ref) / np.linalg.norm(ref)
    err_raw = np.linalg.norm(raw.cpu().numpy()[:, 4:-4, 4:-4] - ref) / np.linalg.norm(ref)
    print(f"FBP vs phantom: on the eEM grid {err_mapped:.3f}, raw FBP grid {err_raw:.3f}")
    assert err_mapped < 0.15 and err_mapped < 0.5 * err_raw


def test_best_kl_iterate_is_reproducible(gpu):
    """The lowest-KL iterate kept by best_kl_eem has the KL recorded for it."""
    import torch
    from sscRaft import learned
    from sscRaft.device import eem_kl_device

    class C(learned.WarmStartConfig):
        n, angles = 128, 64

    p_hat, y, _, ang = learned.sinograms(3, 500, C)
    best, best_kl, best_it, curve = learned.best_kl_eem(y, ang, 60)
    kl, _ = eem_kl_device(y, ang, best)
    np.testing.assert_allclose(kl, best_kl, rtol=1e-12)
    np.testing.assert_array_equal(best_kl, curve.min(axis=1))
    assert torch.isfinite(best).all()


def test_trained_network_gains_over_fbp_and_uniform_starts(gpu):
    """On held-out slices the trained warm start reaches plain eEM's lowest KL
    in fewer iterations than the FBP start, which needs fewer than plain eEM."""
    import os

    from sscRaft import learned
    from sscRaft.device import eem_device
    from sscRaft.nn import UNet

    path = learned.default_weights_path()
    assert os.path.exists(path), "trained weights are part of the package"
    net = UNet.load(path)
    cfg = learned.WarmStartConfig
    p_hat, y, _, ang = learned.sinograms(24, cfg.heldout_first + 5000, cfg)
    _, kl_plain, ex_plain = eem_device(y, ang, cfg.em_iterations)
    kl_plain = learned.valid_kl(kl_plain, ex_plain)
    target, k_plain = kl_plain.min(axis=1), kl_plain.argmin(axis=1)
    x0, accept, xf, (kl0_net, kl0_fbp) = learned.warm_start(net, p_hat, y, ang, cfg)
    K = 60
    _, kn, en = eem_device(y, ang, K, obj=x0.clone())
    _, kf, ef = eem_device(y, ang, K, obj=xf.clone())
    k_net = learned.iterations_to_reach(learned.valid_kl(kn, en), target)
    k_fbp = learned.iterations_to_reach(learned.valid_kl(kf, ef), target)
    print(f"median iterations to plain eEM's lowest KL: plain {np.median(k_plain)}, FBP start "
          f"{np.median(k_fbp)}, learned start {np.median(k_net)}; accepted {accept.sum()}/24; "
          f"start KL median net {np.median(kl0_net):.1f} vs FBP {np.median(kl0_fbp):.1f}")
    assert accept.sum() >= 20
    assert np.median(kl0_net) < 0.5 * np.median(kl0_fbp)
    assert np.median(k_net[k_net >= 0]) < np.median(k_fbp[k_fbp >= 0]) < np.median(k_plain)
    assert (k_net >= 0).sum() >= 22


def test_learned_eem_equals_warm_start_then_eem(gpu, monkeypatch):
    """learned_eem (acceptance folded into the first iteration's KL) gives the
    same acceptance and the same image as the warm start's separate KL checks
    followed by eEM; one slice's network start is spoiled so that it is
    rejected and rerun from its FBP start."""
    import torch

    from sscRaft import learned
    from sscRaft.device import eem_device, eem_kl_device
    from sscRaft.nn import UNet

    net = UNet.load(learned.default_weights_path())
    cfg = learned.WarmStartConfig
    p_hat, y, _, ang = learned.sinograms(8, cfg.heldout_first + 7000, cfg)
    orig = learned.warm_start

    def spoiled(*args, **kwargs):
        x0, acc, xf, kls = orig(*args, **kwargs)
        x0[3].mul_(40.0)  # far worse than the FBP start
        return x0, acc, xf, kls

    monkeypatch.setattr(learned, "warm_start", spoiled)
    x0, _, xf, _ = learned.warm_start(net, p_hat, y, ang, cfg, check=False)
    kn, en = eem_kl_device(y, ang, x0)
    kf, ef = eem_kl_device(y, ang, xf)
    acc_ref = (kn <= kf) & (en <= ef)
    assert not acc_ref[3] and acc_ref.sum() >= 6
    ref = torch.where(torch.from_numpy(acc_ref).cuda()[:, None, None], x0, xf).contiguous()
    eem_device(y, ang, 5, obj=ref, with_kl=False)
    x, acc = learned.learned_eem(net, p_hat, y, ang, 5, cfg)
    np.testing.assert_array_equal(acc, acc_ref)
    assert torch.equal(x, ref)



def test_learned_eem_finishing_slices_start_from_fbp_in_the_same_run(gpu):
    """Slices listed in `finish` whose FBP start needs no more iterations than
    the run has start from their FBP start inside the run: the result equals
    the acceptance rule and eEM over the warm start with those slices' starts
    replaced by their FBP starts; a finishing count above the run's
    iterations keeps the separate run afterwards."""
    import torch

    from sscRaft import learned
    from sscRaft.device import eem_device, eem_kl_device
    from sscRaft.nn import UNet

    net = UNet.load(learned.default_weights_path())
    cfg = learned.WarmStartConfig
    p_hat, y, _, ang = learned.sinograms(6, cfg.heldout_first + 7100, cfg)
    x0, _, xf, _ = learned.warm_start(net, p_hat, y, ang, cfg, check=False)
    kf, ef = eem_kl_device(y, ang, xf)

    def accepted_start(start):
        kn, en = eem_kl_device(y, ang, start)
        acc = torch.from_numpy((kn <= kf) & (en <= ef)).cuda()
        return torch.where(acc[:, None, None], start, xf).contiguous()

    fin = [1, 4]
    x0f = x0.clone()
    x0f[fin] = xf[fin]
    ref = accepted_start(x0f)
    eem_device(y, ang, 5, obj=ref, with_kl=False)
    x, acc = learned.learned_eem(net, p_hat, y, ang, 5, cfg, finish=(np.array(fin), 3))
    assert acc[fin].all()
    assert torch.equal(x, ref)
    # more finishing iterations than the run: those slices are run again, from
    # their FBP start, for 7 iterations after it
    ref2 = accepted_start(x0.clone())
    eem_device(y, ang, 5, obj=ref2, with_kl=False)
    xr = xf[fin].contiguous()
    eem_device(y[fin].contiguous(), ang, 7, obj=xr, with_kl=False)
    x2, _ = learned.learned_eem(net, p_hat, y, ang, 5, cfg, finish=(np.array(fin), 7))
    keep = [i for i in range(6) if i not in fin]
    assert torch.equal(x2[keep], ref2[keep])
    assert torch.equal(x2[fin], xr)

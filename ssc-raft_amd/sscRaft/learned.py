"""Learned eEM warm start: pair generation, training and the time-to-KL
comparison against plain eEM, all on one GPU with libraft's kernels.

Workload (one slice): 512 x 512 random-ellipse phantom, 128 angles on [0, pi)
x 512 bins, Poisson noise at I0 = 1e4, y = max(p_hat, 0).
  * network input: FBP (ramp, padding 2, pixel 2/N) of p_hat, moved onto the
    eEM grid (the FBP grid is the eEM grid flipped in y and shifted one
    column: x_em[Y][X] ~ fbp[N-1-Y][X+1], checked against the oracle);
  * target: the plain eEM iterate (uniform start) with the lowest KL within
    200 iterations.  The reference eEM (sscRaft/em_methods.py:3-70,
    em_RT_kernel.cu:188-350) does not converge at this size: its
    backprojector is a shifted transpose of its projector, and the KL, after
    falling for some tens of iterations, grows without bound (inf / NaN by
    iteration 200).  The lowest-KL iterate is what plain eEM can offer.
  * warm start: pred = max(net(input), floor); accepted per slice only if
    KL(y || A pred) <= KL(y || A fbp_start), else the FBP start is used.
  * time-to-KL: wall time until every slice's KL is at or below the lowest
    KL its plain run reached; learned = FBP + resampling + network + the
    acceptance check + eEM iterations, plain = eEM iterations from ones.

The reference has no learned components; everything here is new.
"""
import time

import numpy as np

from . import phantom as ph
from .device import eem_device, eem_device_sched, eem_kl_device, fbp_device, radon_device
from .geometries.parallel.fbp_methods import fbp_params
from .nn import UNet, _stream, nn_call

__all__ = ["WarmStartConfig", "make_pairs", "fbp_to_em_grid", "train", "evaluate", "warm_start"]


class WarmStartConfig:
    n = 512
    angles = 128
    i0 = 1.0e4
    pad = 2
    em_iterations = 200
    floor = 1.0e-4
    phantom_config = 5     # seed family: 1000 * config + index
    train_first = 0        # training phantoms use indices [train_first, train_first + count)
    heldout_first = 100000  # held-out phantoms start here


def _fbp_dic(n, pad):
    return {"filter": "ramp", "beta/delta": 0.0, "regularization": 1.0, "rotation axis offset": 0.0,
            "blocksize": 0, "energy[eV]": 1.0, "z2[m]": 0.0, "detectorPixel[m]": 2.0 / n, "padding": pad}


def fbp_to_em_grid(fbp, out=None, stream=None):
    """out[s][Y][X] = fbp[s][N-1-Y][min(X+1, N-1)] on the device (raft_nn_resample)."""
    import torch

    S, N, _ = fbp.shape
    if out is None:
        out = torch.empty_like(fbp)
    nn_call("raft_nn_resample", fbp.data_ptr(), out.data_ptr(), S, N, N, N, N, 1.0, 1.0, float(N - 1), -1.0,
            _stream(stream))
    return out


def sinograms(count, first, cfg=WarmStartConfig, chunk=256):
    """(p_hat [count, A, N], y = max(p_hat, 0), phantoms) on the GPU, seeded per slice."""
    import torch

    ang = ph.uniform_angles(cfg.angles)
    imgs, phat = [], []
    for c0 in range(0, count, chunk):
        c = min(chunk, count - c0)
        img = ph.phantoms(cfg.n, c, config=cfg.phantom_config, first=first + c0, device="cuda")
        p = radon_device(img.contiguous(), ang)  # pixel 2/N: the wrapper's nrays*pixel/2 factor is 1
        phat.append(ph.poisson_log_torch(p, cfg.i0, seed=1000 * cfg.phantom_config + first + c0))
        imgs.append(img)
    p_hat = torch.cat(phat)
    return p_hat, torch.clamp(p_hat, min=0.0).contiguous(), torch.cat(imgs), ang


def fbp_start(p_hat, ang, cfg=WarmStartConfig):
    """Network input: FBP of p_hat on the eEM grid."""
    S = p_hat.shape[0]
    paramf, parami = fbp_params(cfg.n, cfg.angles, S, _fbp_dic(cfg.n, cfg.pad))
    f = fbp_device(p_hat.contiguous(), ang, paramf, parami)
    return fbp_to_em_grid(f)


def valid_kl(kl, ex):
    """KL curve with invalid iterates set to +inf.

    The KL skips bins whose projection is not positive (and counts them); a
    diverging eEM iterate with NaN pixels projects to NaN bins, which the sum
    skips, so its KL looks small.  An iterate is valid only while its count of
    skipped bins has not grown since iteration 0 and its KL is finite."""
    bad = ~np.isfinite(kl) | (ex > ex[:, :1])
    return np.where(bad, np.inf, kl)


def best_kl_eem(y, ang, iterations, chunk=512):
    """Plain eEM from ones for `iterations`, keeping each slice's lowest-KL valid iterate.

    Returns (best [S, N, N], best_kl [S], best_iter [S], kl_curve [S, iterations])."""
    import torch

    S, A, N = y.shape
    best = torch.ones((S, N, N), dtype=torch.float32, device=y.device)
    best_kl = np.full(S, np.inf)
    best_it = np.zeros(S, dtype=np.int64)
    curve = np.zeros((S, iterations))
    for c0 in range(0, S, chunk):
        c1 = min(S, c0 + chunk)
        yc = y[c0:c1].contiguous()
        x = torch.ones((c1 - c0, N, N), dtype=torch.float32, device=y.device)
        prev = torch.empty_like(x)
        ex0 = None
        for it in range(iterations):
            prev.copy_(x)
            _, kl, ex = eem_device(yc, ang, 1, obj=x)  # kl = KL of prev; x advances one iteration
            if ex0 is None:
                ex0 = ex[:, 0].copy()
            finite = torch.isfinite(prev).flatten(1).all(dim=1).cpu().numpy()
            k = np.where(finite & np.isfinite(kl[:, 0]) & (ex[:, 0] <= ex0), kl[:, 0], np.inf)
            curve[c0:c1, it] = k
            better = np.isfinite(k) & (k < best_kl[c0:c1])
            if better.any():
                idx = torch.from_numpy(np.nonzero(better)[0]).to(y.device)
                best[c0:c1].index_copy_(0, idx, prev.index_select(0, idx))
                best_kl[c0:c1][better] = k[better]
                best_it[c0:c1][better] = it
    return best, best_kl, best_it, curve


def make_pairs(count, first, cfg=WarmStartConfig, log=print):
    """Training / evaluation pairs: (input [count, N, N], target [count, N, N], y, ang, info)."""
    import torch

    t0 = time.time()
    p_hat, y, _, ang = sinograms(count, first, cfg)
    inp = fbp_start(p_hat, ang, cfg)
    del p_hat
    target, best_kl, best_it, _ = best_kl_eem(y, ang, cfg.em_iterations)
    if not bool(torch.isfinite(target).all()) or not bool(torch.isfinite(inp).all()):
        raise RuntimeError("non-finite training pair")
    log(f"pairs: {count} slices in {time.time() - t0:.1f} s; lowest-KL iteration median {np.median(best_it):.0f} "
        f"(min {best_it.min()}, max {best_it.max()})")
    return inp, target, y, ang, {"best_kl": best_kl, "best_iter": best_it}


def train(net, inp, target, epochs, batch=8, lr=1e-3, lr_end=1e-5, seed=0, log=print, log_every=100):
    """Adam on mean squared error, cosine learning-rate decay; data stays on the device."""
    import torch

    count = inp.shape[0]
    steps = epochs * (count // batch)
    net.autotune(batch, inp.shape[1], inp.shape[2], backward=True)
    rng = np.random.default_rng(seed)
    step = 0
    t0 = time.time()
    hist = []
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(count)).to(inp.device)
        for b0 in range(0, count - batch + 1, batch):
            idx = order[b0:b0 + batch]
            xb = inp.index_select(0, idx).contiguous()
            tb = target.index_select(0, idx).contiguous()
            cur = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * step / max(steps, 1)))
            loss = net.train_step(xb, tb, lr=float(cur))
            step += 1
            if step % log_every == 0 or step == steps:
                l = loss.item()
                hist.append((step, l))
                log(f"epoch {ep} step {step}/{steps} lr {cur:.2e} loss {l:.4e} ({time.time() - t0:.0f} s)")
    return hist


# images per network call in the warm start (fewer, larger persistent launches:
# 512 images of 512^2 take 51.5 ms at 128 per call, 52.0 ms at 64)
INFERENCE_BATCH = 128


def warm_start(net, p_hat, y, ang, cfg=WarmStartConfig, batch=None, check=True):
    """Starting images for eEM: network output (clamped to cfg.floor), with the
    per-slice acceptance rule against the FBP start.

    Returns (x0, accepted mask, fbp start, (KL of the network start, KL of the FBP start))."""
    import torch

    batch = INFERENCE_BATCH if batch is None else batch
    fb = fbp_start(p_hat, ang, cfg)
    x0 = torch.empty_like(fb)
    for nb in {min(batch, fb.shape[0]), fb.shape[0] % batch} - {0}:
        if (nb, cfg.n, cfg.n, 32, 32) not in net.tuned:
            net.autotune(nb, cfg.n, cfg.n)
    for b0 in range(0, fb.shape[0], batch):
        b1 = min(fb.shape[0], b0 + batch)
        net.forward(fb[b0:b1], out=x0[b0:b1], floor=cfg.floor)
    xf = torch.clamp(fb, min=cfg.floor)
    if not check:
        return x0, None, xf, None
    kl_net, ex_net = eem_kl_device(y, ang, x0)
    kl_fbp, ex_fbp = eem_kl_device(y, ang, xf)
    accept = (kl_net <= kl_fbp) & (ex_net <= ex_fbp)
    if not accept.all():
        idx = torch.from_numpy(np.nonzero(~accept)[0]).to(x0.device)
        x0.index_copy_(0, idx, xf.index_select(0, idx))
    return x0, accept, xf, (kl_net, kl_fbp)


def learned_eem(net, p_hat, y, ang, iterations, cfg=WarmStartConfig, batch=None, finish=None):
    """eEM from the learned start, with the acceptance rule of warm_start folded
    into the first iteration: the network start's KL(y || A x0) is the KL the
    first iteration computes (eem_device(..., kl_iterations=1)), so only the
    FBP start needs a projection of its own.  Slices whose network start has a
    higher KL (or more excluded bins) than their FBP start are run again from
    the FBP start.  The result equals warm_start followed by eem_device.
    finish = (slice indices, iterations): slices whose learned start is known
    not to reach their target, and the iterations their FBP start needs.  When
    that is at most `iterations`, they start from their FBP start in the same
    run (which gives them `iterations`); otherwise they are run again from it
    afterwards for that many iterations.

    Returns (x [S, N, N], accepted mask)."""
    import torch

    x0, _, xf, _ = warm_start(net, p_hat, y, ang, cfg, batch=batch, check=False)
    if finish is not None and len(finish[0]) > 0 and 0 < finish[1] <= iterations:
        idx = torch.from_numpy(np.asarray(finish[0], np.int64)).to(x0.device)
        x0.index_copy_(0, idx, xf.index_select(0, idx))
        finish = None
    kl_fbp, ex_fbp = eem_kl_device(y, ang, xf)
    if iterations == 0:
        kl_net, ex_net = eem_kl_device(y, ang, x0)
        x = x0
    else:
        x, kl, ex = eem_device(y, ang, iterations, obj=x0, kl_iterations=1)
        kl_net, ex_net = kl[:, 0], ex[:, 0]
    accept = (kl_net <= kl_fbp) & (ex_net <= ex_fbp)
    if not accept.all():
        idx = torch.from_numpy(np.nonzero(~accept)[0]).to(x.device)
        xr = xf.index_select(0, idx).contiguous()
        if iterations > 0:
            eem_device(y.index_select(0, idx).contiguous(), ang, iterations, obj=xr, with_kl=False)
        x.index_copy_(0, idx, xr)
    if finish is not None and len(finish[0]) > 0:
        idx = torch.from_numpy(np.asarray(finish[0], np.int64)).to(x.device)
        xr = xf.index_select(0, idx).contiguous()
        if finish[1] > 0:
            eem_device(y.index_select(0, idx).contiguous(), ang, int(finish[1]), obj=xr, with_kl=False)
        x.index_copy_(0, idx, xr)
    return x, accept


def iterations_to_reach(curve, target):
    """First k (number of updates) with curve[:, k] <= target per slice, -1 if never."""
    hit = curve <= target[:, None]
    first = np.where(hit.any(axis=1), hit.argmax(axis=1), -1)
    return first


def evaluate(net, count=512, cfg=WarmStartConfig, max_iterations=None, log=print):
    """Time-to-KL of learned-init vs plain eEM on held-out slices (see module docstring)."""
    import torch

    K = cfg.em_iterations if max_iterations is None else max_iterations
    p_hat, y, _, ang = sinograms(count, cfg.heldout_first, cfg)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    # plain eEM: KL curve over K iterations (untimed exploration), then the timed run to the lowest KL
    _, kl_plain, ex_plain = eem_device(y, ang, K)
    kl_plain = valid_kl(kl_plain, ex_plain)
    target = kl_plain.min(axis=1)
    k_plain = kl_plain.argmin(axis=1)
    Kp = int(k_plain.max())
    _, t_plain = timed(lambda: eem_device(y, ang, Kp, with_kl=False))
    # FBP start and learned start: iterations needed (conv tiles tuned beforehand, untimed)
    for nb in {min(INFERENCE_BATCH, count), count % INFERENCE_BATCH} - {0}:
        net.autotune(nb, cfg.n, cfg.n)
    (x0, accept, xf, kl0), t_ws = timed(lambda: warm_start(net, p_hat, y, ang, cfg))
    _, kl_net, ex_net = eem_device(y, ang, K, obj=x0.clone())
    _, kl_fbp, ex_fbp = eem_device(y, ang, K, obj=xf.clone())
    kl_net, kl_fbp = valid_kl(kl_net, ex_net), valid_kl(kl_fbp, ex_fbp)
    k_net = iterations_to_reach(kl_net, target)
    k_fbp = iterations_to_reach(kl_fbp, target)
    reached = k_net >= 0
    Kl = int(k_net[reached].max()) if reached.any() else None
    Kf = int(k_fbp[k_fbp >= 0].max()) if (k_fbp >= 0).any() else None
    res = {"slices": count, "plain_iterations_to_lowest_kl_max": Kp,
           "plain_iterations_median": float(np.median(k_plain)), "plain_seconds": t_plain,
           "learned_iterations_max": Kl, "learned_iterations_median": float(np.median(k_net[reached]))
           if reached.any() else None, "learned_reached": int(reached.sum()),
           "fbp_start_iterations_median": float(np.median(k_fbp[k_fbp >= 0])) if (k_fbp >= 0).any() else None,
           "fbp_start_iterations_max": Kf, "fbp_start_reached": int((k_fbp >= 0).sum()),
           "accepted": int(accept.sum()),
           "kl_start_net_median": float(np.median(kl0[0])), "kl_start_fbp_median": float(np.median(kl0[1])),
           "kl_target_median": float(np.median(target))}
    q = [0.1, 0.25, 0.5, 0.75, 0.9, 0.99, 1.0]
    res["iterations_quantiles"] = {"q": q, "plain": np.quantile(k_plain, q).tolist(),
                                   "fbp_start": np.quantile(k_fbp[k_fbp >= 0], q).tolist() if (k_fbp >= 0).any()
                                   else None,
                                   "learned": np.quantile(k_net[reached], q).tolist() if reached.any() else None}
    res["iterations_sum"] = {"plain": int(k_plain.sum()), "fbp_start": int(k_fbp[k_fbp >= 0].sum()),
                             "learned": int(k_net[reached].sum())}
    miss = np.nonzero(~reached)[0]
    res["not_reached"] = [{"slice": int(s), "target": float(target[s]), "best_kl": float(kl_net[s].min()),
                           "best_iter": int(kl_net[s].argmin()), "kl_start": float(kl_net[s, 0]),
                           "plain_best_iter": int(k_plain[s])} for s in miss[:8]]
    # timed runs over all slices: each start runs until its last slice is
    # there; slices the learned start never brings to their target are
    # finished from their FBP start inside the timed run
    if Kl is not None:
        miss = np.nonzero(~reached & (k_fbp >= 0))[0]
        finish = (miss, int(k_fbp[miss].max())) if miss.size else None

        def learned():
            x, _ = learned_eem(net, p_hat, y, ang, Kl, cfg, finish=finish)
            return x

        learned()  # untimed: tunes and sets up the shapes (as plain's untimed run does)
        _, t_learned = timed(learned)
        res["learned_seconds"] = t_learned
        res["speedup"] = t_plain / t_learned
        done = int(reached.sum()) + int(miss.size)
        res["speedup_note"] = (
            "all slices reach the target from the learned start" if reached.all() else
            f"all {done} of {count} slices counted: {int(reached.sum())} reach the target from the learned start, "
            f"{int(miss.size)} from their FBP start"
            + (f" in the same run (they need {finish[1]} of its {Kl} iterations)" if finish and finish[1] <= Kl
               else f" ({finish[1] if finish else 0} more iterations, timed)"))
    if Kf is not None:
        def fbp_only():
            x = torch.clamp(fbp_start(p_hat, ang, cfg), min=cfg.floor)
            if Kf > 0:
                eem_device(y, ang, Kf, obj=x, with_kl=False)
            return x

        _, t_fbp = timed(fbp_only)
        res["fbp_start_seconds"] = t_fbp
        res["fbp_start_speedup"] = t_plain / t_fbp
    res["warm_start_seconds"] = t_ws
    res["scheduled"] = evaluate_scheduled(net, p_hat, y, ang, x0, xf, k_plain, k_net, k_fbp, cfg, timed)
    log(res)
    return res


def sched_eem(y, ang, x_start, counts):
    """eEM where slice s runs counts[s] iterations (raft_eem_device_sched):
    the slices are gathered in decreasing count order, run, and scattered
    back.  Returns the images [S, N, N]."""
    import torch

    order = np.argsort(-np.asarray(counts), kind="stable")
    idx = torch.from_numpy(order).to(y.device)
    ys = y.index_select(0, idx).contiguous()
    xs = x_start.index_select(0, idx).contiguous()
    eem_device_sched(ys, ang, np.asarray(counts)[order].astype(np.int32), obj=xs)
    x = torch.empty_like(xs)
    x.index_copy_(0, idx, xs)
    return x


def evaluate_scheduled(net, p_hat, y, ang, x0, xf, k_plain, k_net, k_fbp, cfg, timed):
    """Time-to-KL when every slice stops as soon as it reaches its target
    (raft_eem_device_sched): each start runs its own per-slice counts, found
    by the untimed runs of evaluate(), so the batch costs the sum of its
    slices' iterations instead of its slowest slice's count times the batch.
      plain   eEM from ones, k_plain[s] iterations;
      fbp     FBP start, k_fbp[s] iterations;
      learned warm_start (FBP, network, both starts' KL, the acceptance rule),
              then k_net[s] iterations; a slice whose accepted start never
              reaches its target within the explored iterations runs from its
              FBP start for k_fbp[s] iterations instead.
    Slices that no start brings to the target are counted and reported."""
    import torch

    dev = y.device
    k_plain = np.asarray(k_plain, np.int64)
    k_net = np.asarray(k_net, np.int64)
    k_fbp = np.asarray(k_fbp, np.int64)
    fallback = (k_net < 0) & (k_fbp >= 0)
    k_learn = np.where(k_net >= 0, k_net, np.where(fallback, k_fbp, 0))
    fb_idx = torch.from_numpy(np.nonzero(fallback)[0]).to(dev)
    r = {"slices": int(len(k_plain)), "learned_reached": int(((k_net >= 0) | fallback).sum()),
         "learned_from_fbp_start": int(fallback.sum()), "fbp_start_reached": int((k_fbp >= 0).sum()),
         "iterations_sum": {"plain": int(k_plain.sum()), "fbp_start": int(np.maximum(k_fbp, 0).sum()),
                            "learned": int(k_learn.sum())}}
    ones = torch.ones_like(x0)
    sched_eem(y, ang, ones, k_plain)  # warm-up of the scheduled path
    _, t_plain = timed(lambda: sched_eem(y, ang, torch.ones_like(x0), k_plain))

    def fbp_run():
        x = torch.clamp(fbp_start(p_hat, ang, cfg), min=cfg.floor)
        return sched_eem(y, ang, x, np.maximum(k_fbp, 0))

    def learned_run():
        xs, _, xfb, _ = warm_start(net, p_hat, y, ang, cfg)
        if fb_idx.numel():
            xs.index_copy_(0, fb_idx, xfb.index_select(0, fb_idx))
        return sched_eem(y, ang, xs, k_learn)

    _, t_fbp = timed(fbp_run)
    _, t_learn = timed(learned_run)
    r.update(plain_seconds=t_plain, fbp_start_seconds=t_fbp, learned_seconds=t_learn,
             speedup=t_plain / t_learn, fbp_start_speedup=t_plain / t_fbp)
    return r


def default_weights_path():
    import os

    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_warmstart_n512_a128.npz")


def main(argv=None):
    """python -m sscRaft.learned train|evaluate [options] (GPU)."""
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python -m sscRaft.learned")
    ap.add_argument("command", choices=["train", "evaluate"])
    ap.add_argument("--count", type=int, default=4096, help="training slices")
    ap.add_argument("--heldout", type=int, default=512, help="held-out slices for the evaluation")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--weights", default=None, help="output (train) / input (evaluate) .npz")
    ap.add_argument("--init", default=None, help="train: start from these weights instead of a random init")
    ap.add_argument("--log", default=None, help="JSON log file")
    args = ap.parse_args(argv)
    path = args.weights or default_weights_path()
    record = {"command": args.command}

    def log(msg):
        print(msg, flush=True)

    if args.command == "train":
        inp, target, _, _, info = make_pairs(args.count, WarmStartConfig.train_first, log=log)
        record["pairs"] = {"count": args.count, "best_iter_median": float(np.median(info["best_iter"]))}
        net = UNet.load(args.init) if args.init else UNet(seed=args.seed)
        record["init"] = args.init
        t0 = time.time()
        hist = train(net, inp, target, args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
        record["train_seconds"] = time.time() - t0
        record["loss_history"] = hist
        del inp, target
        net.save(path, epochs=args.epochs, count=args.count, batch=args.batch, lr=args.lr)
        log(f"weights -> {path}")
    else:
        net = UNet.load(path)
    record["evaluation"] = evaluate(net, count=args.heldout, log=log)
    if args.log:
        with open(args.log, "w") as f:
            json.dump(record, f, indent=1, default=float)


if __name__ == "__main__":
    main()

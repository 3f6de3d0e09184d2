#!/usr/bin/env python3
"""Benchmark of the parallel-beam hot path on one MI355X.

Headline (`value`): FBP slices/s on 64 slices of 2048 x 2048, 2048 angles on
[0, pi) x 2048 bins, padding 2, ramp filter, detector pixel 2/N, noisy
sinogram (Poisson, I0 = 1e4) already resident in HBM.  One step = one
raft_fbp_device call over the 64 slices (pad + FFT filter + backprojection).

Also reported on the same JSON line:
  roofline      backprojection kernel (k_backproject<float,0>): bound "valu",
                2 FLOP per gathered sample over its average launch time (HIP
                events on the launch stream) against the 157.3 TFLOP/s FP32
                vector peak; beside it the effective gather bandwidth, the
                measured HBM traffic and its fraction of 8 TB/s, the
                counter VALU issue rate (profiles/), and `lds`: the gathers'
                LDS read bytes (4 per sample) over the launch time against
                the device's conflict-free LDS read rate measured at start-up
                by raft_lds_read_probe (the binding limit of the FP32 rows);
  cpu_baseline  the FP32 restatement in oracle/ timed on this host's cores on
                a bounded sample (whole slices of the same workload);
  fp16_rows     the same FBP with FP16-stored filtered rows;
  host_pointers_pcie  the same FBP through the host-pointer entry point
                (getFBPMultiGPU), PCIe copies included;
  radon         Radon-RT of the 64 phantoms (2048^2 -> 2048 x 2048): whole
                call and projector kernel, gathered-sample roofline, FP32 port
                on the host, error vs the FP64 oracle;
  em            plain eEM on 16 slices of 1024^2 (1024 angles, I0 = 1e4,
                uniform start, 200 iterations, FP64 KL per iteration);
  em.learned_init  time-to-KL of the U-Net warm start vs plain eEM on held-out
                512^2 slices (128 angles, I0 = 1e4), with the trained weights
                in ssc-raft_amd/sscRaft/weights/ (null if they are absent);
  c1            one 512^2 slice (512 angles): GPU call time and the FP32 host
                restatement's time;
  next_rows     low-pass filter export, BST FBP and flat/dark correction at
                64 x 2048^2, transmission EM at 16 x 1024^2;
  batch         256 uneven scans (N in {256..1024}): FBP + 200 eEM iterations
                of all of them in one captured graph launch;
  unet          the warm start U-Net's forward / training throughput against
                dense BF16 MFMA (2.5 PFLOP/s).
Parity of the timed paths against the FP64 oracle is the job of tests/ (the
2048^2 FBP case is tests/test_gpu_parity.py::test_fbp_2048); here the oracle
is used only in the CPU baseline leg: the FP32 port is timed, and the FP64
oracle checks the first 4 slices of the timed GPU outputs (FBP with FP32 and
FP16 rows, Radon).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ssc-raft_amd"))

METRIC = "FBP slices/s @2048²; EM time-to-KL, learned-init vs plain, 1 GPU"
HBM_PEAK = 8.0e12  # B/s, MI355X
ERR_SLICES = 4  # slices of each workload checked against the FP64 oracle


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def read_profile_value(name, key):
    """A `key = value` figure from a committed rocprofv3 summary under
    profiles/ (HBM bytes per launch, VALU issue rate), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        vals = {}
        with open(path) as fh:
            for line in fh:
                if "=" in line and not line.lstrip().startswith(("#", "|")):
                    k, v = line.split("=", 1)
                    vals[k.strip()] = v.strip()
        return float(vals[key])
    except (OSError, KeyError, ValueError):
        return None


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, or None."""
    return read_profile_value(name, "hbm_bytes_per_dispatch")


VALU_PEAK = 157.3e12  # FLOP/s, MI355X FP32 vector
LDS_RATE = None  # measured LDS read rate of the device (raft_lds_read_probe), set in main()


def valu_roofline(kernel, samples, launch_s, bytes_gathered, traffic=None, valu_issue=None, lds_bytes_per_sample=4.0):
    """Roofline block of a gather-accumulate kernel (backprojection, projector).

    These kernels do O(N^3) gathers on O(N^2) data served from LDS, so the
    bound is VALU issue, not HBM: `achieved` is 2 FLOP per sample (coordinate
    step + accumulate) over the measured launch time, against the FP32 vector
    peak.  Beside it: the effective gather bandwidth (gathered bytes / time,
    which may exceed HBM bandwidth because LDS serves the gathers), the HBM
    fraction of the measured traffic, and the counter VALU issue rate (wave
    instructions per clock per CU, 1.0 = the issue limit) from profiles/."""
    ach = 2.0 * samples / launch_s
    r = dict(bound="valu", achieved=ach / 1e12, peak=VALU_PEAK / 1e12, unit="TFLOP/s", frac=ach / VALU_PEAK,
             traffic=traffic, kernel=kernel, launch_ms=launch_s * 1e3, samples_per_launch=samples,
             samples_per_s=samples / launch_s, effective_gather_GBps=bytes_gathered / launch_s / 1e9,
             hbm_frac=(traffic / launch_s / HBM_PEAK) if traffic else None, valu_issue_per_clk_per_cu=valu_issue)
    if LDS_RATE:
        # the gathers' LDS reads (16-byte entries of 4 FP32 slices: 4 bytes per
        # sample) against the device's measured conflict-free LDS read rate
        lds = samples * lds_bytes_per_sample
        r["lds"] = dict(read_bytes_per_launch=lds, achieved_TBps=lds / launch_s / 1e12,
                        peak_TBps=LDS_RATE["bytes_per_s"] / 1e12, frac=lds / launch_s / LDS_RATE["bytes_per_s"],
                        peak_note="raft_lds_read_probe: conflict-free 16-byte LDS reads on every CU, "
                                  f"{LDS_RATE['bytes_per_clk_per_cu']:.1f} B/clk/CU at the "
                                  f"{LDS_RATE['clock_khz'] / 1e6:.2f} GHz clock attribute")
    return r


def fbp_dic(N):
    return {"filter": "ramp", "beta/delta": 0.0, "regularization": 1.0, "rotation axis offset": 0.0,
            "blocksize": 0, "energy[eV]": 1.0, "z2[m]": 0.0, "detectorPixel[m]": 2.0 / N, "padding": 2}


def _err(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return dict(rel_l2=float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300)),
                max_abs=float(np.max(np.abs(a - ref))), max_ref=float(np.max(np.abs(ref))))


def cpu_baseline(args, host, ang, N, A):
    """The FP32 restatement in oracle/ on this host's cores, on whole slices;
    then the FP64 oracle on the first ERR_SLICES slices as the checker of the
    GPU outputs kept in `host` (FBP with FP32 and FP16 rows, Radon)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    ns = min(args.cpu_slices, host["y"].shape[0])
    ys = np.ascontiguousarray(host["y"][:ns])
    pf1, pi1 = fbp_params(N, A, ns, fbp_dic(N))
    cores = host_cores()
    t0 = time.perf_counter()
    port = oracle.fbp(ys, ang, pf1, pi1, prec="f32", nthreads=cores)
    dt = time.perf_counter() - t0
    cpu = dict(value=ns / dt, unit="slices/s", cores=cores, kind="port",
               sample=f"{ns} of the 64 slices (2048^2, 2048 angles, pad 2, ramp), "
                      f"FP32 C restatement with FP64 bin indices, {dt:.2f} s",
               host_cpu=host_cpu_model())
    log(f"CPU baseline: {cpu['value']:.3f} slices/s on {cores} threads")
    ne = host["fbp_f32"].shape[0]
    pfe, pie = fbp_params(N, A, ne, fbp_dic(N))
    t0 = time.perf_counter()
    ref = oracle.fbp(np.ascontiguousarray(host["y"][:ne]), ang, pfe, pie, prec="f64", nthreads=cores)
    cpu["errors_vs_fp64_oracle"] = dict(
        slices=ne, gpu_f32_rows=_err(host["fbp_f32"], ref), gpu_f16_rows=_err(host["fbp_f16"], ref),
        cpu_port=_err(port[:min(ns, ne)], ref[:min(ns, ne)]))
    log(f"FP64 oracle on {ne} slices: {time.perf_counter() - t0:.1f} s; GPU rel-L2 "
        f"{cpu['errors_vs_fp64_oracle']['gpu_f32_rows']['rel_l2']:.2e} (FP32 rows), "
        f"{cpu['errors_vs_fp64_oracle']['gpu_f16_rows']['rel_l2']:.2e} (FP16 rows)")
    # Radon: FP32 port on one slice (timed), FP64 oracle on ERR_SLICES slices (checker)
    img = host["img"]
    t0 = time.perf_counter()
    oracle.radon(np.ascontiguousarray(img[:1]), ang, prec="f32", nthreads=cores)
    rdt = time.perf_counter() - t0
    rref = oracle.radon(np.ascontiguousarray(img), ang, prec="f64", nthreads=cores)
    radon = dict(value=1.0 / rdt, unit="slices/s", cores=cores, kind="port",
                 sample="1 of the 64 phantoms (2048^2 -> 2048 angles x 2048 bins), FP32 C restatement",
                 errors_vs_fp64_oracle=dict(slices=img.shape[0], gpu=_err(host["radon"], rref)))
    log(f"Radon CPU: {radon['value']:.3f} slices/s; GPU rel-L2 vs FP64 {radon['errors_vs_fp64_oracle']['gpu']['rel_l2']:.2e}")
    return cpu, radon


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical CPUs)"
    except OSError:
        pass
    return None


def radon_bench(args, torch, dv, img, ang, out):
    """Radon-RT of the 64 phantoms (planar in HBM) into `out`: whole call and
    the projector kernel (HIP events on the launch stream)."""
    S, N, _ = img.shape
    A = len(ang)
    for _ in range(args.warmup):
        dv.radon_device(img, ang, out=out)
    torch.cuda.synchronize()
    dv.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dv.radon_device(img, ang, out=out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    k_ms = dv.last_kernel_ms(dv.KF_PROJ)
    dv.set_profiling(False)
    samples = float(S) * A * N * N
    bytes_alg = samples * 4 + float(S) * N * N * 4 + float(S) * A * N * 4
    ks = k_ms * 1e-3
    r = dict(config=f"Radon-RT {S} x 2048^2 -> {A} angles x {N} bins", slices_per_s=S * args.steps / dt,
             ms_per_step=1e3 * dt / args.steps, kernel="k_project_lds<0>", kernel_ms=k_ms,
             samples_per_s=samples / ks,
             roofline=valu_roofline("k_project_lds<0>", samples, ks, bytes_alg,
                                    valu_issue=read_profile_value("backproject_projector_counters.txt",
                                                                  "valu_issue.k_project_lds<0>")))
    log(f"Radon: {r['slices_per_s']:.1f} slices/s, projector {k_ms:.3f} ms/launch")
    return r


def pcie_bench(args, y, ang, N):
    """The host-pointer entry point (getFBPMultiGPU through sscRaft.fbp) on the
    same 64 slices: host -> device copy, FBP, device -> host copy."""
    import sscRaft

    yh = y.cpu().numpy()
    dic = dict(fbp_dic(N), **{"gpu": [0], "angles[rad]": ang, "method": "RT"})
    sscRaft.fbp(yh, dic=dic)
    reps = 5
    calls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        sscRaft.fbp(yh, dic=dic)
        calls.append(time.perf_counter() - t0)
    dt = sum(calls) / reps
    r = dict(slices_per_s=yh.shape[0] / dt, seconds_per_call=dt, seconds_per_call_min=min(calls),
             seconds_per_call_max=max(calls), bytes_moved=float(yh.nbytes) + float(yh.shape[0]) * N * N * 4,
             note="sscRaft.fbp on host numpy arrays (pageable memory): includes both PCIe copies; every call "
                  "returns a freshly allocated output; mean of 5 calls")
    log(f"FBP via host pointers: {r['slices_per_s']:.1f} slices/s ({dt:.3f} s per call)")
    return r


def fbp_bench(args, torch, dv, ph):
    N = A = 2048
    S = 64
    dev = torch.device("cuda:0")
    ang = ph.uniform_angles(A)
    t0 = time.perf_counter()
    img = ph.phantoms(N, S, config=2, device=dev)
    p = dv.radon_device(img, ang)
    y = ph.poisson_log_torch(p, 1.0e4, seed=2000)
    torch.cuda.synchronize()
    log(f"inputs: {S} x {A} x {N} sinogram generated on the GPU in {time.perf_counter() - t0:.1f} s")
    radon = radon_bench(args, torch, dv, img, ang, p)
    ne = min(ERR_SLICES, S)
    host = dict(y=y[:max(ne, args.cpu_slices)].cpu().numpy(), img=img[:ne].cpu().numpy(), radon=p[:ne].cpu().numpy())
    del img, p
    if args.copy_calibration:
        # elementwise copy kernel: reads and writes exactly 4 * S * A * N bytes
        for _ in range(3):
            y2 = y.clone()
        torch.cuda.synchronize()
        log(f"copy calibration: {3} copies of {4 * y.numel() / 2**20:.1f} MiB")
        del y2
    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    paramf, parami = fbp_params(N, A, S, fbp_dic(N))
    tlo, width, fft_path = dv.fbp_plan_info(ang, paramf, parami)
    out = torch.empty((S, N, N), dtype=torch.float32, device=dev)
    res = {}
    for store in ("f32", "f16"):
        for _ in range(args.warmup):
            dv.fbp_device(y, ang, paramf, parami, out=out, store=store)
        torch.cuda.synchronize()
        dv.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dv.fbp_device(y, ang, paramf, parami, out=out, store=store)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        bp_ms = dv.last_kernel_ms(dv.KF_BP)
        filt_ms = dv.last_kernel_ms(dv.KF_FILTER)
        dv.set_profiling(False)
        res[store] = dict(seconds=dt, ms_per_step=1e3 * dt / args.steps, slices_per_s=S * args.steps / dt,
                          bp_ms=bp_ms, filter_ms=filt_ms)
        host["fbp_" + store] = out[:ne].cpu().numpy()
        log(f"FBP {store}: {res[store]['slices_per_s']:.1f} slices/s, step {res[store]['ms_per_step']:.2f} ms, "
            f"backprojection {bp_ms:.3f} ms/launch, filter {filt_ms:.3f} ms/launch")
    cpu, radon_cpu = (None, None) if args.no_cpu else cpu_baseline(args, host, ang, N, A)
    if radon is not None:
        radon["cpu"] = radon_cpu
    pcie = None if args.no_pcie else pcie_bench(args, y, ang, N)
    # roofline of the backprojection kernel (FP32 rows)
    samples = float(S) * A * N * N
    bytes_alg = samples * 4 + float(S) * A * width * 4 + float(S) * N * N * 4
    bp_s = res["f32"]["bp_ms"] * 1e-3
    roof = valu_roofline("k_backproject<float,0>", samples, bp_s, bytes_alg,
                         traffic=read_traffic("fbp_backproject_2048_traffic.txt"),
                         valu_issue=read_profile_value("backproject_projector_counters.txt",
                                                       "valu_issue.k_backproject<float,0>"))
    roof.update(gathered_bytes_per_launch=bytes_alg, staged_row_width=width, fft_path=bool(fft_path))
    return res, roof, cpu, radon, pcie


def next_rows_bench(args, torch, dv, ph):
    """The rows next to the hot path, device-resident, at the headline shapes:
    low-pass filter export, BST FBP and flat/dark correction on 64 x 2048 x
    2048 data, transmission EM on 16 x 1024^2 (1024 angles)."""
    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    dev = torch.device("cuda:0")
    reps = max(2, min(args.steps, 5))

    def timed(fn, reset=None):
        if reset:
            reset()
        fn()
        torch.cuda.synchronize()
        tot = 0.0
        for _ in range(reps):
            if reset:
                reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            tot += time.perf_counter() - t0
        return tot / reps

    r = {}
    N = A = 2048
    S = 64
    ang = ph.uniform_angles(A)
    img = ph.phantoms(N, S, config=2, device=dev)
    y = ph.poisson_log_torch(dv.radon_device(img, ang), 1.0e4, seed=2000)
    del img
    paramf, parami = fbp_params(N, A, S, fbp_dic(N))
    io = 2.0 * S * A * N * 4
    work = torch.empty_like(y)
    dt = timed(lambda: dv.lowpass_device(work, paramf, parami), reset=lambda: work.copy_(y))
    r["lowpass"] = dict(config="64 x 2048 angles x 2048 bins, pad 2, ramp, in place", seconds=dt,
                        slices_per_s=S / dt, hbm_gbps_in_out=io / dt / 1e9,
                        roofline_frac_in_out=io / dt / HBM_PEAK)
    out = torch.empty((S, N, N), dtype=torch.float32, device=dev)
    dt = timed(lambda: dv.bst_device(y, paramf, parami, out=out))
    r["bst"] = dict(config="64 x 2048 angles x 2048 bins -> 64 x 2048^2, pad 2, ramp", seconds=dt,
                    slices_per_s=S / dt, padded_size=N * 3)
    del out
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    raw = torch.rand((S, A, N), generator=g, device=dev) * 900.0 + 100.0
    flat = torch.rand((S, 2, N), generator=g, device=dev) * 100.0 + 1000.0
    dark = torch.rand((S, N), generator=g, device=dev) * 10.0
    dt = timed(lambda: dv.background_device(work, flat, dark, is_log=True), reset=lambda: work.copy_(raw))
    r["background"] = dict(config="64 x 2048 x 2048 frames, 2 flats blended by angle, dark, -log, in place",
                           seconds=dt, gb_per_s=io / dt / 1e9, roofline_frac=io / dt / HBM_PEAK)
    del raw, flat, dark, work, y
    n3 = A3 = 1024
    S3, K = 16, 20
    ang3 = ph.uniform_angles(A3)
    img3 = ph.phantoms(n3, S3, config=3, device=dev) * 0.5
    count = torch.exp(-ph.poisson_log_torch(dv.radon_device(img3, ang3), 1.0e4, seed=3000)).contiguous()
    flat3 = torch.ones((S3, n3), dtype=torch.float32, device=dev)
    obj = torch.ones((S3, n3, n3), dtype=torch.float32, device=dev)
    dt = timed(lambda: dv.tem_device(count, flat3, ang3, K, obj=obj), reset=lambda: obj.fill_(1.0))
    r["tem"] = dict(config=f"16 x 1024^2, 1024 angles x 1024 bins, {K} iterations", seconds=dt,
                    ms_per_iteration=1e3 * dt / K, samples_per_s=2.0 * S3 * A3 * n3 * n3 * K / dt)
    log(f"next rows: lowpass {r['lowpass']['slices_per_s']:.0f} slices/s, BST {r['bst']['slices_per_s']:.0f} "
        f"slices/s, background {r['background']['gb_per_s']:.0f} GB/s, tEM {r['tem']['ms_per_iteration']:.2f} ms/it")
    return r


def em_bench(args, torch, dv, ph):
    N = A = 1024
    S = 16
    K = args.em_iters
    dev = torch.device("cuda:0")
    ang = ph.uniform_angles(A)
    img = ph.phantoms(N, S, config=3, device=dev)
    p = dv.radon_device(img, ang)
    y = torch.clamp(ph.poisson_log_torch(p, 1.0e4, seed=3000), min=0.0).contiguous()
    del img, p
    # warm-up (kernels, workspace, graph path)
    x = torch.ones((S, N, N), dtype=torch.float32, device=dev)
    dv.eem_device(y, ang, 2, obj=x)
    torch.cuda.synchronize()
    # timed: K iterations replayed from a captured hipGraph, KL every iteration
    x = torch.ones((S, N, N), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, kl, ex = dv.eem_device(y, ang, K, obj=x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # per-kernel device times from a short run with HIP events (direct launches)
    xp = torch.ones((S, N, N), dtype=torch.float32, device=dev)
    dv.set_profiling(True)
    dv.eem_device(y, ang, min(K, 10), obj=xp)
    torch.cuda.synchronize()
    proj_ms = dv.last_kernel_ms(dv.KF_PROJ)
    embp_ms = dv.last_kernel_ms(dv.KF_EMBP)
    dv.set_profiling(False)
    samples_it = 2.0 * S * A * N * N
    half = samples_it / 2.0  # samples of each of the two kernels per iteration
    rows_bytes = float(S) * A * N * 4
    img_bytes = float(S) * N * N * 4
    roof = dict(
        projector=valu_roofline("k_project_lds<1>", half, proj_ms * 1e-3, half * 4 + img_bytes + rows_bytes,
                                valu_issue=read_profile_value("backproject_projector_counters.txt",
                                                              "valu_issue.k_project_lds<1>")),
        backproject_update=valu_roofline("k_backproject<float,1>", half, embp_ms * 1e-3,
                                         half * 4 + rows_bytes + 3 * img_bytes,
                                         valu_issue=read_profile_value("backproject_projector_counters.txt",
                                                                       "valu_issue.k_backproject<float,1>")))
    r = dict(config="16 x 1024^2, 1024 angles x 1024 bins, I0=1e4, uniform init", iterations=K, seconds=dt,
             ms_per_iteration=1e3 * dt / K, slice_iterations_per_s=S * K / dt,
             samples_per_s=samples_it * K / dt, kl_first=float(kl[:, 0].sum()), kl_last=float(kl[:, -1].sum()),
             kl_min=float(kl.min(axis=1).sum()), kl_argmin_median=float(np.median(kl.argmin(axis=1))),
             kl_excluded_last=int(ex[:, -1].sum()), projector_ms=proj_ms, backproject_update_ms=embp_ms,
             roofline=roof)
    log(f"EM plain: {K} iterations in {dt:.2f} s ({r['ms_per_iteration']:.2f} ms/it), KL {r['kl_first']:.4e} -> "
        f"{r['kl_last']:.4e}; projector {proj_ms:.3f} ms, backprojection+update {embp_ms:.3f} ms")
    if not args.no_cpu:
        r["cpu_baseline"] = cpu_baseline_em(torch, dv, y[:1].contiguous(), ang)
    return r


def cpu_baseline_em(torch, dv, y1, ang, K=5):
    """eEM CPU-baseline leg: the FP32 restatement in oracle/ timed on one C3
    slice for one iteration, and the FP64 oracle as the checker of K GPU
    iterations (iterate and per-iteration KL) on the same slice."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    yh = y1.cpu().numpy()
    cores = host_cores()
    t0 = time.perf_counter()
    oracle.eem(yh, ang, 1, prec="f32", with_kl=False, nthreads=cores)
    dt = time.perf_counter() - t0
    xg, klg, exg = dv.eem_device(y1, ang, K)
    xg = xg.cpu().numpy()
    t0 = time.perf_counter()
    xr, klr, exr = oracle.eem(yh, ang, K, nthreads=cores)
    dto = time.perf_counter() - t0
    e = _err(xg, xr)
    r = dict(value=1.0 / dt, unit="slice-iterations/s", cores=cores, kind="port",
             sample="1 slice of the 16 (1024^2, 1024 angles), 1 iteration, FP32 C restatement",
             errors_vs_fp64_oracle=dict(slices=1, iterations=K, iterate=e,
                                        kl_max_rel=float(np.max(np.abs(klg - klr) / np.abs(klr))),
                                        excluded_equal=bool(np.array_equal(exg, exr))))
    log(f"EM CPU: {r['value']:.3f} slice-iterations/s; GPU vs FP64 oracle after {K} iterations: rel-L2 "
        f"{e['rel_l2']:.2e}, KL max rel {r['errors_vs_fp64_oracle']['kl_max_rel']:.2e} ({dto:.1f} s)")
    return r


def c1_bench(args, torch, dv, ph):
    """One 512^2 slice, 512 angles, pad 2, ramp: the CPU-runnable case, timed on
    the GPU (one call, data resident) and with the FP32 restatement on the host."""
    from sscRaft.geometries.parallel.fbp_methods import fbp_params

    N = A = 512
    ang = ph.uniform_angles(A)
    img = ph.phantoms(N, 1, config=1, device=torch.device("cuda:0"))
    y = ph.poisson_log_torch(dv.radon_device(img, ang), 1.0e4, seed=1000)
    pf, pi = fbp_params(N, A, 1, fbp_dic(N))
    out = torch.empty((1, N, N), dtype=torch.float32, device=y.device)
    for _ in range(3):
        dv.fbp_device(y, ang, pf, pi, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        dv.fbp_device(y, ang, pf, pi, out=out)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / 10
    r = {"config": "1 slice 512^2, 512 angles, pad 2, ramp, I0=1e4", "gpu_seconds": gpu_s}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        ys = y.cpu().numpy()
        cores = host_cores()
        t0 = time.perf_counter()
        oracle.fbp(ys, ang, pf, pi, prec="f32", nthreads=cores)
        r["cpu_seconds"] = time.perf_counter() - t0
        r["cpu_cores"] = cores
    log(f"C1: GPU {gpu_s * 1e3:.2f} ms per slice" + (f", CPU {r['cpu_seconds']:.3f} s on {r['cpu_cores']} threads"
                                                     if "cpu_seconds" in r else ""))
    return r


def batch_bench(args, torch, dv, ph):
    """256 independent scans, N from {256, 384, 512, 768, 1024} (angles = bins = N),
    I0 log-uniform in [1e3, 1e5]; FBP (pad 2, ramp, FP16 rows) + 200 eEM
    iterations of every scan in one captured graph launch."""
    import numpy as np

    rng = np.random.default_rng(4000)
    sizes = rng.choice([256, 384, 512, 768, 1024], size=args.batch_scans)
    i0s = 10.0 ** rng.uniform(3.0, 5.0, size=args.batch_scans)
    dev = torch.device("cuda:0")
    tomos, ys, angs, pfs, fouts, ems = [], [], [], [], [], []
    t0 = time.perf_counter()
    for k, n in enumerate(sizes):
        n = int(n)
        ang = ph.uniform_angles(n)
        img = ph.phantoms(n, 1, config=4, first=k, device=dev)
        p = ph.poisson_log_torch(dv.radon_device(img, ang), float(i0s[k]), seed=4000 + k)[0].contiguous()
        tomos.append(p)
        ys.append(torch.clamp(p, min=0.0).contiguous())
        angs.append(ang)
        pfs.append([0, 0, 1, 0, 2.0 / n, 2.0 / n, 0])
        fouts.append(torch.empty((n, n), dtype=torch.float32, device=dev))
        ems.append(torch.ones((n, n), dtype=torch.float32, device=dev))
    torch.cuda.synchronize()
    log(f"C4 inputs: {len(sizes)} scans generated in {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = dv.batch_device(tomos, angs, pfs, pad=2, filter_type=7, store="f16", fbp_out=fouts, ys=ys, em=ems,
                         iterations=args.em_iters)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    branches_ms = None
    if not args.no_batch_compare:
        # the same batch with every stage per bucket inside its graph branch
        ems2 = [torch.ones_like(e) for e in ems]
        os.environ["RAFT_BATCH_QUEUE"] = "0"
        try:
            branches_ms = dv.batch_device(tomos, angs, pfs, pad=2, filter_type=7, store="f16", fbp_out=fouts, ys=ys,
                                          em=ems2, iterations=args.em_iters)
        finally:
            del os.environ["RAFT_BATCH_QUEUE"]
        same = all(bool(torch.equal(a, b)) for a, b in zip(ems, ems2))
        log(f"C4 per-bucket branches: graph {branches_ms / 1e3:.2f} s (eEM outputs identical: {same})")
    n3 = float(np.sum(sizes.astype(np.float64) ** 3))
    r = {"config": f"{len(sizes)} scans, N in {{256,384,512,768,1024}}, angles = bins = N, I0 log-uniform "
                   f"[1e3, 1e5], FBP pad 2 ramp FP16 rows + {args.em_iters} eEM iterations, one graph launch",
         "scans": int(len(sizes)), "sizes_count": {int(v): int((sizes == v).sum()) for v in np.unique(sizes)},
         "mode": "persistent work queue: one launch per stage and iteration over all scans (graph-captured)",
         "graph_seconds": ms / 1e3, "wall_seconds": wall, "scans_per_s": len(sizes) / wall,
         "per_bucket_branches_graph_seconds": None if branches_ms is None else branches_ms / 1e3,
         "em_samples_per_s": 2.0 * n3 * args.em_iters / (ms / 1e3),
         "finite": bool(all(bool(torch.isfinite(f).all()) for f in fouts))}
    log(f"C4: {len(sizes)} scans in {wall:.2f} s (graph {ms / 1e3:.2f} s), eEM "
        f"{r['em_samples_per_s']:.3e} samples/s")
    return r


def unet_bench(args, torch):
    """U-Net forward and training step on 512^2 images (random weights and data):
    FLOP of the 3x3 convolutions / time, against dense BF16 MFMA."""
    from sscRaft.nn import UNet

    net = UNet(seed=0)
    r = {}
    tuned = net.autotune(32, 512, 512)
    net.autotune(8, 512, 512, backward=True)
    r["tuned_variants_b32"] = {(f"{k[0]}@{k[2]}" if isinstance(k[0], str) else f"{k[3]}->{k[4]}@{k[1]}"): v[0]
                               for k, v in tuned.items()}
    for B in (8, 32):
        x = torch.rand((B, 512, 512), device="cuda")
        out = torch.empty_like(x)
        for _ in range(3):
            net.forward(x, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            net.forward(x, out=out)
        torch.cuda.synchronize()
        tf = (time.perf_counter() - t0) / 10
        r[f"forward_b{B}_ms"] = tf * 1e3
        r[f"forward_b{B}_tflops"] = net.flops_forward(B, 512, 512) / tf / 1e12
    x = torch.rand((8, 512, 512), device="cuda")
    t = torch.rand((8, 512, 512), device="cuda")
    for _ in range(2):
        net.train_step(x, t, lr=1e-5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        net.train_step(x, t, lr=1e-5)
    torch.cuda.synchronize()
    tt = (time.perf_counter() - t0) / 5
    r["train_step_b8_ms"] = tt * 1e3
    r["train_step_b8_tflops"] = 3 * net.flops_forward(8, 512, 512, fused=False) / tt / 1e12
    r["roofline"] = {"bound": "mfma", "achieved": r["forward_b32_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": r["forward_b32_tflops"] / 2500.0,
                     "note": "fused inference forward pass, batch 32, FLOP = 2*9*Cin*Cout per output pixel over the "
                             "14 layers (first layer Cin = 1); training counts 3x the padded-input forward"}
    log(f"U-Net 512^2: forward {r['forward_b32_tflops']:.0f} TFLOP/s (batch 32), training step "
        f"{r['train_step_b8_tflops']:.0f} TFLOP/s (batch 8)")
    return r


def learned_bench(args):
    """Time-to-KL of the U-Net warm start vs plain eEM (sscRaft.learned.evaluate)."""
    from sscRaft import learned

    path = learned.default_weights_path()
    if not os.path.exists(path):
        log(f"learned warm start: no trained weights at {path}")
        return {"status": "no trained weights", "weights": os.path.relpath(path, ROOT)}
    from sscRaft.nn import UNet

    net = UNet.load(path)
    r = learned.evaluate(net, count=args.learned_slices, log=log)
    r["weights"] = os.path.relpath(path, ROOT)
    r["config"] = "held-out 512^2 slices, 128 angles x 512 bins, I0=1e4; target = lowest KL of plain eEM in 200 iterations"
    return r


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--learned-slices", type=int, default=512)
    ap.add_argument("--batch-scans", type=int, default=256)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-batch-compare", action="store_true",
                    help="skip the second C4 run with every stage per bucket (graph branches)")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--no-next", action="store_true", help="skip the low-pass / BST / flat-dark / tEM timings")
    ap.add_argument("--no-em", action="store_true")
    ap.add_argument("--em-iters", type=int, default=200)
    ap.add_argument("--only", choices=["all", "fbp", "em", "batch", "next"], default="all",
                    help="profiling runs: time only one workload (the JSON line keeps the same fields)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of the 64 x 2048 x 2048 sinogram (known byte count) so "
                         "that rocprofv3 FETCH_SIZE / WRITE_SIZE can be scaled")
    args = ap.parse_args()
    if args.only == "em":
        args.no_cpu = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the slab sharding over several GPUs is "
              f"not exercised by this benchmark.", file=sys.stderr)
        sys.exit(2)
    import torch

    from sscRaft import device as dv
    from sscRaft import phantom as ph

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    global LDS_RATE
    LDS_RATE = dv.lds_read_probe(0)
    log(f"LDS read probe: {LDS_RATE['bytes_per_s'] / 1e12:.1f} TB/s "
        f"({LDS_RATE['bytes_per_clk_per_cu']:.1f} B/clk/CU)")
    if args.only == "batch":
        print(json.dumps({"metric": METRIC, "batch": batch_bench(args, torch, dv, ph)}), flush=True)
        return
    if args.only == "next":
        print(json.dumps({"metric": METRIC, "next_rows": next_rows_bench(args, torch, dv, ph)}), flush=True)
        return
    if args.only == "em":
        em = em_bench(args, torch, dv, ph)
        if not args.no_learned:
            em["learned_init"] = learned_bench(args)
        print(json.dumps({"metric": METRIC, "em": em}), flush=True)
        return
    res, roof, cpu, radon, pcie = fbp_bench(args, torch, dv, ph)
    em = None if (args.no_em or args.only == "fbp") else em_bench(args, torch, dv, ph)
    if em is not None and not args.no_learned:
        em["learned_init"] = learned_bench(args)
    c1 = None if args.only == "fbp" else c1_bench(args, torch, dv, ph)
    nxt = None if (args.only == "fbp" or args.no_next) else next_rows_bench(args, torch, dv, ph)
    unet = None if args.only == "fbp" else unet_bench(args, torch)
    batch = None if (args.no_batch or args.only == "fbp") else batch_bench(args, torch, dv, ph)
    r32 = res["f32"]
    line = {
        "metric": METRIC,
        "value": r32["slices_per_s"],
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-ellipse phantoms, Radon-RT, Poisson I0=1e4)",
        "config": {"workload": "FBP-RT 64 slices x 2048 angles x 2048 bins -> 64 x 2048^2, pad 2, ramp, "
                               "pixel 2/N, FP32 rows", "slices": 64, "nrays": 2048, "nangles": 2048, "padding": 2,
                   "filter": "ramp"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp16_rows": {"value": res["f16"]["slices_per_s"], "unit": "slices/s", "bp_ms": res["f16"]["bp_ms"],
                      "filter_ms": res["f16"]["filter_ms"]},
        "filter_ms": r32["filter_ms"],
        "host_pointers_pcie": pcie,
        "radon": radon,
        "em": em,
        "c1": c1,
        "next_rows": nxt,
        "unet": unet,
        "batch": batch,
        "autotune": dv.autotune_choices(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

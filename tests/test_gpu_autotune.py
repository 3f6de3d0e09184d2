"""The tile autotuner on the GPU: the first call of a geometry times every
candidate (backprojector angle chunks 32/16/48 and, for FP32 rows, the pair
form that takes two slice groups per workgroup at 32 or 16 angles per round
and the scalar-prefetch sweep at 16 or 32; staged-projector ray-chunk counts) and keeps the fastest; every candidate
computes the same values, so a choice changes speed only -- checked here by
forcing each candidate through the cache file and comparing outputs bit for
bit.  The slice counts leave the last pair of slice groups without its second
group (FBP) or with a partial second group (eEM)."""
import json

import numpy as np
import pytest

import oracle
from sscRaft import phantom as ph

pytestmark = pytest.mark.gpu

PAIR = 1024  # BP_PAIR: the pair-form flag of a backprojector variant
SPF = 2048  # BP_SPF: the scalar-prefetch sweep flag
BP_VARIANTS = [16, 32, 48, PAIR | 16, PAIR | 32, SPF | 16, SPF | 32]


def _force(tmp_path, monkeypatch, entries):
    from sscRaft import device as dv

    f = tmp_path / "tune.json"
    f.write_text(json.dumps({k: {"choice": v, "ms": []} for k, v in entries.items()}))
    monkeypatch.setenv("RAFT_AUTOTUNE_CACHE", str(f))
    dv.autotune_clear()


def test_tuner_times_candidates_and_every_chunk_agrees(gpu, tmp_path, monkeypatch):
    import torch
    from sscRaft import device as dv

    n, a, s = 200, 180, 10
    y, ang = ph.poisson_log(oracle.radon(ph.phantoms(n, s, config=7), ph.uniform_angles(a)), 1e4, seed=5), \
        ph.uniform_angles(a)
    paramf, parami = oracle.fbp_params(n, a, s, pad=2, filter_type=7, pixel=2.0 / n)
    t = torch.from_numpy(y).cuda()
    monkeypatch.setenv("RAFT_AUTOTUNE_CACHE", str(tmp_path / "fresh.json"))
    dv.autotune_clear()
    ref = dv.fbp_device(t, ang, paramf, parami).cpu().numpy()
    choices = dv.autotune_choices()
    keys = [k for k in choices if k.startswith(f"bp/n={n}/A={a}/")]
    assert len(keys) == 1, choices
    rec = choices[keys[0]]
    assert sorted(c for c, _ in rec["ms"]) == BP_VARIANTS and rec["choice"] in BP_VARIANTS
    assert json.loads((tmp_path / "fresh.json").read_text())[keys[0]]["choice"] == rec["choice"]
    print(f"tuned {keys[0]}: {rec}")
    for chunk in BP_VARIANTS:
        _force(tmp_path, monkeypatch, {keys[0]: chunk})
        out = dv.fbp_device(t, ang, paramf, parami).cpu().numpy()
        np.testing.assert_array_equal(out, ref)
    pf, pi = paramf, parami
    assert np.linalg.norm(ref - oracle.fbp(y, ang, pf, pi)) / np.linalg.norm(oracle.fbp(y, ang, pf, pi)) < 1e-5


BP_VARIANTS_F16 = [16, 32, 48, SPF | 16, SPF | 32]


def test_fp16_rows_variants_agree(gpu, tmp_path, monkeypatch):
    """FBP with FP16-stored rows: the tuner times the angle chunks and the
    scalar-prefetch sweep, and every forced variant gives the same image bit
    for bit."""
    import torch
    from sscRaft import device as dv

    n, a, s = 160, 150, 9
    y, ang = ph.poisson_log(oracle.radon(ph.phantoms(n, s, config=7), ph.uniform_angles(a)), 1e4, seed=6), \
        ph.uniform_angles(a)
    paramf, parami = oracle.fbp_params(n, a, s, pad=2, filter_type=7, pixel=2.0 / n)
    t = torch.from_numpy(y).cuda()
    monkeypatch.setenv("RAFT_AUTOTUNE_CACHE", str(tmp_path / "fresh16.json"))
    dv.autotune_clear()
    ref = dv.fbp_device(t, ang, paramf, parami, store="f16").cpu().numpy()
    choices = dv.autotune_choices()
    keys = [k for k in choices if k.startswith(f"bp/n={n}/A={a}/") and "/store=1/" in k]
    assert len(keys) == 1, choices
    rec = choices[keys[0]]
    assert sorted(c for c, _ in rec["ms"]) == BP_VARIANTS_F16 and rec["choice"] in BP_VARIANTS_F16
    print(f"tuned {keys[0]}: {rec}")
    for chunk in BP_VARIANTS_F16:
        _force(tmp_path, monkeypatch, {keys[0]: chunk})
        out = dv.fbp_device(t, ang, paramf, parami, store="f16").cpu().numpy()
        np.testing.assert_array_equal(out, ref)


def test_eem_and_radon_variants_agree(gpu, tmp_path, monkeypatch):
    """eEM iterates are bit-identical whichever backprojector chunk and
    projector ray-chunk count run (the KL differs only in summation order);
    Radon likewise for the ray-chunk count."""
    import torch
    from sscRaft import device as dv

    n, a, s, K = 96, 96, 5, 3
    img = ph.phantoms(n, s, config=7, first=20)
    ang = ph.uniform_angles(a)
    y = torch.from_numpy(np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=9), 0).astype(np.float32)).cuda()
    monkeypatch.setenv("RAFT_AUTOTUNE_CACHE", str(tmp_path / "fresh.json"))
    dv.autotune_clear()
    x0, kl0, _ = dv.eem_device(y, ang, K)
    p0 = dv.radon_device(torch.from_numpy(img).cuda(), ang)
    choices = dv.autotune_choices()
    bp_key = [k for k in choices if k.startswith(f"bp/n={n}/A={a}/") and "mode=1" in k]
    pr_key = [k for k in choices if k.startswith("proj/") and "em=1" in k]
    rd_key = [k for k in choices if k.startswith("proj/") and "em=0" in k]
    assert len(bp_key) == 1 and len(pr_key) == 1 and len(rd_key) == 1, choices
    rcs = [c for c, _ in choices[pr_key[0]]["ms"]]
    assert len(rcs) >= 2
    assert sorted(c for c, _ in choices[bp_key[0]]["ms"]) == BP_VARIANTS
    for chunk, rc in [(16, rcs[0]), (48, rcs[-1]), (32, rcs[1]), (PAIR | 32, rcs[0]), (PAIR | 16, rcs[1]),
                      (SPF | 16, rcs[0]), (SPF | 32, rcs[-1])]:
        _force(tmp_path, monkeypatch, {bp_key[0]: chunk, pr_key[0]: rc, rd_key[0]: rc})
        x1, kl1, _ = dv.eem_device(y, ang, K)
        assert torch.equal(x1, x0)
        np.testing.assert_allclose(kl1, kl0, rtol=1e-12, atol=0)
        assert torch.equal(dv.radon_device(torch.from_numpy(img).cuda(), ang), p0)

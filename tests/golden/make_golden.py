"""Writes the golden fixtures in tests/golden/ from the FP64 CPU oracle.

The reference ships no test vectors (it has no tests), so these fixtures are
produced by the project's oracle after it is pinned by tests/test_oracle.py
(exact-rational brute force, closed forms, an independent numpy
restatement).  They guard the oracle against regressions and give the GPU
tests a reference that needs no oracle run.  Regenerate with
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ssc-raft_amd"))

import oracle  # noqa: E402
from sscRaft import phantom as ph  # noqa: E402

CASES_FBP = [  # (name, n, angles, slices, pad, filter, pixel, offset)
    ("fbp_n33_a17_p2_ramp", 33, 17, 2, 2, 7, 2.0 / 33, 0.0),
    ("fbp_n48_a40_p1_hann_off", 48, 40, 1, 1, 5, 0.37, 1.5),
    ("fbp_n17_a20_p2_ramp_direct", 17, 20, 1, 2, 7, 1.0, 0.0),
    ("fbp_n40_a31_p0_none", 40, 31, 1, 0, 0, 1.0, 0.0),
]


def main():
    for name, n, a, s, pad, ft, px, off in CASES_FBP:
        img = ph.phantoms(n, s, config=5, first=n)
        ang = ph.uniform_angles(a)
        y = ph.poisson_log(oracle.radon(img, ang), 1e4, seed=n + a)
        paramf, parami = oracle.fbp_params(n, a, s, pad=pad, filter_type=ft, pixel=px, offset=off)
        out = oracle.fbp(y, ang, paramf, parami, nthreads=1)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), tomo=y, angles=ang, paramf=paramf, parami=parami,
                            obj=out)
    for name, ny, nx, a in [("radon_24x24_a19", 24, 24, 19), ("radon_21x33_a12", 21, 33, 12)]:
        img = ph.phantoms(nx, 2, config=6, first=ny)[:, :ny, :]
        img = np.ascontiguousarray(img)
        ang = np.sort(np.random.default_rng(ny).uniform(0, 2 * np.pi, a)).astype(np.float32)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), phantom=img, angles=ang,
                            proj=oracle.radon(img, ang, nthreads=1))
    n = a = 32
    img = ph.phantoms(n, 2, config=7, first=0)
    ang = ph.uniform_angles(a)
    y = np.maximum(ph.poisson_log(oracle.radon(img, ang), 1e4, seed=77), 0).astype(np.float32)
    out = {}
    for k in (1, 5, 20):
        x, kl, ex = oracle.eem(y, ang, k, nthreads=1)
        out[f"obj_k{k}"] = x
        out[f"kl_k{k}"] = kl
        out[f"ex_k{k}"] = ex
    np.savez_compressed(os.path.join(HERE, "eem_n32_a32.npz"), tomo=y, angles=ang, **out)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()

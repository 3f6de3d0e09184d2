"""Pins the CPU float64 U-Net reference (oracle/unet_ref.py) that the GPU
kernels are checked against: convolution semantics and tap order against a
brute-force loop, pooling / upsampling on hand cases, bf16 rounding on known
values, and the autograd gradient of the whole network against central finite
differences."""
import numpy as np
import pytest
import torch

import unet_ref as U


def brute_conv(x, wflat, b):
    """y[co][h][w] = b[co] + sum_{dy,dx,ci} x[ci][h+dy][w+dx] W[co][3(dy+1)+(dx+1)][ci], zero outside."""
    ci_n, H, W = x.shape
    co_n = wflat.shape[0]
    y = np.zeros((co_n, H, W))
    for co in range(co_n):
        for h in range(H):
            for w in range(W):
                s = b[co]
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if 0 <= h + dy < H and 0 <= w + dx < W:
                            s += float(np.dot(x[:, h + dy, w + dx], wflat[co, 3 * (dy + 1) + (dx + 1)]))
                y[co, h, w] = s
    return y


def test_conv3x3_tap_order_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6))
    wflat = rng.standard_normal((3, 9, 2))
    b = rng.standard_normal(3)
    w4 = torch.tensor(wflat).reshape(3, 3, 3, 2).permute(0, 3, 1, 2)
    y = U.conv3x3(torch.tensor(x)[None], w4, torch.tensor(b))[0].numpy()
    np.testing.assert_allclose(y, brute_conv(x, wflat, b), rtol=1e-12, atol=1e-12)


def test_unpack_layout_matches_flat_order():
    base = 32
    n = sum(co * 9 * ci + co for _, co, ci in U.layer_table(base)) + base + 1
    theta = torch.arange(n, dtype=torch.float64)
    p = U.unpack(theta, base)
    w, b = p["enc1a"]
    # W[co=1][tap=4 (dy=dx=0)][ci=2] sits at (1*9 + 4)*32 + 2
    assert w[1, 2, 1, 1].item() == (1 * 9 + 4) * 32 + 2
    assert b[0].item() == 32 * 9 * 32
    hw, hb = p["head"]
    assert hb.item() == n - 1 and hw[0].item() == n - 1 - base


def test_pool_and_upsample_hand_cases():
    x = torch.arange(16, dtype=torch.float64).reshape(1, 1, 4, 4)
    np.testing.assert_array_equal(U.pool2(x)[0, 0].numpy(), [[2.5, 4.5], [10.5, 12.5]])
    u = U.up2(torch.tensor([[[[1.0, 2.0], [3.0, 4.0]]]], dtype=torch.float64))[0, 0].numpy()
    np.testing.assert_array_equal(u, [[1, 1, 2, 2], [1, 1, 2, 2], [3, 3, 4, 4], [3, 3, 4, 4]])


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -2.5, 0.0, 3.0e38], np.float64)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6); 1+2^-9 rounds down
    exp = np.array([1.0, 1.0, 1 + 2 ** -6, 1.0, -2.5, 0.0, U.bf16_value(U.bf16_bits(3.0e38))])
    np.testing.assert_array_equal(U.round_bf16(torch.tensor(vals)).numpy(), exp)
    np.testing.assert_array_equal(U.bf16_value(U.bf16_bits(vals)), exp.astype(np.float32))


def test_layer_table_matches_the_package_layout():
    from sscRaft import unet_layout as SP

    assert U.layer_table() == SP.layer_table()
    off, n = SP.param_layout()
    assert n == sum(co * 9 * ci + co for _, co, ci in U.layer_table()) + 32 + 8


def test_gradient_matches_finite_differences():
    from sscRaft import unet_layout as SP

    theta = SP.init_params(seed=3)
    rng = np.random.default_rng(1)
    # larger head weights so every layer gets a sizeable gradient
    off, _ = SP.param_layout()
    h0, hb = off["head"][0], off["head"][1]
    theta[h0:hb] = rng.standard_normal(hb - h0) * 0.3
    img = rng.random((1, 16, 16))
    tgt = rng.random((1, 16, 16))
    loss, grad, _ = U.loss_and_grad(theta, img, tgt, bf16_weights=False)
    idx = [off[name][0] + k for name in ("enc1a", "enc2b", "bot_b", "dec3a", "dec1b")
           for k in rng.choice(off[name][1] - off[name][0], 3, replace=False)]
    idx += [off[name][1] for name in ("enc1b", "dec2a")] + [h0 + 3, hb]
    th = torch.tensor(theta.astype(np.float64))
    eps = 1e-5
    for i in idx:
        tp, tm = th.clone(), th.clone()
        tp[i] += eps
        tm[i] -= eps
        fp = ((U.forward(tp, torch.tensor(img), bf16_weights=False) - torch.tensor(tgt)) ** 2).mean().item()
        fm = ((U.forward(tm, torch.tensor(img), bf16_weights=False) - torch.tensor(tgt)) ** 2).mean().item()
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-4 * abs(fd), (i, fd, grad[i])

"""CPU float64 reference of the warm start U-Net (test infrastructure only).

Restates, with torch CPU float64 ops and autograd, the network the HIP kernels
of csrc/nn.hip run (structure in sscRaft/nn.py): the same flat parameter
vector (conv W [Cout][9][Cin] with tap = 3 (dy+1) + (dx+1), then b [Cout];
head w [32], b [1]), the same zero-padded 32-channel input, 2x2 average
pooling, nearest 2x upsampling, [encoder | upsampled] concatenation, ReLU
after every 3x3 convolution, 1x1 head, residual, mean squared error.

Pinned in tests/test_unet_ref.py by finite differences (gradients), a
brute-force loop (3x3 convolution semantics and tap order) and hand-computed
pooling / upsampling cases.  The reference ssc-raft has no learned part, so
nothing here cites it.
"""
import numpy as np
import torch
import torch.nn.functional as F

def layer_table(base=32):
    c1, c2, c3, c4 = base, 2 * base, 4 * base, 8 * base
    return [("enc1a", c1, 32), ("enc1b", c1, c1), ("enc2a", c2, c1), ("enc2b", c2, c2), ("enc3a", c3, c2),
            ("enc3b", c3, c3), ("bot_a", c4, c3), ("bot_b", c4, c4), ("dec3a", c3, c3 + c4), ("dec3b", c3, c3),
            ("dec2a", c2, c2 + c3), ("dec2b", c2, c2), ("dec1a", c1, c1 + c2), ("dec1b", c1, c1)]


def unpack(theta, base=32):
    """Views of the flat parameter tensor: {name: (W4 [Cout][Cin][3][3], b)}, head (w, b)."""
    out = {}
    pos = 0
    for name, co, ci in layer_table(base):
        w = theta[pos:pos + co * 9 * ci].reshape(co, 3, 3, ci).permute(0, 3, 1, 2)
        pos += co * 9 * ci
        out[name] = (w, theta[pos:pos + co])
        pos += co
    out["head"] = (theta[pos:pos + base], theta[pos + base:pos + base + 1])
    return out


def conv3x3(x, w4, b):
    """y[n][co][h][w] = b[co] + sum_{ci,ky,kx} x[n][ci][h+ky-1][w+kx-1] w4[co][ci][ky][kx], zero padded."""
    return F.conv2d(x, w4, b, padding=1)


def pool2(x):
    return F.avg_pool2d(x, 2)


def up2(x):
    return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)


def forward(theta, img, base=32, bf16_weights=True):
    """Prediction [n, h, w] for img [n, h, w] (torch float64, CPU)."""
    p = unpack(theta, base)
    if bf16_weights:
        p = {k: (round_bf16(w), b) if k != "head" else (w, b) for k, (w, b) in p.items()}
    n, h, w = img.shape
    x0 = torch.zeros((n, 32, h, w), dtype=img.dtype)
    x0[:, 0] = round_bf16(img)
    r = F.relu

    def c(name, x):
        return r(conv3x3(x, *p[name]))

    e1 = c("enc1b", c("enc1a", x0))
    e2 = c("enc2b", c("enc2a", pool2(e1)))
    e3 = c("enc3b", c("enc3a", pool2(e2)))
    e4 = c("bot_b", c("bot_a", pool2(e3)))
    d3 = c("dec3b", c("dec3a", torch.cat([e3, up2(e4)], 1)))
    d2 = c("dec2b", c("dec2a", torch.cat([e2, up2(d3)], 1)))
    d1 = c("dec1b", c("dec1a", torch.cat([e1, up2(d2)], 1)))
    hw, hb = p["head"]
    z = torch.einsum("nchw,c->nhw", d1, hw) + hb[0]
    return img + z


def round_bf16(t):
    """Round a float64 tensor holding float32 values to the nearest bf16 value
    (ties to even), as float64."""
    a = t.detach().to(torch.float32).numpy().astype(np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    r = torch.from_numpy(u.astype(np.uint32).view(np.float32).astype(np.float64))
    if t.requires_grad:
        return t + (r - t).detach()  # value of r, gradient of the identity
    return r


def loss_and_grad(theta_np, img_np, target_np, base=32, bf16_weights=True):
    """(mse, d mse / d theta, pred) in float64 on the CPU."""
    theta = torch.tensor(np.asarray(theta_np, np.float64), requires_grad=True)
    img = torch.tensor(np.asarray(img_np, np.float64))
    tgt = torch.tensor(np.asarray(target_np, np.float64))
    pred = forward(theta, img, base, bf16_weights)
    loss = ((pred - tgt) ** 2).mean()
    loss.backward()
    return float(loss.detach()), theta.grad.numpy().copy(), pred.detach().numpy()


def bf16_bits(a):
    """uint16 bf16 bit patterns (round to nearest even) of a float array."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)

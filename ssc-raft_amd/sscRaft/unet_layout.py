"""Parameter layout and initialisation of the warm start U-Net (no device
code; shared by sscRaft/nn.py and the tests).

Flat FP32 parameter vector, per 3x3 convolution in layer_table() order:
W [Cout][9][Cin] (tap = 3 (dy+1) + (dx+1)) then b [Cout]; then the 1x1 head
w [32], b [1] and 7 padding elements.
"""
import numpy as np


def layer_table(base=32):
    """[(name, Cout, Cin)] of the 3x3 convolutions, in parameter order."""
    c1, c2, c3, c4 = base, 2 * base, 4 * base, 8 * base
    return [("enc1a", c1, 32), ("enc1b", c1, c1), ("enc2a", c2, c1), ("enc2b", c2, c2), ("enc3a", c3, c2),
            ("enc3b", c3, c3), ("bot_a", c4, c3), ("bot_b", c4, c4), ("dec3a", c3, c3 + c4), ("dec3b", c3, c3),
            ("dec2a", c2, c2 + c3), ("dec2b", c2, c2), ("dec1a", c1, c1 + c2), ("dec1b", c1, c1)]


def param_layout(base=32):
    """Offsets (in elements) of every weight / bias in the flat parameter vector.

    conv layer: W [Cout][9][Cin] then b [Cout]; head: w [base] then b [1],
    padded to a multiple of 8 elements (16-byte aligned bf16 weights).
    """
    off = {}
    pos = 0
    for name, co, ci in layer_table(base):
        off[name] = (pos, pos + co * 9 * ci, co, ci)
        pos += co * 9 * ci + co
    off["head"] = (pos, pos + base, 1, base)
    pos += base + 8
    return off, pos


def init_params(seed=0, base=32):
    """He-normal conv weights (fan-in 9 Cin; the padded input layer counts one
    channel), zero biases, small head weights (the net starts near the identity)."""
    off, n = param_layout(base)
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    for name, co, ci in layer_table(base):
        w0, b0, _, _ = off[name]
        fan_in = 9 * (1 if name == "enc1a" else ci)
        w = rng.standard_normal((co, 9, ci)) * np.sqrt(2.0 / fan_in)
        if name == "enc1a":
            w[:, :, 1:] = 0.0
        p[w0:b0] = w.reshape(-1)
    h0, hb, _, _ = off["head"]
    p[h0:hb] = rng.standard_normal(base) * (0.01 / np.sqrt(base))
    return p

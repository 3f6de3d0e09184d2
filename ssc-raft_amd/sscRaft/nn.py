"""Host side of the learned eEM warm start network (kernels: csrc/nn.hip).

A 4-level 2-D U-Net, base 32 channels: two 3x3 convolutions + ReLU per level
(32, 64, 128, 256 channels), 2x2 average pooling down, nearest 2x upsampling
up, skip concatenation [encoder | upsampled], a 1x1 head and a residual
connection to the input image:  pred = img + head(net(img)).

All arithmetic runs in the HIP kernels of libraft.so (BF16 MFMA, FP32
accumulation, FP32 master weights, Adam); torch tensors only hold device
memory and provide the stream.  There is no CPU fallback: without a GPU and
libraft.so the constructor raises.

The reference (ssc-raft) has no learned components; this module serves the
warm start of its eEM path (get_eEM_RT_MultiGPU, sscRaft/em_methods.py:3-70).
"""
import ctypes

import numpy as np

from .rafttypes import libraft
from .unet_layout import init_params, layer_table, param_layout

__all__ = ["UNet", "layer_table", "param_layout", "init_params", "nn_call", "mfma_probe"]

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_ll = ctypes.c_longlong

_SIGNATURES = {
    "raft_nn_mfma_probe": ([_vp] * 7, _i),
    "raft_nn_conv3x3": ([_vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_conv3x3_ex": ([_vp, _i, _i, _i, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp,
                            _i, _i, _vp], _i),
    "raft_nn_conv_first": ([_vp, _f, _vp, _i, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_conv_first2": ([_vp, _f, _vp, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _i, _i, _vp],
                            _i),
    "raft_nn_conv3x3_head": ([_vp, _i, _i, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _f, _i, _vp], _i),
    "raft_nn_conv_weight_flip": ([_vp, _vp, _i, _i, _vp], _i),
    "raft_nn_wgrad_workspace": ([_i, _i, _i, _i, _i], _ll),
    "raft_nn_conv3x3_wgrad": ([_vp, _i, _i, _vp, _i, _i, _vp, _vp, _vp, _ll, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_relu_mask": ([_vp, _i, _i, _vp, _i, _i, _vp, _i, _i, _ll, _i, _vp], _i),
    "raft_nn_pool2": ([_vp, _i, _i, _vp, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_skip_grad": ([_vp, _i, _i, _vp, _i, _i, _vp, _i, _i, _vp, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_up2": ([_vp, _i, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _vp], _i),
    "raft_nn_input": ([_vp, _vp, _ll, _f, _vp], _i),
    "raft_nn_head_workspace": ([_ll], _ll),
    "raft_nn_head_loss_slots": ([_ll], _ll),
    "raft_nn_head": ([_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _ll, _f, _i, _i, _vp],
                     _i),
    "raft_nn_adam": ([_vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _i, _f, _vp], _i),
    "raft_nn_to_bf16": ([_vp, _vp, _ll, _vp], _i),
    "raft_nn_resample": ([_vp, _vp, _i, _i, _i, _i, _i, _d, _d, _d, _d, _vp], _i),
}
for _name, (_args, _res) in _SIGNATURES.items():
    _fn = getattr(libraft, _name)
    _fn.argtypes = _args
    _fn.restype = _res

_ERRORS = {1: ValueError, 2: RuntimeError, 3: RuntimeError, 4: NotImplementedError}


def nn_call(name, *args):
    """Call a raft_nn_* entry point; raise on a non-zero status."""
    rc = getattr(libraft, name)(*args)
    if rc != 0:
        raise _ERRORS.get(rc, RuntimeError)(f"{name}: status {rc}")
    return rc


def _stream(stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def mfma_probe(A16, B16, A32, B32):
    """Run the MFMA lane-layout probe; inputs are uint16 bf16 bit arrays."""
    import torch

    dev = torch.device("cuda")
    t = [torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(dev) for a in (A16, B16, A32, B32)]
    o16 = torch.zeros(64 * 4, dtype=torch.float32, device=dev)
    o32 = torch.zeros(64 * 16, dtype=torch.float32, device=dev)
    nn_call("raft_nn_mfma_probe", t[0].data_ptr(), t[1].data_ptr(), o16.data_ptr(), t[2].data_ptr(), t[3].data_ptr(),
            o32.data_ptr(), _stream(None))
    torch.cuda.synchronize()
    return o16.cpu().numpy().reshape(64, 4), o32.cpu().numpy().reshape(64, 16)


class _Act:
    """One NHWC bf16 activation buffer (int16 storage) of `ch` channels."""

    def __init__(self, torch, n, h, w, ch, dev):
        self.t = torch.zeros((n * h * w * ch,), dtype=torch.int16, device=dev)
        self.h, self.w, self.ch = h, w, ch

    def ptr(self, c0=0):
        return self.t.data_ptr() + 2 * c0


class UNet:
    """The warm start U-Net on one GPU (see module docstring)."""

    def __init__(self, params=None, seed=0, base=32, device="cuda"):
        import torch

        if base != 32:
            raise ValueError("the 1x1 head kernel is built for 32 channels")
        if not torch.cuda.is_available():
            raise RuntimeError("UNet needs a HIP device (no CPU fallback)")
        self.torch = torch
        self.dev = torch.device(device)
        self.base = base
        self.layers = layer_table(base)
        self.offsets, self.n = param_layout(base)
        p = init_params(seed, base) if params is None else np.asarray(params, np.float32)
        if p.shape != (self.n,):
            raise ValueError(f"expected {self.n} parameters, got {p.shape}")
        self.params = torch.from_numpy(p.copy()).to(self.dev)
        self.params_bf = torch.zeros(self.n, dtype=torch.int16, device=self.dev)
        self.grads = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.v = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.step_count = 0
        self.flip = {name: torch.zeros(co * 9 * ci, dtype=torch.int16, device=self.dev)
                     for name, co, ci in self.layers}
        self._bufs = None
        self._shape = None
        self._fbufs = None
        self._fshape = None
        self.tuned = {}  # (n, h, w, Cin, Cout) -> conv tile variant (see autotune)
        self.refresh()

    # ---------------------------------------------------------------- params
    def refresh(self, stream=None):
        """bf16 copies and backward-data (flipped) weights from the FP32 master."""
        st = _stream(stream)
        nn_call("raft_nn_to_bf16", self.params.data_ptr(), self.params_bf.data_ptr(), self.n, st)
        self._flip_all(st)

    def _flip_all(self, st):
        for name, co, ci in self.layers:
            w0 = self.offsets[name][0]
            nn_call("raft_nn_conv_weight_flip", self.params_bf.data_ptr() + 2 * w0, self.flip[name].data_ptr(), co, ci,
                    st)

    def numpy_params(self):
        self.torch.cuda.synchronize()
        return self.params.cpu().numpy().copy()

    def save(self, path, **meta):
        """Weights file: 3x3 conv weights as bf16 bit patterns (the convolutions
        only ever read their bf16 copy), biases and the head in FP32."""
        p = self.numpy_params()
        conv = np.zeros(self.n, dtype=bool)
        for name, co, ci in self.layers:
            w0, b0 = self.offsets[name][0], self.offsets[name][1]
            conv[w0:b0] = True
        u = p.view(np.uint32).astype(np.uint64)
        bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        np.savez_compressed(path, conv_bf16=bits[conv], other_f32=p[~conv], n=self.n, base=self.base,
                            adam_steps=self.step_count, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path, allow_pickle=False) as z:
            base = int(z["base"])
            if "params" in z.files:
                return cls(params=z["params"], base=base, device=device)
            offsets, n = param_layout(base)
            if int(z["n"]) != n:
                raise ValueError(f"{path}: {int(z['n'])} parameters, expected {n}")
            conv = np.zeros(n, dtype=bool)
            for name, co, ci in layer_table(base):
                conv[offsets[name][0]:offsets[name][1]] = True
            p = np.zeros(n, np.float32)
            p[conv] = (z["conv_bf16"].astype(np.uint32) << 16).view(np.float32)
            p[~conv] = z["other_f32"]
            return cls(params=p, base=base, device=device)

    # --------------------------------------------------------------- buffers
    def _buffers(self, n, h, w):
        if self._shape == (n, h, w):
            return self._bufs
        if h % 8 or w % 8:
            raise ValueError("image height and width must be multiples of 8 (three 2x2 poolings)")
        t, dev = self.torch, self.dev
        c1, c2, c3, c4 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        L = [(h, w), (h // 2, w // 2), (h // 4, w // 4), (h // 8, w // 8)]
        spec = {
            "x0": (0, 32), "a1": (0, c1), "cat1": (0, c1 + c2), "p1": (1, c1), "a2": (1, c2), "cat2": (1, c2 + c3),
            "p2": (2, c2), "a3": (2, c3), "cat3": (2, c3 + c4), "p3": (3, c3), "a4": (3, c4), "e4": (3, c4),
            "a5": (2, c3), "d3": (2, c3), "a6": (1, c2), "d2": (1, c2), "a7": (0, c1), "d1": (0, c1),
            # gradients
            "g_d1": (0, c1), "g_a7": (0, c1), "g_cat1": (0, c1 + c2), "g_d2": (1, c2), "g_a6": (1, c2),
            "g_cat2": (1, c2 + c3), "g_d3": (2, c3), "g_a5": (2, c3), "g_cat3": (2, c3 + c4), "g_e4": (3, c4),
            "g_a4": (3, c4), "g_p3": (3, c3), "g_e3": (2, c3), "g_a3": (2, c3), "g_p2": (2, c2), "g_e2": (1, c2),
            "g_a2": (1, c2), "g_p1": (1, c1), "g_e1": (0, c1), "g_a1": (0, c1),
        }
        self._bufs = None
        bufs = {k: _Act(t, n, L[lv][0], L[lv][1], ch, dev) for k, (lv, ch) in spec.items()}
        ws = 0
        for name, co, ci in self.layers:
            lv = {"enc1": 0, "dec1": 0, "enc2": 1, "dec2": 1, "enc3": 2, "dec3": 2, "bot_": 3}[name[:4]]
            ws = max(ws, libraft.raft_nn_wgrad_workspace(n, L[lv][0], L[lv][1], ci, co))
        P = n * h * w
        bufs["wg_work"] = t.zeros(int(ws), dtype=t.float32, device=dev)
        bufs["head_work"] = t.zeros(int(libraft.raft_nn_head_workspace(P)), dtype=t.float32, device=dev)
        bufs["loss"] = t.zeros(int(libraft.raft_nn_head_loss_slots(P)), dtype=t.float64, device=dev)
        bufs["levels"] = L
        self._bufs, self._shape = bufs, (n, h, w)
        return bufs

    # ------------------------------------------------------------ primitives
    def _conv(self, name, x, cx0, y, cy0, n, hw, st, relu=True):
        co, ci = self.offsets[name][2], self.offsets[name][3]
        w0, b0 = self.offsets[name][0], self.offsets[name][1]
        v = self.tuned.get((n, hw[0], hw[1], ci, co), 0)
        nn_call("raft_nn_conv3x3", x.ptr(), x.ch, cx0, self.params_bf.data_ptr() + 2 * w0,
                self.params.data_ptr() + 4 * b0, y.ptr(), y.ch, cy0, n, hw[0], hw[1], ci, co, int(relu), v, st)

    POOL_VARIANTS = (6, 7, 8, 9, 10, 11, 12, 13)  # persistent forms (pooled output in the epilogue)

    def _conv_ex(self, name, x, pool, x2, y, n, hw, st, relu=True, yp=None):
        """Fused-input convolution (inference): x pooled 2x2 when `pool`, and
        input channels past x's own read from x2 upsampled 2x; yp (optional)
        receives the 2x2 average of the output."""
        co, ci = self.offsets[name][2], self.offsets[name][3]
        w0, b0 = self.offsets[name][0], self.offsets[name][1]
        v = self.tuned.get((n, hw[0], hw[1], ci, co), 0)
        if yp is not None and v not in self.POOL_VARIANTS:
            v = 0
        split = ci if x2 is None else x.ch
        nn_call("raft_nn_conv3x3_ex", x.ptr(), x.ch, 0, int(pool), None if x2 is None else x2.ptr(),
                0 if x2 is None else x2.ch, 0, split, self.params_bf.data_ptr() + 2 * w0,
                self.params.data_ptr() + 4 * b0, y.ptr(), y.ch, 0, n, hw[0], hw[1], ci, co, int(relu), v,
                None if yp is None else yp.ptr(), 0 if yp is None else yp.ch, 0, st)

    def _conv_bwd_data(self, name, g, gx, cgx0, n, hw, st):
        co, ci = self.offsets[name][2], self.offsets[name][3]
        v = self.tuned.get((n, hw[0], hw[1], co, ci), 0)
        nn_call("raft_nn_conv3x3", g.ptr(), g.ch, 0, self.flip[name].data_ptr(), None, gx.ptr(), gx.ch, cgx0, n, hw[0],
                hw[1], co, ci, 0, v, st)

    # ---------------------------------------------------------------- tuning
    CONV_VARIANTS = (1, 2, 3, 5, 6, 7, 9, 10, 11, 13)  # the 16x64 forms (8, 12) measured slower on every shape
    HEAD_VARIANTS = (6, 9, 10, 13)  # persistent forms with one 32-channel output tile (head / first-layer modes)

    def autotune(self, n, h, w, backward=False, repeats=5, stream=None):
        """Pick, per convolution shape of this network at batch n and image h x w,
        the fastest tile variant of raft_nn_conv3x3 by timing each candidate
        (HIP events on the launch stream); the choice is kept in self.tuned and
        used by every later call with that shape.  The fused first two layers and
        the fused last layer + head are timed on their own (keys ("first2", n, h, w)
        and ("head", n, h, w)).  Returns {key: (variant, us)}."""
        t = self.torch
        st = _stream(stream)
        cur = t.cuda.current_stream() if stream is None else stream
        L = [(h, w), (h // 2, w // 2), (h // 4, w // 4), (h // 8, w // 8)]
        lv = {"enc1": 0, "dec1": 0, "enc2": 1, "dec2": 1, "enc3": 2, "dec3": 2, "bot_": 3}
        shapes = set()
        for name, co, ci in self.layers:
            hh, ww = L[lv[name[:4]]]
            shapes.add((n, hh, ww, ci, co))
            if backward and name != "enc1a":
                shapes.add((n, hh, ww, co, ci))
        result = {}
        for (nb, hh, ww, ci, co) in sorted(shapes):
            x = t.zeros(nb * hh * ww * ci, dtype=t.int16, device=self.dev)
            wt = t.zeros(co * 9 * ci, dtype=t.int16, device=self.dev)
            y = t.zeros(nb * hh * ww * co, dtype=t.int16, device=self.dev)
            best = None
            for v in self.CONV_VARIANTS:
                if v in (1, 4, 7, 8, 11, 12) and co % 64:
                    continue
                args = (x.data_ptr(), ci, 0, wt.data_ptr(), None, y.data_ptr(), co, 0, nb, hh, ww, ci, co, 1, v, st)
                nn_call("raft_nn_conv3x3", *args)
                e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                e0.record(cur)
                for _ in range(repeats):
                    nn_call("raft_nn_conv3x3", *args)
                e1.record(cur)
                e1.synchronize()
                us = 1e3 * e0.elapsed_time(e1) / repeats
                if best is None or us < best[1]:
                    best = (v, us)
            self.tuned[(nb, hh, ww, ci, co)] = best[0]
            result[(nb, hh, ww, ci, co)] = best
        if h % 2 == 0 and w % 2 == 0:
            # the fused first two layers and the fused last layer + head, timed on
            # their own (their variants are not the plain convolution's best)
            c = self.base
            img = t.zeros((n, h, w), dtype=t.float32, device=self.dev)
            wt = t.zeros(c * 9 * c, dtype=t.int16, device=self.dev)
            y = t.zeros(n * h * w * c, dtype=t.int16, device=self.dev)
            yp = t.zeros(n * (h // 2) * (w // 2) * c, dtype=t.int16, device=self.dev)
            hw_ = t.zeros(c + 1, dtype=t.float32, device=self.dev)
            pred = t.zeros((n, h, w), dtype=t.float32, device=self.dev)
            calls = {
                "first2": lambda v: nn_call("raft_nn_conv_first2", img.data_ptr(), 1.0, wt.data_ptr(), c, None, 1,
                                            wt.data_ptr(), None, y.data_ptr(), c, 0, n, h, w, 1, v, yp.data_ptr(), c,
                                            0, st),
                "head": lambda v: nn_call("raft_nn_conv3x3_head", y.data_ptr(), c, 0, wt.data_ptr(), None, n, h, w, c,
                                          1, v, hw_.data_ptr(), hw_.data_ptr() + 4 * c, img.data_ptr(),
                                          pred.data_ptr(), 0.0, 1, st)}
            for kind, call in calls.items():
                best = None
                for v in self.HEAD_VARIANTS:
                    call(v)
                    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                    e0.record(cur)
                    for _ in range(repeats):
                        call(v)
                    e1.record(cur)
                    e1.synchronize()
                    us = 1e3 * e0.elapsed_time(e1) / repeats
                    if best is None or us < best[1]:
                        best = (v, us)
                self.tuned[(kind, n, h, w)] = best[0]
                result[(kind, n, h, w)] = best
        return result

    def _wgrad(self, name, x, cx0, g, n, hw, st):
        co, ci = self.offsets[name][2], self.offsets[name][3]
        w0, b0 = self.offsets[name][0], self.offsets[name][1]
        work = self._bufs["wg_work"]
        nn_call("raft_nn_conv3x3_wgrad", x.ptr(), x.ch, cx0, g.ptr(), g.ch, 0, self.grads.data_ptr() + 4 * w0,
                self.grads.data_ptr() + 4 * b0, work.data_ptr(), work.numel(), n, hw[0], hw[1], ci, co, 0, st)

    def _mask(self, g, act, ca0, n_pix, st):
        nn_call("raft_nn_relu_mask", g.ptr(), g.ch, 0, act.ptr(), act.ch, ca0, g.ptr(), g.ch, 0, n_pix, g.ch, st)

    # --------------------------------------------------------------- forward
    def _forward(self, img, n, h, w, st):
        b = self._buffers(n, h, w)
        L = b["levels"]
        c1, c2, c3 = self.base, 2 * self.base, 4 * self.base
        nn_call("raft_nn_input", img.data_ptr(), b["x0"].ptr(), n * h * w, 1.0, st)
        self._conv("enc1a", b["x0"], 0, b["a1"], 0, n, L[0], st)
        self._conv("enc1b", b["a1"], 0, b["cat1"], 0, n, L[0], st)
        nn_call("raft_nn_pool2", b["cat1"].ptr(), b["cat1"].ch, 0, b["p1"].ptr(), c1, 0, n, L[1][0], L[1][1], c1, st)
        self._conv("enc2a", b["p1"], 0, b["a2"], 0, n, L[1], st)
        self._conv("enc2b", b["a2"], 0, b["cat2"], 0, n, L[1], st)
        nn_call("raft_nn_pool2", b["cat2"].ptr(), b["cat2"].ch, 0, b["p2"].ptr(), c2, 0, n, L[2][0], L[2][1], c2, st)
        self._conv("enc3a", b["p2"], 0, b["a3"], 0, n, L[2], st)
        self._conv("enc3b", b["a3"], 0, b["cat3"], 0, n, L[2], st)
        nn_call("raft_nn_pool2", b["cat3"].ptr(), b["cat3"].ch, 0, b["p3"].ptr(), c3, 0, n, L[3][0], L[3][1], c3, st)
        self._conv("bot_a", b["p3"], 0, b["a4"], 0, n, L[3], st)
        self._conv("bot_b", b["a4"], 0, b["e4"], 0, n, L[3], st)
        nn_call("raft_nn_up2", b["e4"].ptr(), b["e4"].ch, 0, b["cat3"].ptr(), b["cat3"].ch, c3, n, L[3][0], L[3][1],
                b["e4"].ch, 0, st)
        self._conv("dec3a", b["cat3"], 0, b["a5"], 0, n, L[2], st)
        self._conv("dec3b", b["a5"], 0, b["d3"], 0, n, L[2], st)
        nn_call("raft_nn_up2", b["d3"].ptr(), c3, 0, b["cat2"].ptr(), b["cat2"].ch, c2, n, L[2][0], L[2][1], c3, 0, st)
        self._conv("dec2a", b["cat2"], 0, b["a6"], 0, n, L[1], st)
        self._conv("dec2b", b["a6"], 0, b["d2"], 0, n, L[1], st)
        nn_call("raft_nn_up2", b["d2"].ptr(), c2, 0, b["cat1"].ptr(), b["cat1"].ch, c1, n, L[1][0], L[1][1], c2, 0, st)
        self._conv("dec1a", b["cat1"], 0, b["a7"], 0, n, L[0], st)
        self._conv("dec1b", b["a7"], 0, b["d1"], 0, n, L[0], st)
        return b

    def _forward_fused(self, img, n, h, w, st, out, floor):
        """Inference forward with the first layer computed inside the second
        one's halo load, each level's 2x2 average written by its last
        convolution's epilogue, upsampling and skip concatenations folded into
        the convolutions' input loads, and the head into the last
        convolution's epilogue: the same values as _forward followed by the
        head, up to the order in which the first layer's 9 exact products are
        summed (pooling / upsampling / head are bit-exact).  Writes the
        prediction to out."""
        b = self._fused_buffers(n, h, w)
        L = [(h, w), (h // 2, w // 2), (h // 4, w // 4), (h // 8, w // 8)]
        w0, b0 = self.offsets["enc1a"][0], self.offsets["enc1a"][1]
        w1, b1 = self.offsets["enc1b"][0], self.offsets["enc1b"][1]
        v = self.tuned.get(("first2", n, h, w), self.tuned.get((n, h, w, self.base, self.base), 0))
        nn_call("raft_nn_conv_first2", img.data_ptr(), 1.0, self.params_bf.data_ptr() + 2 * w0, 32,
                self.params.data_ptr() + 4 * b0, 1, self.params_bf.data_ptr() + 2 * w1, self.params.data_ptr() + 4 * b1,
                b["e1"].ptr(), b["e1"].ch, 0, n, h, w, 1, v if v in self.HEAD_VARIANTS else 0, b["p1"].ptr(),
                b["p1"].ch, 0, st)
        self._conv_ex("enc2a", b["p1"], 0, None, b["a2"], n, L[1], st)
        self._conv_ex("enc2b", b["a2"], 0, None, b["e2"], n, L[1], st, yp=b["p2"])
        self._conv_ex("enc3a", b["p2"], 0, None, b["a3"], n, L[2], st)
        self._conv_ex("enc3b", b["a3"], 0, None, b["e3"], n, L[2], st, yp=b["p3"])
        self._conv_ex("bot_a", b["p3"], 0, None, b["a4"], n, L[3], st)
        self._conv_ex("bot_b", b["a4"], 0, None, b["e4"], n, L[3], st)
        self._conv_ex("dec3a", b["e3"], 0, b["e4"], b["a5"], n, L[2], st)
        self._conv_ex("dec3b", b["a5"], 0, None, b["d3"], n, L[2], st)
        self._conv_ex("dec2a", b["e2"], 0, b["d3"], b["a6"], n, L[1], st)
        self._conv_ex("dec2b", b["a6"], 0, None, b["d2"], n, L[1], st)
        self._conv_ex("dec1a", b["e1"], 0, b["d2"], b["a7"], n, L[0], st)
        co, ci = self.offsets["dec1b"][2], self.offsets["dec1b"][3]
        w1, b1 = self.offsets["dec1b"][0], self.offsets["dec1b"][1]
        h0, hb = self.offsets["head"][0], self.offsets["head"][1]
        v = self.tuned.get(("head", n, h, w), self.tuned.get((n, h, w, ci, co), 0))
        nn_call("raft_nn_conv3x3_head", b["a7"].ptr(), b["a7"].ch, 0, self.params_bf.data_ptr() + 2 * w1,
                self.params.data_ptr() + 4 * b1, n, h, w, ci, 1, v if v in self.HEAD_VARIANTS else 0,
                self.params.data_ptr() + 4 * h0, self.params.data_ptr() + 4 * hb, img.data_ptr(), out.data_ptr(),
                float(floor if floor is not None else 0.0), int(floor is not None), st)

    def _fused_buffers(self, n, h, w):
        if self._fshape == (n, h, w):
            return self._fbufs
        if h % 8 or w % 8:
            raise ValueError("image height and width must be multiples of 8 (three 2x2 poolings)")
        c1, c2, c3, c4 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        L = [(h, w), (h // 2, w // 2), (h // 4, w // 4), (h // 8, w // 8)]
        spec = {"e1": (0, c1), "p1": (1, c1), "p2": (2, c2), "p3": (3, c3), "a2": (1, c2), "e2": (1, c2), "a3": (2, c3), "e3": (2, c3),
                "a4": (3, c4), "e4": (3, c4), "a5": (2, c3), "d3": (2, c3), "a6": (1, c2), "d2": (1, c2),
                "a7": (0, c1)}
        self._fbufs = None
        self._fbufs = {k: _Act(self.torch, n, L[lv][0], L[lv][1], ch, self.dev) for k, (lv, ch) in spec.items()}
        self._fshape = (n, h, w)
        return self._fbufs

    def _check_img(self, img):
        t = self.torch
        if not (isinstance(img, t.Tensor) and img.is_cuda and img.dtype == t.float32 and img.is_contiguous()):
            raise ValueError("expected a contiguous float32 CUDA tensor [n, h, w]")
        if img.dim() == 2:
            img = img[None]
        if img.dim() != 3:
            raise ValueError("expected [n, h, w]")
        return img

    def forward(self, img, out=None, floor=None, stream=None, fused=True):
        """pred = img + head(net(img)) for img [n, h, w] float32 on the device.

        floor: clamp the prediction from below (an eEM start must be positive).
        fused: fold input conversion, pooling, upsampling and concatenations
        into the convolutions and the head into the last one (False: the
        training path's separate kernels).
        """
        img = self._check_img(img)
        n, h, w = img.shape
        st = _stream(stream)
        if out is None:
            out = self.torch.empty_like(img)
        elif not (isinstance(out, self.torch.Tensor) and out.is_cuda and out.dtype == self.torch.float32
                  and out.is_contiguous() and out.numel() == img.numel()):
            raise ValueError("out must be a contiguous float32 CUDA tensor with the input's size")
        if fused:
            self._forward_fused(img, n, h, w, st, out, floor)
            return out
        b = self._forward(img, n, h, w, st)
        h0, hb = self.offsets["head"][0], self.offsets["head"][1]
        nn_call("raft_nn_head", b["d1"].ptr(), b["d1"].ch, 0, self.params.data_ptr() + 4 * h0,
                self.params.data_ptr() + 4 * hb, img.data_ptr(), out.data_ptr(), None, None, 0, 0, None, None, None,
                None, n * h * w, float(floor if floor is not None else 0.0), int(floor is not None), 0, st)
        return out

    # -------------------------------------------------------------- backward
    def loss_and_grad(self, img, target, pred=None, stream=None):
        """Mean squared error of pred vs target; leaves d(loss)/d(params) in self.grads.

        Returns the loss as a 0-d float64 device tensor (no host sync)."""
        img = self._check_img(img)
        target = self._check_img(target)
        n, h, w = img.shape
        if target.shape != img.shape:
            raise ValueError("target shape differs from the input shape")
        st = _stream(stream)
        b = self._forward(img, n, h, w, st)
        L = b["levels"]
        P = n * h * w
        if pred is None:
            pred = self.torch.empty_like(img)
        h0, hb = self.offsets["head"][0], self.offsets["head"][1]
        g = self.grads.data_ptr()
        nn_call("raft_nn_head", b["d1"].ptr(), b["d1"].ch, 0, self.params.data_ptr() + 4 * h0,
                self.params.data_ptr() + 4 * hb, img.data_ptr(), pred.data_ptr(), target.data_ptr(), b["g_d1"].ptr(),
                b["g_d1"].ch, 0, g + 4 * h0, g + 4 * hb, b["loss"].data_ptr(), b["head_work"].data_ptr(), P, 0.0, 0, 0,
                st)
        c1, c2, c3 = self.base, 2 * self.base, 4 * self.base
        P1, P2, P3, P4 = [n * a * c for a, c in L]
        # level 1 decoder
        self._mask(b["g_d1"], b["d1"], 0, P1, st)
        self._wgrad("dec1b", b["a7"], 0, b["g_d1"], n, L[0], st)
        self._conv_bwd_data("dec1b", b["g_d1"], b["g_a7"], 0, n, L[0], st)
        self._mask(b["g_a7"], b["a7"], 0, P1, st)
        self._wgrad("dec1a", b["cat1"], 0, b["g_a7"], n, L[0], st)
        self._conv_bwd_data("dec1a", b["g_a7"], b["g_cat1"], 0, n, L[0], st)
        nn_call("raft_nn_up2", b["g_cat1"].ptr(), b["g_cat1"].ch, c1, b["g_d2"].ptr(), c2, 0, n, L[1][0], L[1][1], c2,
                1, st)
        # level 2 decoder
        self._mask(b["g_d2"], b["d2"], 0, P2, st)
        self._wgrad("dec2b", b["a6"], 0, b["g_d2"], n, L[1], st)
        self._conv_bwd_data("dec2b", b["g_d2"], b["g_a6"], 0, n, L[1], st)
        self._mask(b["g_a6"], b["a6"], 0, P2, st)
        self._wgrad("dec2a", b["cat2"], 0, b["g_a6"], n, L[1], st)
        self._conv_bwd_data("dec2a", b["g_a6"], b["g_cat2"], 0, n, L[1], st)
        nn_call("raft_nn_up2", b["g_cat2"].ptr(), b["g_cat2"].ch, c2, b["g_d3"].ptr(), c3, 0, n, L[2][0], L[2][1], c3,
                1, st)
        # level 3 decoder
        self._mask(b["g_d3"], b["d3"], 0, P3, st)
        self._wgrad("dec3b", b["a5"], 0, b["g_d3"], n, L[2], st)
        self._conv_bwd_data("dec3b", b["g_d3"], b["g_a5"], 0, n, L[2], st)
        self._mask(b["g_a5"], b["a5"], 0, P3, st)
        self._wgrad("dec3a", b["cat3"], 0, b["g_a5"], n, L[2], st)
        self._conv_bwd_data("dec3a", b["g_a5"], b["g_cat3"], 0, n, L[2], st)
        c4 = 8 * self.base
        nn_call("raft_nn_up2", b["g_cat3"].ptr(), b["g_cat3"].ch, c3, b["g_e4"].ptr(), c4, 0, n, L[3][0], L[3][1], c4,
                1, st)
        # bottleneck
        self._mask(b["g_e4"], b["e4"], 0, P4, st)
        self._wgrad("bot_b", b["a4"], 0, b["g_e4"], n, L[3], st)
        self._conv_bwd_data("bot_b", b["g_e4"], b["g_a4"], 0, n, L[3], st)
        self._mask(b["g_a4"], b["a4"], 0, P4, st)
        self._wgrad("bot_a", b["p3"], 0, b["g_a4"], n, L[3], st)
        self._conv_bwd_data("bot_a", b["g_a4"], b["g_p3"], 0, n, L[3], st)
        # encoder level 3
        nn_call("raft_nn_skip_grad", b["g_cat3"].ptr(), b["g_cat3"].ch, 0, b["g_p3"].ptr(), c3, 0, b["cat3"].ptr(),
                b["cat3"].ch, 0, b["g_e3"].ptr(), c3, 0, n, L[2][0], L[2][1], c3, st)
        self._wgrad("enc3b", b["a3"], 0, b["g_e3"], n, L[2], st)
        self._conv_bwd_data("enc3b", b["g_e3"], b["g_a3"], 0, n, L[2], st)
        self._mask(b["g_a3"], b["a3"], 0, P3, st)
        self._wgrad("enc3a", b["p2"], 0, b["g_a3"], n, L[2], st)
        self._conv_bwd_data("enc3a", b["g_a3"], b["g_p2"], 0, n, L[2], st)
        # encoder level 2
        nn_call("raft_nn_skip_grad", b["g_cat2"].ptr(), b["g_cat2"].ch, 0, b["g_p2"].ptr(), c2, 0, b["cat2"].ptr(),
                b["cat2"].ch, 0, b["g_e2"].ptr(), c2, 0, n, L[1][0], L[1][1], c2, st)
        self._wgrad("enc2b", b["a2"], 0, b["g_e2"], n, L[1], st)
        self._conv_bwd_data("enc2b", b["g_e2"], b["g_a2"], 0, n, L[1], st)
        self._mask(b["g_a2"], b["a2"], 0, P2, st)
        self._wgrad("enc2a", b["p1"], 0, b["g_a2"], n, L[1], st)
        self._conv_bwd_data("enc2a", b["g_a2"], b["g_p1"], 0, n, L[1], st)
        # encoder level 1
        nn_call("raft_nn_skip_grad", b["g_cat1"].ptr(), b["g_cat1"].ch, 0, b["g_p1"].ptr(), c1, 0, b["cat1"].ptr(),
                b["cat1"].ch, 0, b["g_e1"].ptr(), c1, 0, n, L[0][0], L[0][1], c1, st)
        self._wgrad("enc1b", b["a1"], 0, b["g_e1"], n, L[0], st)
        self._conv_bwd_data("enc1b", b["g_e1"], b["g_a1"], 0, n, L[0], st)
        self._mask(b["g_a1"], b["a1"], 0, P1, st)
        self._wgrad("enc1a", b["x0"], 0, b["g_a1"], n, L[0], st)
        return b["loss"][0] / P

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8, gscale=1.0, stream=None):
        st = _stream(stream)
        self.step_count += 1
        nn_call("raft_nn_adam", self.params.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                self.params_bf.data_ptr(), self.n, lr, b1, b2, eps, self.step_count, gscale, st)
        self._flip_all(st)

    def train_step(self, img, target, lr, stream=None):
        loss = self.loss_and_grad(img, target, stream=stream)
        self.adam_step(lr, stream=stream)
        return loss

    def flops_forward(self, n, h, w, fused=True):
        """Multiply-add FLOPs (x2) of the 3x3 convolutions of one forward pass.
        fused: the first layer has its one real input channel (the training
        path convolves the 32-channel zero-padded input)."""
        lv = {"enc1": 0, "dec1": 0, "enc2": 1, "dec2": 1, "enc3": 2, "dec3": 2, "bot_": 3}
        tot = 0
        for name, co, ci in self.layers:
            s = 2 ** lv[name[:4]]
            if fused and name == "enc1a":
                ci = 1
            tot += 2 * 9 * co * ci * n * (h // s) * (w // s)
        return tot

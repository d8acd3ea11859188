"""Learned warm start: a periodic 3D convolution network rho -> the 6 unit-strain correctors, used as
the initial guess u0 of the PCG (nlfiber/linear.py:212-216 takes u0; the reference warm-starts only
from the previous solve, nlfiber/fiber.py:41-56).

Architecture (3 levels, widths 16/32/64, channels-last, periodic 3x3x3 convolutions, leaky ReLU):

    x0 (rho, 16 ch padded) -> e1a -> e1 (16) -> pool -> e2a -> e2 (32) -> pool -> e3a -> e3 (64)
    d2 = conv(concat(up(e3), e2)) (32) ; d1 = conv(concat(up(d2), e1)) (16) ; out = conv(d1) (18 of 32)

Every convolution, its data and weight gradients, pooling, up-sampling gradients, the loss and the
Adam step run in libnlfh.so (BF16 MFMA with FP32 accumulation, FP32 master weights); torch only
holds the buffers. The output, times the cell side n, is the corrector field: channel 3 I + c is
component c of the corrector of unit strain I.
"""
import ctypes

import numpy as np
import torch

from . import _lib

LAYERS = [
    # name, ca, a_half, cb, cout, coutp, act, level (0 = full resolution)
    ("e1a", 16, 0, 0, 16, 16, 1, 0),
    ("e1", 16, 0, 0, 16, 16, 1, 0),
    ("e2a", 16, 0, 0, 32, 32, 1, 1),
    ("e2", 32, 0, 0, 32, 32, 1, 1),
    ("e3a", 32, 0, 0, 64, 64, 1, 2),
    ("e3", 64, 0, 0, 64, 64, 1, 2),
    ("d2", 64, 1, 32, 32, 32, 1, 1),
    ("d1", 32, 1, 16, 16, 16, 1, 0),
    ("out", 16, 0, 0, 32, 32, 0, 0),  # 18 corrector channels; channels 18..31 stay zero
]
NOUT = 18
COUT_HEAD = 32


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _st():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class WarmStartNet:
    """The corrector predictor. `params` (optional) is a dict of float32 numpy arrays
    {name: w [27, coutp, cin], name + '.b': b [cout]} (the format of save / load)."""

    def __init__(self, seed=0, params=None):
        self.lib = _lib.load()
        self.dev = torch.device("cuda", torch.cuda.current_device())
        g = np.random.default_rng(seed)
        self.w, self.b, self.wh = {}, {}, {}
        self.m1, self.m2 = {}, {}
        for name, ca, ah, cb, cout, coutp, act, lvl in LAYERS:
            cin = ca + cb
            if params is not None:
                w = np.asarray(params[name], np.float32)
                b = np.asarray(params[name + ".b"], np.float32)
            else:
                w = (g.standard_normal((27, coutp, cin)) * np.sqrt(2.0 / (27 * cin))).astype(np.float32)
                if name == "out":
                    w *= 0.1
                    w[:, NOUT:, :] = 0.0  # unused head channels: zero weights, zero gradients
                b = np.zeros(cout, np.float32)
            self.w[name] = torch.tensor(w, device=self.dev)
            self.b[name] = torch.tensor(b, device=self.dev)
            self.wh[name] = torch.empty(w.shape, dtype=torch.int16, device=self.dev)
            self._bf(self.w[name], self.wh[name])
            self.m1[name] = (torch.zeros_like(self.w[name]), torch.zeros_like(self.b[name]))
            self.m2[name] = (torch.zeros_like(self.w[name]), torch.zeros_like(self.b[name]))
        self.step = 0
        self._part = None

    # --- helpers ---------------------------------------------------------------------------
    def _bf(self, src_f32, dst_i16):
        _lib.check(self.lib.nlfh_net_f2bf(_p(src_f32), _p(dst_i16), src_f32.numel(), _st()))

    def params(self):
        out = {}
        for name, *_ in LAYERS:
            out[name] = self.w[name].cpu().numpy()
            out[name + ".b"] = self.b[name].cpu().numpy()
        return out

    def save(self, path, **meta):
        np.savez(path, **self.params(), **{"meta_" + k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            params = {k: f[k] for k in f.files if not k.startswith("meta_")}
        return cls(params=params)

    def input(self, rho):
        """(B, n, n, n) densities -> bf16 channels-last input with channel 0 = 2 rho - 1."""
        rho = rho.to(self.dev, torch.float64).contiguous()
        B, n = rho.shape[0], rho.shape[1]
        xh = torch.empty((B, n, n, n, 16), dtype=torch.int16, device=self.dev)
        _lib.check(self.lib.nlfh_net_input(_p(rho), rho.numel(), _p(xh), _st()))
        return xh

    def _conv(self, name, a, b, n, B, out_f=None, out_h=None):
        _, ca, ah, cb, cout, coutp, act, _ = [l for l in LAYERS if l[0] == name][0]
        _lib.check(self.lib.nlfh_net_conv(_p(a), ca, ah, _p(b), cb, _p(self.wh[name]), _p(self.b[name]), n, B, cout,
                                          coutp, act, 0, _p(out_f), _p(out_h), _st()))

    # --- forward / backward ---------------------------------------------------------------
    def forward(self, x0, keep=False):
        """x0 bf16 (B, n, n, n, 16) -> out float32 (B, n^3, 18)."""
        B, n = x0.shape[0], x0.shape[1]
        h = {1: n // 2, 2: n // 4}
        dev, i16 = self.dev, torch.int16

        def buf(c, lvl, dtype=i16):
            s = n if lvl == 0 else h[lvl]
            return torch.empty((B, s, s, s, c), dtype=dtype, device=dev)

        A = {"x0": x0}
        A["e1a"] = buf(16, 0); self._conv("e1a", x0, None, n, B, out_h=A["e1a"])
        A["e1"] = buf(16, 0); self._conv("e1", A["e1a"], None, n, B, out_h=A["e1"])
        A["p1"] = buf(16, 1)
        _lib.check(self.lib.nlfh_net_pool(_p(A["e1"]), _p(A["p1"]), n, B, 16, _st()))
        A["e2a"] = buf(32, 1); self._conv("e2a", A["p1"], None, h[1], B, out_h=A["e2a"])
        A["e2"] = buf(32, 1); self._conv("e2", A["e2a"], None, h[1], B, out_h=A["e2"])
        A["p2"] = buf(32, 2)
        _lib.check(self.lib.nlfh_net_pool(_p(A["e2"]), _p(A["p2"]), h[1], B, 32, _st()))
        A["e3a"] = buf(64, 2); self._conv("e3a", A["p2"], None, h[2], B, out_h=A["e3a"])
        A["e3"] = buf(64, 2); self._conv("e3", A["e3a"], None, h[2], B, out_h=A["e3"])
        A["d2"] = buf(32, 1); self._conv("d2", A["e3"], A["e2"], h[1], B, out_h=A["d2"])
        A["d1"] = buf(16, 0); self._conv("d1", A["d2"], A["e1"], n, B, out_h=A["d1"])
        out = torch.empty((B, n, n, n, COUT_HEAD), dtype=torch.float32, device=dev)
        A["out"] = out
        _lib.check(self.lib.nlfh_net_conv(_p(A["d1"]), 16, 0, None, 0, _p(self.wh["out"]), _p(self.b["out"]), n, B,
                                          COUT_HEAD, COUT_HEAD, 0, 0, _p(out), None, _st()))
        if keep:
            self._acts = A
        return out

    def _part_buf(self):
        if self._part is None:
            self._part = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=self.dev)
        return self._part

    def backward(self, dout):
        """dout float32 (B, n, n, n, 32): gradient of the loss w.r.t. the head output. Returns
        {name: (dW, db)}."""
        A = self._acts
        B, n = dout.shape[0], dout.shape[1]
        h = {0: n, 1: n // 2, 2: n // 4}
        dev = self.dev
        part = self._part_buf()
        grads = {}
        lib = self.lib

        def layer(name):
            return [l for l in LAYERS if l[0] == name][0]

        def wb_grads(name, dy, y, a, b, s):
            _, ca, ah, cb, cout, coutp, act, _ = layer(name)
            nv = B * s ** 3
            g = torch.empty((B, s, s, s, coutp), dtype=torch.int16, device=dev)
            _lib.check(lib.nlfh_net_act_grad(_p(dy), _p(y) if act else None, _p(g), dy.numel(), act, _st()))
            dw = torch.empty_like(self.w[name])
            _lib.check(lib.nlfh_net_wgrad(_p(g), coutp, _p(a), ca, ah, _p(b), cb, s, B, _p(part), part.numel(), _p(dw),
                                          _st()))
            db = torch.empty_like(self.b[name])
            _lib.check(lib.nlfh_net_bias_grad(_p(dy), _p(y) if act else None, nv, cout, act, _p(part), part.numel(),
                                              _p(db), _st()))
            grads[name] = (dw, db)
            return g

        def dgrad(name, g, s):
            """input gradient (f32, ca + cb channels at resolution s) of layer `name`."""
            _, ca, ah, cb, cout, coutp, act, _ = layer(name)
            cin = ca + cb
            wt = torch.empty((27, cin, coutp), dtype=torch.int16, device=dev)
            _lib.check(lib.nlfh_net_wflip(_p(self.wh[name]), coutp, cin, _p(wt), cin, _st()))
            d = torch.empty((B, s, s, s, cin), dtype=torch.float32, device=dev)
            _lib.check(lib.nlfh_net_conv(_p(g), coutp, 0, None, 0, _p(wt), None, s, B, cin, cin, 0, 0, _p(d), None,
                                         _st()))
            return d

        # head
        g = wb_grads("out", dout, None, A["d1"], None, h[0])
        dd1 = dgrad("out", g, h[0])
        # d1 = conv(concat(up(d2), e1))
        g = wb_grads("d1", dd1, A["d1"], A["d2"], A["e1"], h[0])
        din = dgrad("d1", g, h[0])  # (B, n^3, 32 + 16)
        dd2 = torch.empty((B, h[1], h[1], h[1], 32), dtype=torch.float32, device=dev)
        _lib.check(lib.nlfh_net_upsum(_p(din), 48, 0, _p(dd2), h[0], B, 32, 1.0, 0, _st()))
        de1 = torch.empty((B, h[0], h[0], h[0], 16), dtype=torch.float32, device=dev)
        _lib.check(lib.nlfh_net_slice_add(_p(din), 48, 32, _p(de1), B * h[0] ** 3, 16, 0, _st()))
        # d2 = conv(concat(up(e3), e2))
        g = wb_grads("d2", dd2, A["d2"], A["e3"], A["e2"], h[1])
        din = dgrad("d2", g, h[1])  # (B, (n/2)^3, 64 + 32)
        de3 = torch.empty((B, h[2], h[2], h[2], 64), dtype=torch.float32, device=dev)
        _lib.check(lib.nlfh_net_upsum(_p(din), 96, 0, _p(de3), h[1], B, 64, 1.0, 0, _st()))
        de2 = torch.empty((B, h[1], h[1], h[1], 32), dtype=torch.float32, device=dev)
        _lib.check(lib.nlfh_net_slice_add(_p(din), 96, 64, _p(de2), B * h[1] ** 3, 32, 0, _st()))
        # e3 = conv(e3a), e3a = conv(p2), p2 = pool(e2)
        g = wb_grads("e3", de3, A["e3"], A["e3a"], None, h[2])
        de3a = dgrad("e3", g, h[2])
        g = wb_grads("e3a", de3a, A["e3a"], A["p2"], None, h[2])
        dp2 = dgrad("e3a", g, h[2])
        _lib.check(lib.nlfh_net_updist(_p(dp2), _p(de2), h[1], B, 32, 0.125, 1, _st()))
        # e2 = conv(e2a), e2a = conv(p1), p1 = pool(e1)
        g = wb_grads("e2", de2, A["e2"], A["e2a"], None, h[1])
        de2a = dgrad("e2", g, h[1])
        g = wb_grads("e2a", de2a, A["e2a"], A["p1"], None, h[1])
        dp1 = dgrad("e2a", g, h[1])
        _lib.check(lib.nlfh_net_updist(_p(dp1), _p(de1), h[0], B, 16, 0.125, 1, _st()))
        # e1 = conv(e1a), e1a = conv(x0)
        g = wb_grads("e1", de1, A["e1"], A["e1a"], None, h[0])
        de1a = dgrad("e1", g, h[0])
        wb_grads("e1a", de1a, A["e1a"], A["x0"], None, h[0])
        return grads

    def loss_and_grad(self, out, target):
        """MSE over the 18 corrector channels; out (B, n, n, n, 32), target (B, n, n, n, 18) f32."""
        full = torch.zeros_like(out)
        full[..., :NOUT] = target
        grad = torch.empty_like(out)
        parts = torch.empty(256, dtype=torch.float64, device=self.dev)
        len18 = out.numel() // COUT_HEAD * NOUT
        _lib.check(self.lib.nlfh_net_mse(_p(out), _p(full), _p(grad), out.numel(), _p(parts), 256, _st()))
        # the kernel normalises by all 32 channels; rescale to the mean over the 18 used ones
        scale = out.numel() / len18
        grad.mul_(scale)
        grad[..., NOUT:] = 0.0
        return float(parts.sum().item() * scale), grad

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        for name, (dw, db) in grads.items():
            m1w, m1b = self.m1[name]
            m2w, m2b = self.m2[name]
            _lib.check(self.lib.nlfh_net_adam(_p(self.w[name]), _p(dw), _p(m1w), _p(m2w), _p(self.wh[name]),
                                              dw.numel(), lr, b1, b2, eps, self.step, wd, _st()))
            _lib.check(self.lib.nlfh_net_adam(_p(self.b[name]), _p(db.contiguous()), _p(m1b), _p(m2b), None,
                                              db.numel(), lr, b1, b2, eps, self.step, 0.0, _st()))

    # --- use as a warm start ----------------------------------------------------------------
    def predict(self, rho):
        """rho (n, n, n) or (B, n, n, n) -> correctors (B, 6, n, n, n, 3) float64 (reference layout per
        cell: (6, nx, ny, nz, 3)), mean-free per component."""
        r = rho if isinstance(rho, torch.Tensor) else torch.as_tensor(np.asarray(rho))
        if r.dim() == 3:
            r = r.unsqueeze(0)
        n = r.shape[1]
        out = self.forward(self.input(r))
        B = out.shape[0]
        # head channels 0..17 (right-hand side x 3 + component) scaled by n, mean-free per channel, float64, written
        # in the planar layout (B, 6, 3, n, n, n) by one read-out; returned as a (B, 6, n, n, n, 3) view of it
        planar = torch.empty((B, 6, 3, n, n, n), dtype=torch.float64, device=self.dev)
        part = torch.empty(B * 513 * NOUT, dtype=torch.float64, device=self.dev)
        _lib.check(self.lib.nlfh_net_readout(_p(out), COUT_HEAD, NOUT, n ** 3, B, float(n), _p(part), part.numel(),
                                             _p(planar), _st()))
        return planar.permute(0, 1, 3, 4, 5, 2)

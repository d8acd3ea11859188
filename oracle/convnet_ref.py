"""Float64 torch-CPU restatement of the learned warm-start network (test infrastructure only).

TEST INFRASTRUCTURE ONLY: the GPU tests compare the library's BF16-MFMA network kernels (forward,
data and weight gradients, pooling, up-sampling, loss) against this reference; the product never
imports it. The network has no counterpart in the reference repository's linear path; its layout
follows nlfiber/prior/unet.py:83-139 (3 levels, widths 16/32/64) with circular instead of symmetric
padding because the inputs are whole periodic cells.

`bf16=True` rounds weights, the input and every stored activation to bfloat16 (round to nearest even)
as the GPU kernels do, so that the forward pass can be compared tightly; gradients are exact float64.
"""
import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [
    ("e1a", 16, 0, 0, 16, 16, 1, 0),
    ("e1", 16, 0, 0, 16, 16, 1, 0),
    ("e2a", 16, 0, 0, 32, 32, 1, 1),
    ("e2", 32, 0, 0, 32, 32, 1, 1),
    ("e3a", 32, 0, 0, 64, 64, 1, 2),
    ("e3", 64, 0, 0, 64, 64, 1, 2),
    ("d2", 64, 1, 32, 32, 32, 1, 1),
    ("d1", 32, 1, 16, 16, 16, 1, 0),
    ("out", 16, 0, 0, 32, 32, 0, 0),
]


class _RoundBF16(torch.autograd.Function):
    """bfloat16 rounding (nearest even) in the forward pass, identity in the backward pass, so the
    reference gradients stay exact float64 functions of the rounded forward values."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g


def _r(x, bf16):
    return _RoundBF16.apply(x) if bf16 else x


def conv(x, w, b, act, bf16):
    """x (B, n, n, n, cin) channels-last; w (27, cout, cin) with tap t = kx*9 + ky*3 + kz reading
    x[v + (kx-1, ky-1, kz-1)] periodically."""
    cout, cin = w.shape[1], w.shape[2]
    W = w.reshape(3, 3, 3, cout, cin).permute(3, 4, 0, 1, 2)
    xt = F.pad(x.permute(0, 4, 1, 2, 3), (1, 1, 1, 1, 1, 1), mode="circular")
    y = F.conv3d(xt, W) + b[None, :, None, None, None]
    y = y.permute(0, 2, 3, 4, 1)
    if act:
        y = F.leaky_relu(y, 0.1)
    return y


def pool(x):
    B, n, _, _, C = x.shape
    return x.reshape(B, n // 2, 2, n // 2, 2, n // 2, 2, C).mean(dim=(2, 4, 6))


def up(x):
    return x.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)


def forward(params, rho, bf16=True):
    """params: {name: w (27, coutp, cin) tensor, name + '.b': b (cout,)} float64 tensors (leaves for
    autograd); rho (B, n, n, n). Returns the head output (B, n, n, n, 32)."""
    P = {k: _r(v, bf16) if not k.endswith(".b") else v for k, v in params.items()}
    B, n = rho.shape[0], rho.shape[1]
    x0 = torch.zeros((B, n, n, n, 16), dtype=torch.float64)
    x0[..., 0] = 2.0 * rho - 1.0
    x0 = _r(x0, bf16)

    def c(name, x, act=1):
        return _r(conv(x, P[name], P[name + ".b"], act, bf16), bf16)

    e1 = c("e1", c("e1a", x0))
    p1 = _r(pool(e1), bf16)
    e2 = c("e2", c("e2a", p1))
    p2 = _r(pool(e2), bf16)
    e3 = c("e3", c("e3a", p2))
    d2 = c("d2", torch.cat([up(e3), e2], dim=-1))
    d1 = c("d1", torch.cat([up(d2), e1], dim=-1))
    return conv(d1, P["out"], P["out.b"], 0, bf16)


def loss(out, target):
    """mean squared error over the 18 corrector channels."""
    return ((out[..., :18] - target) ** 2).mean()

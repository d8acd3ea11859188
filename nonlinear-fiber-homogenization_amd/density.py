"""Density map on the GPU: pre-density phi (symmetric octant) -> physical density rho (full cell) with exact
volume, and its vector-Jacobian product (nlfiber/density.py:76-91, nlfiber/symmetry.py:26-102).

    u   = exact_symmetrize(F mirror(sym_perm(phi)) / std)
    rho = sigmoid(beta (u - eta)),   mean(rho) = v

Every step runs in libnlfh.so (nlfh_dmap_*): the periodic hat filter is a direct stencil (its support
is |offset| <= r voxels), the orbit values are computed once per orbit, eta by bisection and Newton in
one block. `rho` is differentiable in torch autograd with the reference's implicit eta derivative.
Three-dimensional cells only (the reference's d = 2 path is not on the GPU).
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _RhoFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, dm, beta):
        ctx.dm, ctx.beta = dm, beta
        ctx.save_for_backward(phi)
        return dm._rho_raw(phi, beta)

    @staticmethod
    def backward(ctx, g):
        (phi,) = ctx.saved_tensors
        return ctx.dm.rho_vjp(phi, ctx.beta, g.unsqueeze(0))[0], None, None


class DensityMap:
    """phi (symmetric octant, (n_o)^3) -> rho (full cell, (2 n_o)^3) with volume v at sharpness beta."""

    def __init__(self, n_o, d=3, r=1.5, v=0.3):
        if d != 3:
            raise NotImplementedError("the GPU density map handles three-dimensional cells (d = 3)")
        self.n_o, self.d, self.r, self.v = int(n_o), d, float(r), float(v)
        self.n = 2 * self.n_o
        self.shape = (self.n,) * 3
        self._lib = _lib.load()
        self._dev = torch.device("cuda", torch.cuda.current_device())
        h = ctypes.c_void_p()
        _lib.check(self._lib.nlfh_dmap_create(self.n_o, self.r, self.v, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.nlfh_dmap_destroy(h)
            except Exception:
                pass
            self._h = None

    def _phi(self, phi):
        t = phi if isinstance(phi, torch.Tensor) else torch.as_tensor(np.asarray(phi))
        t = t.detach().to(device=self._dev, dtype=torch.float64).contiguous()
        if tuple(t.shape) != (self.n_o,) * 3:
            raise ValueError(f"phi must be ({self.n_o},)*3, got {tuple(t.shape)}")
        return t

    def smooth(self, phi):
        """Filtered, scaled, exactly symmetrised full-cell field u (nlfiber/density.py:80-84)."""
        p = self._phi(phi)
        u = torch.empty(self.shape, dtype=torch.float64, device=self._dev)
        _lib.check(self._lib.nlfh_dmap_smooth(self._h, _ptr(p), _ptr(u), _stream()))
        return u

    def _rho_raw(self, phi, beta):
        p = self._phi(phi)
        rho = torch.empty(self.shape, dtype=torch.float64, device=self._dev)
        _lib.check(self._lib.nlfh_dmap_rho(self._h, _ptr(p), float(beta), _ptr(rho), None, _stream()))
        return rho

    def rho(self, phi, beta):
        """rho(phi, beta) (nlfiber/density.py:86-87); differentiable in phi when phi requires grad."""
        if isinstance(phi, torch.Tensor) and phi.requires_grad:
            return _RhoFn.apply(phi, self, float(beta))
        return self._rho_raw(phi, beta)

    def eta(self, phi, beta):
        """The threshold eta (nlfiber/density.py:89-91)."""
        p = self._phi(phi)
        rho = torch.empty(self.shape, dtype=torch.float64, device=self._dev)
        eta = ctypes.c_double()
        _lib.check(self._lib.nlfh_dmap_rho(self._h, _ptr(p), float(beta), _ptr(rho), ctypes.byref(eta), _stream()))
        return eta.value

    def rho_vjp(self, phi, beta, rbar):
        """(d rho / d phi)^T rbar for rows rbar (k, n, n, n) -> (k, n_o, n_o, n_o)."""
        p = self._phi(phi)
        rb = rbar.detach().to(device=self._dev, dtype=torch.float64).contiguous()
        if rb.dim() == 3:
            rb = rb.unsqueeze(0)
        out = torch.empty((rb.shape[0],) + (self.n_o,) * 3, dtype=torch.float64, device=self._dev)
        for i in range(0, rb.shape[0], 3):
            blk = rb[i:i + 3].contiguous()
            _lib.check(self._lib.nlfh_dmap_rho_vjp(self._h, _ptr(p), float(beta), _ptr(blk), blk.shape[0],
                                                   _ptr(out[i:i + 3]), _stream()))
        return out

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
import itertools

import numpy as np
import torch

from . import _lib
from . import topology


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

    def binary(self, phi, up=1, v=None, remove_floating=True):
        """Binary cell of side up * 2 n_o (nlfiber/density.py:96-121): the smoothed field u (on the GPU), upsampled
        periodically by `up` and symmetrised again, thresholded at its (1 - v) quantile; with remove_floating, the
        pieces that do not span the cell are dropped and the kept solid grows back to volume v through the void
        voxels touching it, highest u first (whole orbits at a time, since u is exactly symmetric), until it is
        within one voxel of v. Host arrays (numpy bool)."""
        v = self.v if v is None else float(v)
        u = self.smooth(phi).cpu().numpy()
        if up > 1:
            u = exact_symmetrize_host(topology.upsample_periodic(u, up))
        b = u > np.quantile(u, 1.0 - v)
        if remove_floating:
            for _ in range(500):
                b = topology.remove_floating(b)
                missing = v - b.mean()
                if missing * b.size < 1 or not b.any():
                    break
                front = topology.dilate_periodic(b) & ~b
                cand = np.sort(u[front])[::-1]
                if cand.size == 0:
                    break
                k = min(max(int(round(missing * b.size)), 1), cand.size)
                b = b | (front & (u >= cand[k - 1]))
            b = topology.remove_floating(b)
        return b

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


def exact_symmetrize_host(full):
    """Cubic (Oh) symmetrisation of a full cell (2 n_o)^3 in which every orbit gets one value exactly
    (nlfiber/symmetry.py:61-66): the 8 mirror images averaged axis by axis (x, then y, then z), the octant averaged
    over the 6 axis permutations (accumulated in itertools order), each octant entry replaced by the entry at
    its sorted index, and the octant mirrored back out. numpy float64; the GPU density map does the same for
    its own cells."""
    x = np.asarray(full, dtype=np.float64)
    for ax in range(3):
        h = x.shape[ax] // 2
        lo = np.take(x, np.arange(h), axis=ax)
        hi = np.flip(np.take(x, np.arange(h, 2 * h), axis=ax), axis=ax)
        x = 0.5 * (lo + hi)
    acc = 0
    for perm in itertools.permutations(range(3)):
        acc = acc + np.transpose(x, perm)
    octant = acc / 6
    srt = np.sort(np.indices(octant.shape).reshape(3, -1), axis=0)
    octant = octant[srt[0], srt[1], srt[2]].reshape(octant.shape)
    for ax in range(3):
        octant = np.concatenate([octant, np.flip(octant, axis=ax)], axis=ax)
    return octant

"""The fiber constraint map on the GPU: phi (symmetric octant) -> normalized stiffness invariants
c = (C11, C12, C44) / E, with Jacobian rows through the density map and the correctors held fixed
(nlfiber/fiber.py:17-81).

`Fiber.value` and `Fiber.value_jac` keep the reference's signatures and warm start (the correctors of
the last solve start the next one, nlfiber/fiber.py:41-56). The density map, the 2-RHS PCG, the
per-element energies and every vector-Jacobian product run in libnlfh.so; the three Jacobian rows are
one batched density-map VJP of the three rho-gradients of c. `binary_check` binarizes on the host
(DensityMap.binary, periodic connectivity in `topology`) and homogenizes the binary cell on the GPU.
"""
import itertools

import numpy as np
import torch

from .density import DensityMap
from .linear import LinearHomogenization, cubic_invariants, cubic_leakage  # noqa: F401 (re-exported as in the reference)


class Fiber:
    def __init__(self, n_o, d=3, v=0.3, r=1.5, E=1.0, nu=0.3, emin=1e-3, p=3.0, tol=1e-9, precision="fp64"):
        """precision (not in the reference): the linear solver's, "fp64" (the reference's PCG) or "mixed"."""
        self.n_o, self.d, self.v = n_o, d, v
        self.dm = DensityMap(n_o, d, r, v)
        self.lh = LinearHomogenization(self.dm.shape, E=E, nu=nu, emin=emin, p=p, tol=tol, precision=precision)
        self._mat = dict(E=E, nu=nu, emin=emin, p=p, tol=tol)
        self._lh_bin = {}
        self.u = None  # warm start of the correctors
        self.solves = 0

    def _u0(self):
        if self.u is None:
            return torch.zeros((2,) + self.lh.shape + (3,), dtype=torch.float64, device=self.dm._dev)
        return self.u

    def cfun(self, phi, beta, u0):
        """c(phi) and (correctors, iterations); differentiable in phi through torch autograd."""
        c, u, it = self.lh.invariants(self.dm.rho(phi, beta), u0)
        return c, (u, it)

    def value(self, phi, beta, warm=True):
        u0 = self._u0() if warm else torch.zeros_like(self._u0())
        c, u, it = self.lh.invariants(self.dm.rho(phi, beta), u0)
        self.u = u.detach()
        self.solves += 1
        return np.asarray(c.detach())

    def value_jac(self, phi, beta, warm=True):
        """(c, J) with J (3, n_o, n_o, n_o) = d c / d phi (correctors held fixed, as the reference's
        jax.vjp through `invariants` does)."""
        u0 = self._u0() if warm else torch.zeros_like(self._u0())
        rho = self.dm.rho(phi, beta).requires_grad_(True)
        c, u, it = self.lh.invariants(rho, u0)
        rows = [torch.autograd.grad(c[i], rho, retain_graph=i < 2)[0] for i in range(3)]
        J = self.dm.rho_vjp(phi, beta, torch.stack(rows))
        self.u = u.detach()
        self.solves += 1
        return np.asarray(c.detach()), J

    def rho(self, phi, beta):
        return self.dm.rho(phi, float(beta))

    def lh_for(self, shape):
        if shape not in self._lh_bin:
            self._lh_bin[shape] = LinearHomogenization(shape, **self._mat)
        return self._lh_bin[shape]

    def binary_check(self, phi, up=1, full=False):
        """Binarize phi (upsampled by `up`), drop floating pieces and homogenize the binary cell
        (nlfiber/fiber.py:67-81). Returns dict(vol, cell, c) and, with full=True, C (6x6), c and the cubic
        leakage from the full C."""
        b = self.dm.binary(phi, up=up)
        lh = self.lh_for(b.shape)
        rho = torch.as_tensor(b, dtype=torch.float64, device=self.dm._dev)
        out = dict(vol=float(b.mean()), cell=b)
        if full:
            C, _, _ = lh.C_full(rho)
            C = np.asarray(C)
            out.update(C=C, c=cubic_invariants(C) / lh.E, leak=cubic_leakage(C))
        else:
            c, _, _ = lh.invariants(rho)
            out["c"] = np.asarray(c.detach().cpu() if isinstance(c, torch.Tensor) else c)
        self.solves += 1
        return out


def rel_error(c, target):
    """nlfiber/fiber.py:84-86."""
    c, t = np.asarray(c), np.asarray(target)
    return float(np.linalg.norm(c - t) / np.linalg.norm(t))


def sigma_min_normalized(J):
    """Scale the Jacobian's rows to unit length and return its least singular value, a rank test
    (nlfiber/fiber.py:96-100)."""
    Jf = np.asarray(J.detach().cpu() if isinstance(J, torch.Tensor) else J).reshape(J.shape[0], -1)
    Jf = Jf / np.linalg.norm(Jf, axis=1, keepdims=True)
    return float(np.linalg.svd(Jf, compute_uv=False).min())


def gram_logdet(J, metric=None):
    """The log-volume log det(J G J^T) spanned by the Jacobian's rows under the metric G (nlfiber/fiber.py:89-93);
    `metric` applies G to one row (identity if None)."""
    Jt = J if isinstance(J, torch.Tensor) else torch.as_tensor(np.asarray(J))
    Jf = Jt.reshape(Jt.shape[0], -1)
    GJ = Jf if metric is None else torch.stack([torch.as_tensor(metric(r)).reshape(-1) for r in Jt])
    return float(torch.linalg.slogdet(Jf @ GJ.T.to(Jf))[1])


def random_start(n_o, d=3, seed=0, kmax=3, nterms=40):
    """A seeded pre-density for starting a design: a sum of a few low-frequency cosines with the cell's symmetry,
    shifted and scaled to zero mean and unit spread (nlfiber/fiber.py:103-121); a host array."""
    rng = np.random.default_rng(seed)
    n = 2 * n_o
    x = (np.arange(n_o) + 0.5) / n
    f = np.zeros((n_o,) * d)
    for _ in range(nterms):
        k = rng.integers(0, kmax + 1, size=d)
        if not k.any():
            continue
        a = rng.standard_normal() / np.linalg.norm(k)
        term = np.ones(())
        for ax in range(d):
            shp = [1] * d
            shp[ax] = n_o
            term = term * np.cos(2 * np.pi * k[ax] * x).reshape(shp)
        f = f + a * term
    perms = list(itertools.permutations(range(d)))
    f = sum(np.transpose(f, p) for p in perms) / len(perms)
    return (f - f.mean()) / (f.std() + 1e-12)

"""Derivatives on the GPU against the oracle and central differences.

Mirrors the reference's tests/test_core.py:101-122: the invariants' Jacobian in rho (correctors held
fixed) and the adjoint VJP of the correctors, each checked by central differences of the GPU
solver and against the oracle's closed-form gradient.
"""
import numpy as np
import pytest
import torch

from oracle import linear_ref as O

pytestmark = pytest.mark.gpu


def _cell(pkg, n=8, seed=4):
    rng = np.random.default_rng(seed)
    return pkg.cells.exact_symmetrize(pkg.cells.density(rng.standard_normal((n, n, n)), r=1.0, beta=4.0, v=0.3,
                                                        filter_kernel="hat"))


def test_invariants_gradient_vs_oracle_and_fd(gpu_pkg):
    rho = _cell(gpu_pkg)
    lh = gpu_pkg.LinearHomogenization(rho.shape, tol=1e-12)
    r = torch.tensor(rho, device="cuda", requires_grad=True)
    c, u2, _ = lh.invariants(r)
    J = torch.stack([torch.autograd.grad(c[i], r, retain_graph=True)[0] for i in range(3)]).cpu().numpy()
    ref = O.LinearRef(rho.shape, tol=1e-12)
    cr, u2r, _ = ref.invariants(rho)
    assert np.abs(c.detach().numpy() - cr).max() < 1e-10 * np.abs(cr).max()
    assert np.abs(u2.cpu().numpy() - u2r).max() < 1e-8 * np.abs(u2r).max()
    Jr = ref.invariants_grad(rho, u2.cpu().numpy())
    assert np.abs(J - Jr).max() < 1e-10 * np.abs(Jr).max()
    d = np.random.default_rng(5).standard_normal(rho.shape)
    h = 1e-5
    cp = lh.invariants(rho + h * d)[0].detach().numpy()
    cm = lh.invariants(rho - h * d)[0].detach().numpy()
    fd = (cp - cm) / (2 * h)
    np.testing.assert_allclose(np.tensordot(J, d, 3), fd, rtol=1e-5, atol=1e-10)


def test_corrector_vjp_vs_fd(gpu_pkg):
    rho = _cell(gpu_pkg, seed=6)
    lh = gpu_pkg.LinearHomogenization(rho.shape, tol=1e-12)
    s0 = lh.stiffness_scale(rho)
    Es = np.eye(6)[:2]
    s = s0.clone().requires_grad_(True)
    u = lh.correctors(s, Es, None)
    f = (u ** 2).sum()
    (g,) = torch.autograd.grad(f, s)
    d = torch.tensor(np.random.default_rng(7).standard_normal(rho.shape), device="cuda")
    h = 1e-6
    fp = (lh.solve(s0 + h * d, Es)[0] ** 2).sum().item()
    fm = (lh.solve(s0 - h * d, Es)[0] ** 2).sum().item()
    fd = (fp - fm) / (2 * h)
    assert abs(float((g * d).sum()) - fd) < 1e-4 * abs(fd)
    # and against the oracle's adjoint
    ref = O.LinearRef(rho.shape, tol=1e-12)
    un = u.detach().cpu().numpy()
    gr = ref.correctors_vjp(s0.cpu().numpy(), Es, un, 2 * un)
    assert np.abs(g.cpu().numpy() - gr).max() < 1e-6 * np.abs(gr).max()


def test_eight_strains_correctors_vjp_and_bilinear(gpu_pkg):
    """More than six macroscopic strains: the solve, the load, the residual, the adjoint solve and C run in
    blocks of at most six (the library's limit per call); the reference's custom_vjp takes any number
    (nlfiber/linear.py:118-134). The gradient of sum(u^2) + sum(C) through the correctors and the bilinear
    form, and C itself, are checked against the oracle."""
    rho = _cell(gpu_pkg, seed=9)
    lh = gpu_pkg.LinearHomogenization(rho.shape, tol=1e-12)
    s0 = lh.stiffness_scale(rho)
    rng = np.random.default_rng(10)
    Es = np.concatenate([np.eye(6), rng.standard_normal((2, 6))])
    s = s0.clone().requires_grad_(True)
    u = lh.correctors(s, Es, None)
    assert u.shape == (8,) + rho.shape + (3,)
    C = lh.bilinear(s, Es, u)
    f = (u ** 2).sum()
    (g,) = torch.autograd.grad(f, s, retain_graph=True)
    (gC,) = torch.autograd.grad(C.sum(), s)
    ref = O.LinearRef(rho.shape, tol=1e-12)
    sn = s0.cpu().numpy()
    ur, _ = ref.solve(sn, Es)
    un = u.detach().cpu().numpy()
    assert np.abs(un - ur).max() < 1e-8 * np.abs(ur).max()
    Cr = ref.bilinear(sn, Es, un)
    assert np.abs(np.asarray(C.detach()) - Cr).max() < 1e-10 * np.abs(Cr).max()
    gr = ref.correctors_vjp(sn, Es, un, 2 * un)
    assert np.abs(g.cpu().numpy() - gr).max() < 1e-6 * np.abs(gr).max()
    b = ref.element_energy(sn, Es, un)
    gCr = b.sum(axis=(0, 1)) / b[0, 0].size  # d sum_IJ C_IJ / ds_e with the correctors fixed
    assert np.abs(gC.cpu().numpy() - gCr).max() < 1e-10 * np.abs(gCr).max()
    # load and residual in blocks: the residual of the correctors is at the solver tolerance
    r = lh.residual(s0, Es, u.detach())
    bnorm = lh.load(s0, Es).flatten(1).norm(dim=1)
    assert (r.flatten(1).norm(dim=1) / bnorm).max() < 1e-9

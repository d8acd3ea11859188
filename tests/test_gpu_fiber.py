"""Density map and fiber constraint map on the GPU against the oracle (nlfiber/density.py,
nlfiber/symmetry.py, nlfiber/fiber.py): values, exact volume and orbit symmetry, the density-map VJP,
and the invariants' Jacobian in phi against the oracle chain and central differences (the reference's
tests/test_core.py:43-53, 101-109)."""
import numpy as np
import pytest
import torch

from oracle import density_ref as DR
from oracle import linear_ref as O

pytestmark = pytest.mark.gpu


def _phi(n_o, seed):
    return DR.sym_perm(np.random.default_rng(seed).standard_normal((n_o,) * 3))


@pytest.mark.parametrize("n_o,r,v", [(4, 1.5, 0.27), (6, 2.5, 0.3), (8, 1.5, 0.2)])
def test_density_map_matches_oracle(gpu_pkg, n_o, r, v):
    from nlfh_amd.density import DensityMap
    phi = _phi(n_o, n_o)
    dm = DensityMap(n_o, r=r, v=v)
    u = dm.smooth(phi).cpu().numpy()
    ur = DR.smooth(phi, r)
    assert np.abs(u - ur).max() < 1e-12 * np.abs(ur).max()
    for beta in (2.0, 8.0, 32.0):
        rho = dm.rho(phi, beta).cpu().numpy()
        rr = DR.rho(phi, beta, r, v)
        assert np.abs(rho - rr).max() < 1e-10
        assert abs(rho.mean() - v) < 1e-12
        assert abs(dm.eta(phi, beta) - DR.solve_eta(ur, v, beta)) < 1e-10
        # every orbit carries one value exactly
        assert np.array_equal(rho, rho[::-1]) and np.array_equal(rho, np.transpose(rho, (2, 0, 1)))
        assert np.array_equal(rho, rho[:, ::-1, :]) and np.array_equal(rho, np.transpose(rho, (1, 0, 2)))


def test_density_map_vjp_matches_oracle_and_autograd(gpu_pkg):
    from nlfh_amd.density import DensityMap
    n_o, r, v, beta = 4, 1.5, 0.27, 4.0
    phi = _phi(n_o, 2)
    dm = DensityMap(n_o, r=r, v=v)
    rng = np.random.default_rng(5)
    rbar = rng.standard_normal((3,) + dm.shape)
    g = dm.rho_vjp(phi, beta, torch.tensor(rbar)).cpu().numpy()
    gr = DR.rho_vjp(phi, beta, r, v, rbar)
    assert np.abs(g - gr).max() < 1e-10 * np.abs(gr).max()
    # autograd through DensityMap.rho: d sum(rho^3) / d phi against central differences
    p = torch.tensor(phi, device="cuda", requires_grad=True)
    f = (dm.rho(p, beta) ** 3).sum()
    (gp,) = torch.autograd.grad(f, p)
    d = _phi(n_o, 3)
    h = 1e-6
    fd = ((dm.rho(phi + h * d, beta) ** 3).sum() - (dm.rho(phi - h * d, beta) ** 3).sum()).item() / (2 * h)
    assert abs(float((gp.cpu().numpy() * d).sum()) - fd) < 1e-7 * abs(fd) + 1e-10


def test_fiber_value_and_jacobian(gpu_pkg):
    """c and J = dc/dphi (correctors fixed) against the oracle chain (density map, 2-RHS solve, invariant
    gradient, density-map VJP) and central differences; the warm start is kept."""
    from nlfh_amd.fiber import Fiber
    n_o, v, r, beta = 4, 0.3, 1.5, 8.0
    phi = _phi(n_o, 7)
    fib = Fiber(n_o, v=v, r=r, tol=1e-12)
    c, J = fib.value_jac(phi, beta)
    rho = DR.rho(phi, beta, r, v)
    ref = O.LinearRef(rho.shape, emin=1e-3, tol=1e-12)
    cr, u2r, _ = ref.invariants(rho)
    assert np.abs(c - cr).max() < 1e-10 * np.abs(cr).max()
    Jrho = ref.invariants_grad(rho, u2r)
    Jr = DR.rho_vjp(phi, beta, r, v, Jrho)
    Jg = J.cpu().numpy()
    assert np.abs(Jg - Jr).max() < 1e-8 * np.abs(Jr).max()
    d = _phi(n_o, 8)
    h = 1e-5
    cp = Fiber(n_o, v=v, r=r, tol=1e-12).value(phi + h * d, beta)
    cm = Fiber(n_o, v=v, r=r, tol=1e-12).value(phi - h * d, beta)
    np.testing.assert_allclose(np.tensordot(Jg, d, 3), (cp - cm) / (2 * h), rtol=1e-5, atol=1e-10)
    # warm start: the correctors of the last solve start the next one
    assert fib.u is not None and fib.solves == 1
    c2 = fib.value(phi, beta)
    assert fib.lh.last_info["it"] <= 2 and np.abs(c2 - c).max() < 1e-10 * np.abs(c).max()


def _schwarz_octant(n_o):
    x = 2 * np.pi * (np.arange(n_o) + 0.5) / (2 * n_o)
    G = np.meshgrid(x, x, x, indexing="ij")
    return -sum(np.cos(g) for g in G)


def test_fiber_invariant_gradient_reference_case(gpu_pkg):
    """The reference's tests/test_core.py:100-108 on the GPU Fiber (default tolerance 1e-9): J against
    central differences of cold solves, and the rank check of the normalised Jacobian."""
    from nlfh_amd.fiber import Fiber, random_start, sigma_min_normalized
    fib = Fiber(4, v=0.3, r=1.5)
    phi = DR.sym_perm(random_start(4, seed=4, kmax=1) + _schwarz_octant(4))
    c, J = fib.value_jac(phi, 4.0)
    d = DR.sym_perm(np.random.default_rng(5).standard_normal((4, 4, 4)))
    h = 1e-5
    fd = (fib.value(phi + h * d, 4.0, warm=False) - fib.value(phi - h * d, 4.0, warm=False)) / (2 * h)
    np.testing.assert_allclose(np.tensordot(J.cpu().numpy(), d, 3), fd, rtol=1e-5, atol=1e-10)
    assert sigma_min_normalized(J) > 1e-3


def test_identity_filter_radius(gpu_pkg):
    """r = 1.0 (the reference's test_pipeline fibers): the hat kernel is the identity."""
    from nlfh_amd.density import DensityMap
    phi = _phi(4, 11)
    dm = DensityMap(4, r=1.0, v=0.3)
    np.testing.assert_allclose(dm.smooth(phi).cpu().numpy(), DR.smooth(phi, 1.0), rtol=0, atol=1e-12)
    assert abs(dm.rho(phi, 4.0).mean().item() - 0.3) < 1e-12


@pytest.mark.parametrize("n_o,up", [(8, 1), (6, 2)])
def test_binary_check_matches_oracle(gpu_pkg, n_o, up):
    """Fiber.binary_check (nlfiber/fiber.py:67-81): the binary cell is the oracle's (same voxels), it has
    one spanning solid piece, its volume is within one voxel of v, every orbit is kept or dropped whole, and
    the GPU's C of it equals the FP64 oracle's C."""
    from nlfh_amd.fiber import Fiber
    from nlfh_amd.linear import cubic_invariants
    x = 2 * np.pi * (np.arange(n_o) + 0.5) / (2 * n_o)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = -(np.cos(X) + np.cos(Y) + np.cos(Z)) + 0.5 * _phi(n_o, 40)  # a perturbed Schwarz P octant
    fib = Fiber(n_o, v=0.3, r=1.5)
    out = fib.binary_check(phi, up=up, full=True)
    b = out["cell"]
    ref = DR.binary(phi, 1.5, 0.3, up=up)
    assert b.shape == (2 * n_o * up,) * 3 and np.array_equal(b, ref)
    lab, n = gpu_pkg.topology.periodic_label(b)
    assert n == 1 and gpu_pkg.topology.spanning(lab, 1)
    # grown back to v (whole orbits, so it may pass v by a few orbits of up to 48 voxels)
    assert (0.3 - out["vol"]) * b.size < 1.0 and out["vol"] < 0.32
    assert np.array_equal(b, b[::-1]) and np.array_equal(b, np.transpose(b, (1, 2, 0)))
    lr = O.LinearRef(b.shape, emin=1e-3, tol=1e-9)
    Cr, _, _ = lr.C_full(b.astype(np.float64))
    assert np.abs(out["C"] - Cr).max() < 1e-7 * np.abs(Cr).max()
    assert np.allclose(out["c"], cubic_invariants(Cr), rtol=1e-7, atol=0)
    c_only = fib.binary_check(phi, up=up)["c"]
    assert np.allclose(c_only, out["c"], rtol=1e-7, atol=0)

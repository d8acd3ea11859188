"""Solver parity on the GPU: C from the plain FP64 PCG against the oracle, the closed forms and the
reference's own test properties (tests/test_core.py:77-142 of the reference).

The gate is |dC|/max|C| <= 1e-6 (the workload's requirement); where both sides solve to
tol 1e-10 the agreement is far tighter and the tests ask for it.
"""
import os

import numpy as np
import pytest
import torch

from oracle import linear_ref as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _grey(pkg, n, seed=0, v=0.35, beta=3.0):
    rng = np.random.default_rng(seed)
    return pkg.cells.density(rng.standard_normal((n, n, n)), r=1.0, beta=beta, v=v, filter_kernel="hat")


def test_C_full_grey_cell_matches_oracle_and_iterations(gpu_pkg):
    rho = gpu_pkg.cells.exact_symmetrize(_grey(gpu_pkg, 8))
    lh = gpu_pkg.LinearHomogenization(rho.shape, tol=1e-10)
    C, u, it = lh.C_full(rho)
    Cr, ur, itr = O.LinearRef(rho.shape, tol=1e-10).C_full(rho)
    C = np.asarray(C)
    assert _rel(C, Cr) < 1e-10
    assert abs(it - itr) <= 1
    np.testing.assert_allclose(C, C.T, atol=1e-12)
    assert gpu_pkg.cubic_leakage(C) < 1e-8
    assert _rel(u.cpu().numpy(), ur) < 1e-6


def test_uniform_solid_is_isotropic_material(gpu_pkg):
    lh = gpu_pkg.LinearHomogenization((4, 4, 4), E=2.0, nu=0.25)
    C, _, it = lh.C_full(np.ones((4, 4, 4)))
    np.testing.assert_allclose(np.asarray(C), gpu_pkg.iso_D(2.0, 0.25), atol=1e-10)
    assert it == 0


def test_laminate_closed_form(gpu_pkg):
    rng = np.random.default_rng(11)
    sx = rng.uniform(0.05, 1.0, 12)
    rho = np.broadcast_to(sx[:, None, None], (12, 8, 6)).copy()
    lh = gpu_pkg.LinearHomogenization(rho.shape, emin=0.0, p=1.0, tol=1e-12)
    C, _, _ = lh.C_full(rho)
    assert _rel(np.asarray(C), O.laminate_C(sx)) < 1e-10


def test_warm_start_and_invariants(gpu_pkg):
    rho = gpu_pkg.cells.exact_symmetrize(_grey(gpu_pkg, 8, seed=2))
    lh = gpu_pkg.LinearHomogenization(rho.shape, tol=1e-10)
    C, u, it = lh.C_full(rho)
    C2, u2, it2 = lh.C_full(rho, u0=u)
    assert it2 < it
    assert _rel(np.asarray(C2), np.asarray(C)) < 1e-12
    c, uu, itc = lh.invariants(rho)
    np.testing.assert_allclose(np.asarray(c.detach()), gpu_pkg.cubic_invariants(C), rtol=1e-8)


def test_plane_strain_equals_extruded_3d(gpu_pkg):
    rng = np.random.default_rng(3)
    base = rng.standard_normal((8, 8))
    base = 0.5 * (base + base.T)
    rho2 = 1.0 / (1.0 + np.exp(-3.0 * base))
    c2, _, _ = gpu_pkg.LinearHomogenization(rho2.shape, tol=1e-11).invariants(rho2)
    rho3 = np.repeat(rho2[..., None], 3, axis=2)
    C3, _, _ = gpu_pkg.LinearHomogenization(rho3.shape, tol=1e-11).C_full(rho3)
    C3 = np.asarray(C3)
    np.testing.assert_allclose(np.asarray(c2.detach()), [C3[0, 0], C3[1, 0], C3[5, 5]], rtol=1e-8)


def test_hashin_shtrikman_bounds(gpu_pkg):
    rho = (gpu_pkg.cells.schwarz_cell(12, beta=64.0, v=0.3) > 0.5).astype(float)
    C, _, _ = gpu_pkg.LinearHomogenization(rho.shape).C_full(rho)
    c11, c12, c44 = gpu_pkg.cubic_invariants(C)
    K = (c11 + 2 * c12) / 3
    Khs, Ghs = gpu_pkg.hashin_shtrikman(rho.mean())
    assert K < Khs and c44 < Ghs


def test_maxit_exhaustion_is_reported(gpu_pkg):
    rho = gpu_pkg.cells.exact_symmetrize(_grey(gpu_pkg, 8, seed=5))
    lh = gpu_pkg.LinearHomogenization(rho.shape, maxit=3)
    _, _, it = lh.C_full(rho)
    assert it == 3
    assert lh.last_info["converged"] is False


def test_config0_schwarz_cell_golden(gpu_pkg):
    """32^3 Schwarz-P cell, emin 1e-4, FP64 PCG to 1e-8: C within 1e-6 of the stored oracle C."""
    g = np.load(os.path.join(GOLDEN, "schwarz32.npz"))
    lh = gpu_pkg.LinearHomogenization(g["rho"].shape, emin=float(g["emin"]), tol=1e-8)
    C, _, it = lh.C_full(g["rho"])
    err = _rel(np.asarray(C), g["C"])
    print(f"config0 32^3: it={it} (oracle {int(g['it'])}) max|dC|/max|C|={err:.3e}", flush=True)
    assert err < 1e-6
    assert abs(it - int(g["it"])) <= 2

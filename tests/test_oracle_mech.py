"""Pins of the finite-strain oracle (oracle/mech_ref.py) against what the mathematics fixes: central
differences of its energies and stresses, the small-strain limit of the neo-Hookean tangent, and the
linear material against the linear oracle's homogenized stiffness (nlfiber/mech.py:1-17 states that
the finite-strain cell reduces to the linear C^H in the small-strain limit)."""
import numpy as np
import pytest

from oracle import linear_ref as LO
from oracle import mech_ref as M


def _rand_F(rng, n, amp=0.15):
    return np.eye(3) + amp * rng.standard_normal((n, 3, 3))


@pytest.mark.parametrize("material", ["neohooke", "linear"])
@pytest.mark.parametrize("wls", [True, False])
def test_material_stress_and_tangent_are_derivatives(material, wls):
    rng = np.random.default_rng(3)
    lam, mu = LO.lame(1.0, 0.3)
    mat = M.MatRef(material, lam, mu)
    F = _rand_F(rng, 5)
    gam = rng.uniform(0.05, 1.0, 5) if wls else None
    psi, P, A = mat.eval(F, gam)
    h = 1e-6
    for i in range(3):
        for j in range(3):
            d = np.zeros((3, 3))
            d[i, j] = h
            pp, Pp, _ = mat.eval(F + d, gam)
            pm, Pm, _ = mat.eval(F - d, gam)
            np.testing.assert_allclose(P[:, i, j], (pp - pm) / (2 * h), rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(A[:, :, :, i, j], (Pp - Pm) / (2 * h), rtol=1e-6, atol=1e-8)
    # major symmetry of a hyperelastic tangent
    np.testing.assert_allclose(A, np.transpose(A, (0, 3, 4, 1, 2)), atol=1e-12)


def test_neohooke_tangent_at_identity_is_linear():
    lam, mu = LO.lame(1.0, 0.3)
    A = M.A_neo(np.eye(3)[None], lam, mu)[0]
    np.testing.assert_allclose(A, M.A_lin(lam, mu), atol=1e-14)
    assert abs(M.psi_neo(np.eye(3), lam, mu)) < 1e-15


@pytest.mark.parametrize("material,wls", [("neohooke", (500.0, 0.01)), ("neohooke", None), ("linear", None)])
def test_cell_gradient_and_hessian_product_are_derivatives(material, wls):
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.0, 1.0, (3, 4, 2))
    cell = M.CellRef(rho, material=material, wls=wls)
    w = 0.05 * rng.standard_normal(cell.shape + (3,))
    Fb = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    gw, P = cell.stress(w, Fb)
    dw = rng.standard_normal(w.shape)
    dF = rng.standard_normal((3, 3))
    h = 1e-6
    num = (cell.energy(w + h * dw, Fb + h * dF) - cell.energy(w - h * dw, Fb - h * dF)) / (2 * h)
    assert num == pytest.approx(np.sum(gw * dw) + np.sum(P * dF), rel=1e-6)
    A = cell.tangent(w, Fb)
    hw, hP = cell.hvp(A, dw, dF)
    gp, Pp = cell.stress(w + h * dw, Fb + h * dF)
    gm, Pm = cell.stress(w - h * dw, Fb - h * dF)
    np.testing.assert_allclose(hw, (gp - gm) / (2 * h), rtol=1e-5, atol=1e-9 * np.abs(hw).max())
    np.testing.assert_allclose(hP, (Pp - Pm) / (2 * h), rtol=1e-5, atol=1e-9 * np.abs(hP).max())


def test_linear_material_energy_at_corrector_is_half_C():
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.0, 1.0, (4, 4, 4))
    lin = LO.LinearRef(rho.shape, tol=1e-12, maxit=4000)
    C, u, _ = lin.C_full(rho)
    cell = M.CellRef(rho, material="linear", wls=None)
    for I in range(6):
        H = np.zeros((3, 3))
        i, j = LO.VOIGT[I]
        if i == j:
            H[i, i] = 1.0
        else:
            H[i, j] = H[j, i] = 0.5
        Fb = np.eye(3) + H
        assert cell.energy(u[I], Fb) == pytest.approx(0.5 * C[I, I], rel=1e-10)
        g, _ = cell.stress(u[I], Fb)
        g0, _ = cell.stress(np.zeros_like(u[I]), Fb)
        assert np.linalg.norm(g) < 1e-8 * np.linalg.norm(g0)


def test_dense_newton_small_strain_matches_linear_C():
    """Full control at Fbar = I + t H: Pbar = t C:H + O(t^2) for the neo-Hookean cell (mech.py:1-17)."""
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.2, 1.0, (3, 3, 3))
    lin = LO.LinearRef(rho.shape, tol=1e-12, maxit=4000)
    C, _, _ = lin.C_full(rho)
    cell = M.CellRef(rho, material="neohooke", wls=(500.0, 0.01))
    n = np.array([1.0, -0.3, 0.2, 0.4, -0.2, 0.1])
    H = np.zeros((3, 3))
    for I, (i, j) in enumerate(LO.VOIGT):
        if i == j:
            H[i, i] += n[I]
        else:
            H[i, j] += 0.5 * n[I]
            H[j, i] += 0.5 * n[I]
    errs = []
    for t in (1e-3, 5e-4):
        w, Fb = M.newton_dense(cell, np.eye(3) + t * H)
        _, P = cell.stress(w, Fb)
        Pv = np.array([P[i, j] if i == j else 0.5 * (P[i, j] + P[j, i]) for i, j in LO.VOIGT])
        errs.append(np.abs(Pv - t * C @ n).max() / (t * np.abs(C @ n).max()))
    assert errs[0] < 5e-3 and errs[1] < 0.6 * errs[0]  # first-order convergence in t


@pytest.mark.parametrize("material,wls", [("neohooke", (500.0, 0.01)), ("neohooke", (20.0, 0.3)), ("linear", None)])
def test_density_sensitivity_is_the_rho_derivative(material, wls):
    rng = np.random.default_rng(12)
    rho = rng.uniform(0.05, 0.95, (3, 3, 2))
    w = 0.05 * rng.standard_normal(rho.shape + (3,))
    v = rng.standard_normal(rho.shape + (3,))
    Fb = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    dF = rng.standard_normal((3, 3))
    g = M.CellRef(rho, material=material, wls=wls).rho_sensitivity(w, Fb, v, dF)
    d = rng.standard_normal(rho.shape)
    h = 1e-6
    num = (M.CellRef(rho + h * d, material=material, wls=wls).work(w, Fb, v, dF) -
           M.CellRef(rho - h * d, material=material, wls=wls).work(w, Fb, v, dF)) / (2 * h)
    assert np.sum(g * d) == pytest.approx(num, rel=1e-6, abs=1e-12)


def test_toec_derivatives_second_order_is_linear_C_and_symmetric():
    """toec's second derivative equals the linear C at the linear correctors (tests/test_mech.py:36-41);
    the third derivative, from central differences, is symmetric in all three indices."""
    rng = np.random.default_rng(17)
    rho = rng.uniform(0.2, 1.0, (3, 3, 3))
    lin = LO.LinearRef(rho.shape, tol=1e-12, maxit=4000)
    C, chi, _ = lin.C_full(rho)
    C2, T3 = M.toec_derivatives(M.CellRef(rho), chi)
    np.testing.assert_allclose(C2, C, atol=1e-9 * np.abs(C).max())
    np.testing.assert_allclose(T3, np.transpose(T3, (1, 0, 2)), atol=1e-7 * np.abs(T3).max())
    np.testing.assert_allclose(T3, np.transpose(T3, (0, 2, 1)), atol=1e-7 * np.abs(T3).max())


@pytest.mark.parametrize("material,wls", [("neohooke", (500.0, 0.01)), ("neohooke", None), ("linear", None)])
def test_toec_closed_form_of_the_backward_matches_oracle(pkg, material, wls):
    """The torch closed form that toec's backward differentiates (mech.toec_values_t; CPU here) against the
    oracle's second derivatives (explicit tangent) and third derivatives (central differences of it)."""
    import types

    import torch
    rng = np.random.default_rng(21)
    rho = rng.uniform(0.1, 1.0, (4, 5, 3))
    chi = 0.3 * rng.standard_normal((6,) + rho.shape + (3,))
    C2, T3 = M.toec_derivatives(M.CellRef(rho, material=material, wls=wls), chi)
    lam, mu = pkg.mech.lame(1.0, 0.3)
    cell = types.SimpleNamespace(shape=rho.shape, lam=lam, mu=mu, material=material)
    s = 1e-3 + (1 - 1e-3) * rho ** 3
    gam = np.ones_like(rho)
    if wls is not None:
        b, r0 = wls
        gam = (np.tanh(b * r0) + np.tanh(b * (rho ** 3 - r0))) / (np.tanh(b * r0) + np.tanh(b * (1 - r0)))
    v = pkg.mech.toec_values_t(cell, torch.tensor(chi).reshape(6, -1), torch.tensor(s / (8 * rho.size)).reshape(-1),
                               torch.tensor(gam).reshape(-1)).numpy()
    C, T = pkg.mech._toec_assemble_t(torch.tensor(v))
    assert np.abs(C.numpy() - C2).max() < 1e-13 * np.abs(C2).max()
    d3 = np.zeros((6, 6, 6))
    for q, (I, J, K) in enumerate(pkg.mech._TRIPLES):
        for t in {(I, J, K), (I, K, J), (J, I, K), (J, K, I), (K, I, J), (K, J, I)}:
            d3[t] = v[21 + q]
    assert np.abs(d3 - T3).max() <= 1e-6 * max(np.abs(T3).max(), 1e-300)

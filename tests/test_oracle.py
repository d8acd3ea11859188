"""Pins of the FP64 oracle (CPU only).

Each check is independent of the formula it checks: an exact tensor-product assembly of Ke, the
rigid-body null space, brute-force dense assembly of K, a dense direct solve, closed forms
(uniform solid, laminate), bounds (Reuss/Voigt, Hashin-Shtrikman), symmetry and scaling laws, the
preconditioner as the exact inverse of K(1), central differences for the two derivatives, and the
reference's own test properties (tests/test_core.py:77-142 of the reference). A planted mistake in
the oracle (a dropped term, a wrong sign or index, a transposed operand) fails at least one.
"""
import os

import numpy as np
import pytest

from oracle import linear_ref as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_Ke_equals_tensor_product_assembly_and_null_space():
    for E, nu in ((1.0, 0.3), (2.0, 0.25), (0.7, 0.45)):
        Ke, G, D = O.element_matrices(E, nu)
        Kt = O.Ke_tensor_product(E, nu)
        assert np.abs(Ke - Kt).max() < 1e-14 * np.abs(Kt).max()
        w = np.linalg.eigvalsh(Ke)
        assert (w > 1e-10).sum() == 18  # 6 rigid modes
        X = O.NODES.astype(float)
        for c in range(3):  # translations
            t = np.zeros(24)
            t[c::3] = 1.0
            assert np.abs(Ke @ t).max() < 1e-14
        for a, b in ((0, 1), (0, 2), (1, 2)):  # infinitesimal rotations
            r = np.zeros(24)
            r[a::3] = -X[:, b]
            r[b::3] = X[:, a]
            assert np.abs(Ke @ r).max() < 1e-14
        # G = D times the mean strain-displacement operator: each row sums to zero over nodes
        assert np.abs(G.reshape(6, 8, 3).sum(axis=1)).max() < 1e-15


def test_apply_K_equals_dense_assembly_and_load():
    rng = np.random.default_rng(0)
    for shape in ((3, 4, 5), (2, 3, 2), (4, 4, 1)):
        s = rng.uniform(0.1, 1.0, shape)
        lr = O.LinearRef(shape)
        K = O.dense_K(shape, s, O.Ke_tensor_product())
        u = rng.standard_normal(shape + (3,))
        assert np.abs(lr.apply_K(s, u).ravel() - K @ u.ravel()).max() < 1e-13
        assert np.abs(K - K.T).max() < 1e-14
        assert np.abs(lr.apply_K(s, np.ones(shape + (3,)))).max() < 1e-13  # K const = 0
        f = lr.load(s, np.eye(6))
        assert np.abs(f.sum(axis=(1, 2, 3))).max() < 1e-13  # loads are mean-free


def test_uniform_solid_is_isotropic_material():
    C, _, it = O.LinearRef((4, 4, 4), E=2.0, nu=0.25).C_full(np.ones((4, 4, 4)))
    np.testing.assert_allclose(C, O.iso_D(2.0, 0.25), atol=1e-12)
    assert it == 0


def test_precond_is_exact_inverse_of_reference_medium():
    rng = np.random.default_rng(1)
    for shape in ((4, 6, 5), (8, 8, 8), (6, 6, 1)):
        lr = O.LinearRef(shape)
        assert np.abs(lr.Kinv.imag).max() < 1e-13 * np.abs(lr.Kinv.real).max()
        assert np.abs(lr.Kinv - np.swapaxes(lr.Kinv, -1, -2)).max() < 1e-13
        u = O.proj(rng.standard_normal((1,) + shape + (3,)))
        Ku = np.stack([lr.apply_K(np.ones(shape), u[0])])
        assert np.abs(lr.precond(Ku) - u).max() < 1e-9 * np.abs(u).max()
    # PCG converges in one iteration for a constant stiffness
    lr = O.LinearRef((6, 6, 6), tol=1e-12)
    _, _, it = lr.C_full(np.full((6, 6, 6), 0.7))
    assert it <= 1


def test_laminate_closed_form_and_dense_solve():
    rng = np.random.default_rng(2)
    sx = rng.uniform(0.05, 1.0, 6)
    s = np.broadcast_to(sx[:, None, None], (6, 3, 4)).copy()
    lr = O.LinearRef(s.shape, emin=0.0, p=1.0, tol=1e-13)
    u, _ = lr.solve(s, np.eye(6))
    assert np.abs(lr.bilinear(s, np.eye(6), u) - O.laminate_C(sx)).max() < 1e-12
    s = rng.uniform(0.01, 1.0, (3, 4, 3))
    lr = O.LinearRef(s.shape, emin=0.0, p=1.0, tol=1e-13)
    u, _ = lr.solve(s, np.eye(6))
    C = lr.bilinear(s, np.eye(6), u)
    Cd, _ = O.dense_C(s.shape, s)
    assert np.abs(C - Cd).max() < 1e-12


def test_bounds_symmetry_scaling_and_shift():
    rng = np.random.default_rng(3)
    s = rng.uniform(0.05, 1.0, (5, 4, 6))
    lr = O.LinearRef(s.shape, emin=0.0, p=1.0, tol=1e-12)
    C = lr.bilinear(s, np.eye(6), lr.solve(s, np.eye(6))[0])
    D = O.iso_D(1.0, 0.3)
    np.testing.assert_allclose(C, C.T, atol=1e-12)
    assert np.linalg.eigvalsh(np.mean(s) * D - C).min() > -1e-12          # Voigt
    assert np.linalg.eigvalsh(C - D / np.mean(1.0 / s)).min() > -1e-12     # Reuss
    C2 = lr.bilinear(2.5 * s, np.eye(6), lr.solve(2.5 * s, np.eye(6))[0])
    assert np.abs(C2 - 2.5 * C).max() < 1e-11
    ss = np.roll(s, (2, 1, 3), axis=(0, 1, 2))
    C3 = lr.bilinear(ss, np.eye(6), lr.solve(ss, np.eye(6))[0])
    assert np.abs(C3 - C).max() < 1e-11
    # axis permutation x <-> y swaps the Voigt indices accordingly
    st = np.transpose(s, (1, 0, 2))
    lt = O.LinearRef(st.shape, emin=0.0, p=1.0, tol=1e-12)
    Ct = lt.bilinear(st, np.eye(6), lt.solve(st, np.eye(6))[0])
    P = [1, 0, 2, 4, 3, 5]
    assert np.abs(Ct - C[np.ix_(P, P)]).max() < 1e-11


def test_cubic_cell_properties_and_hs_bound():
    g = np.load(os.path.join(GOLDEN, "grey8.npz"))
    lr = O.LinearRef(g["rho"].shape, emin=float(g["emin"]), tol=1e-11)
    C, u, it = lr.C_full(g["rho"])
    assert np.abs(C - g["C"]).max() < 1e-12 * np.abs(C).max()
    assert O.cubic_leakage(C) < 1e-8
    c, u2, _ = lr.invariants(g["rho"])
    np.testing.assert_allclose(c, O.cubic_invariants(C), rtol=1e-8)
    b = (g["rho"] > 0.5).astype(float)
    Cb, _, _ = O.LinearRef(b.shape, tol=1e-10).C_full(b)
    c11, c12, c44 = O.cubic_invariants(Cb)
    Khs, Ghs = O.hashin_shtrikman(b.mean())
    assert (c11 + 2 * c12) / 3 < Khs and c44 < Ghs


def test_golden_operator_outputs_reproduce():
    g = np.load(os.path.join(GOLDEN, "grey8.npz"))
    lr = O.LinearRef(g["rho"].shape, emin=float(g["emin"]))
    s = lr.stiffness_scale(g["rho"])
    Kw = np.stack([lr.apply_K(s, x) for x in g["w"]])
    assert np.abs(Kw - g["Kw"]).max() < 1e-13 * np.abs(g["Kw"]).max()
    assert np.abs(lr.precond(g["w"]) - g["Mw"]).max() < 1e-13 * np.abs(g["Mw"]).max()


def test_invariants_gradient_fd():
    g = np.load(os.path.join(GOLDEN, "grey8.npz"))
    rho = g["rho"]
    lr = O.LinearRef(rho.shape, tol=1e-12)
    c, u2, _ = lr.invariants(rho)
    J = lr.invariants_grad(rho, u2)
    d = np.random.default_rng(5).standard_normal(rho.shape)
    h = 1e-5
    fd = (lr.invariants(rho + h * d)[0] - lr.invariants(rho - h * d)[0]) / (2 * h)
    np.testing.assert_allclose(np.tensordot(J, d, 3), fd, rtol=1e-5, atol=1e-10)


def test_corrector_vjp_fd():
    g = np.load(os.path.join(GOLDEN, "grey8.npz"))
    lr = O.LinearRef(g["rho"].shape, tol=1e-12)
    s0 = lr.stiffness_scale(g["rho"])
    Es = np.eye(6)[:2]
    f = lambda s: float(np.sum(lr.solve(s, Es)[0] ** 2))  # noqa: E731
    u, _ = lr.solve(s0, Es)
    gr = lr.correctors_vjp(s0, Es, u, 2 * u)
    d = np.random.default_rng(6).standard_normal(s0.shape)
    h = 1e-6
    fd = (f(s0 + h * d) - f(s0 - h * d)) / (2 * h)
    assert abs(float(np.vdot(gr, d)) - fd) < 1e-4 * abs(fd)


def test_float32_recursion_option_converges():
    g = np.load(os.path.join(GOLDEN, "grey8.npz"))
    lr = O.LinearRef(g["rho"].shape, tol=1e-5)
    s = lr.stiffness_scale(g["rho"])
    b = -lr.load(s, np.eye(6))
    x, it = lr.cg(s, b, np.zeros_like(b), dtype=np.float32)
    r = b - np.stack([lr.apply_K(s, xi) for xi in x])
    assert np.sqrt((r ** 2).sum(axis=(1, 2, 3, 4))).max() / np.sqrt((b ** 2).sum(axis=(1, 2, 3, 4))).min() < 1e-4


@pytest.mark.parametrize("n", [8, 12])
def test_cell_generator_volume_and_symmetry(pkg, n):
    rho, meta = pkg.cells.random_cell(n, seed=n)
    assert abs(rho.mean() - meta["v"]) < 1e-12
    for p in ((1, 0, 2), (2, 1, 0), (0, 2, 1)):
        assert np.array_equal(rho, np.transpose(rho, p))
    assert np.array_equal(rho, rho[::-1]) and np.array_equal(rho, rho[:, ::-1]) and np.array_equal(rho, rho[:, :, ::-1])
    r0 = pkg.cells.schwarz_cell(n)
    assert abs(r0.mean() - 0.3) < 1e-12


# --- density map (nlfiber/density.py, nlfiber/symmetry.py) -----------------------------------------
from oracle import density_ref as DR  # noqa: E402


def test_density_symmetry_helpers():
    """mirror/fold round trip, sym_perm an orthogonal projection, the mirrored cell has every symmetry
    (the reference's tests/test_core.py:20-41), and exact_symmetrize equals the 48-map group average."""
    rng = np.random.default_rng(0)
    x = DR.sym_perm(rng.standard_normal((5, 5, 5)))
    full = DR.mirror(x)
    np.testing.assert_allclose(DR.fold(full), x, atol=1e-14)
    y = rng.standard_normal((5, 5, 5))
    assert abs(np.vdot(DR.sym_perm(y), x) - np.vdot(y, x)) < 1e-10
    f = DR.mirror(DR.sym_perm(rng.standard_normal((4, 4, 4))))
    np.testing.assert_allclose(f, f[::-1], atol=0)
    np.testing.assert_allclose(f, np.transpose(f, (2, 0, 1)), atol=0)
    g = rng.standard_normal((8, 8, 8))
    e = DR.exact_symmetrize(g)
    np.testing.assert_allclose(e, DR.group_average(g), atol=1e-14)
    assert np.array_equal(e, e[::-1]) and np.array_equal(e, np.transpose(e, (1, 2, 0)))
    # the group average is self-adjoint
    h = rng.standard_normal((8, 8, 8))
    assert abs(np.vdot(DR.group_average(g), h) - np.vdot(g, DR.group_average(h))) < 1e-10


def test_density_filter_is_the_periodic_hat_stencil():
    """The rfftn filter equals the brute-force periodic convolution with the hat kernel."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 6, 6))
    for r in (1.5, 2.5):
        k = DR.hat_kernel(x.shape, r)
        want = np.zeros_like(x)
        for i, j, l in zip(*np.nonzero(k)):
            want += k[i, j, l] * np.roll(x, (i, j, l), axis=(0, 1, 2))
        np.testing.assert_allclose(DR.filt(x, r), want, atol=1e-13)


def test_density_exact_volume_and_vjp():
    """mean rho = v for every beta and the adjoint against central differences of sum(rho^3)
    (the reference's tests/test_core.py:43-53)."""
    rng = np.random.default_rng(2)
    phi = DR.sym_perm(rng.standard_normal((4, 4, 4)))
    for beta in (2.0, 8.0, 32.0):
        assert abs(DR.rho(phi, beta, 1.5, 0.27).mean() - 0.27) < 1e-12
    beta = 4.0
    r = DR.rho(phi, beta, 1.5, 0.27)
    g = DR.rho_vjp(phi, beta, 1.5, 0.27, 3 * r ** 2)[0]
    d = DR.sym_perm(rng.standard_normal((4, 4, 4)))
    h = 1e-6
    fd = (np.sum(DR.rho(phi + h * d, beta, 1.5, 0.27) ** 3) - np.sum(DR.rho(phi - h * d, beta, 1.5, 0.27) ** 3)) / (2 * h)
    assert abs(np.vdot(g, d) - fd) < 1e-7 * abs(fd) + 1e-10
    # a non-symmetric direction too (the map symmetrises its input)
    d2 = rng.standard_normal((4, 4, 4))
    fd2 = (np.sum(DR.rho(phi + h * d2, beta, 1.5, 0.27) ** 3) - np.sum(DR.rho(phi - h * d2, beta, 1.5, 0.27) ** 3)) / (2 * h)
    assert abs(np.vdot(g, d2) - fd2) < 1e-7 * abs(fd2) + 1e-10


def test_jacobi_diagonal_and_spectral_bound():
    """jacobi_diag is diag K(s) of the dense assembly (every component), and jacobi_bound = lambda_max(Ke) / Ke_00
    dominates the spectrum of D^-1 K(s) for random, high-contrast and single-stiff-voxel fields (the last comes
    close to it: the Rayleigh quotient of an element's top mode in a stiff voxel surrounded by soft ones)."""
    rng = np.random.default_rng(3)
    Ke = O.Ke_tensor_product()
    bound = O.jacobi_bound()
    assert abs(bound - np.linalg.eigvalsh(Ke).max() / Ke[0, 0]) < 1e-12
    tops = []
    for shape, kind in (((3, 4, 3), "uniform01"), ((4, 4, 2), "contrast"), ((4, 4, 4), "voxel")):
        if kind == "uniform01":
            s = rng.uniform(0.1, 1.0, shape)
        elif kind == "contrast":
            s = np.where(rng.uniform(size=shape) < 0.5, 1.0, 1e-4)
        else:
            s = np.full(shape, 1e-6)
            s[1, 2, 1] = 1.0
        K = O.dense_K(shape, s, Ke)
        d = O.jacobi_diag(s).reshape(-1)
        np.testing.assert_allclose(np.diag(K).reshape(-1, 3), np.repeat(d[:, None], 3, axis=1), rtol=1e-13)
        dh = 1.0 / np.sqrt(np.repeat(d, 3))
        lam = np.linalg.eigvalsh(dh[:, None] * K * dh[None, :])
        assert lam.min() > -1e-12 and lam.max() <= bound * (1 + 1e-12)
        tops.append(lam.max())
    assert tops[2] > 0.9 * bound  # nearly attained by an isolated stiff voxel


def test_cheb_smooth_is_the_chebyshev_residual_polynomial():
    """The three-term recurrence of cheb_smooth against the closed form: the error after k steps is
    P_k(D^-1 K) e_0 with P_k(t) = T_k((theta - t) / delta) / T_k(theta / delta), evaluated through the
    eigendecomposition of D^-1/2 K D^-1/2 of the dense assembly (e = x - u* for any solution u* of K u = -f)."""
    rng = np.random.default_rng(4)
    shape = (3, 3, 2)
    s = rng.uniform(0.05, 1.0, shape)
    ref = O.LinearRef(shape)
    Es = np.stack([O.unit_strain(0), O.unit_strain(4)])
    u0 = rng.standard_normal((2,) + shape + (3,))
    K = O.dense_K(shape, s, O.Ke_tensor_product())
    d = np.repeat(O.jacobi_diag(s).reshape(-1), 3)
    w, Q = np.linalg.eigh(K / np.sqrt(d)[:, None] / np.sqrt(d)[None, :])
    f = ref.load(s, Es).reshape(2, -1)
    ustar = np.stack([np.linalg.lstsq(K, -fi, rcond=None)[0] for fi in f])
    hi = O.jacobi_bound()
    for degree, lo in ((1, hi / 10), (5, hi / 30), (17, hi / 100)):
        x = O.cheb_smooth(ref, s, Es, u0, degree, lo, hi).reshape(2, -1)
        th, de = 0.5 * (hi + lo), 0.5 * (hi - lo)
        T = np.polynomial.chebyshev.Chebyshev.basis(degree)
        P = T((th - w) / de) / T(th / de)
        for I in range(2):
            e0 = u0[I].reshape(-1) - ustar[I]
            ek = (Q @ (P * (Q.T @ (np.sqrt(d) * e0)))) / np.sqrt(d)
            np.testing.assert_allclose(x[I] - ustar[I], ek, rtol=0, atol=1e-10 * np.abs(e0).max())
    # on [lo, hi] the polynomial is at most 1 / T_k(theta / delta): the energy error of high modes falls
    x = O.cheb_smooth(ref, s, Es, u0, 17, hi / 100, hi).reshape(2, -1)
    e0, ek = u0[0].reshape(-1) - ustar[0], x[0] - ustar[0]
    hi_modes = w > hi / 100
    c0 = (Q.T @ (np.sqrt(d) * e0))[hi_modes]
    ck = (Q.T @ (np.sqrt(d) * ek))[hi_modes]
    assert np.linalg.norm(ck) < 0.5 * np.linalg.norm(c0)


def test_network_reference_gradients_match_central_differences():
    """The float64 network reference (oracle/convnet_ref.py, unrounded) that the GPU gradients are checked
    against: its autograd gradients of the MSE loss along random directions of every layer's weights and
    biases equal central differences of the loss (h = 1e-6)."""
    import torch
    from oracle import convnet_ref as R
    g = torch.Generator().manual_seed(0)
    params = {}
    for name, ca, ah, cb, cout, coutp, act, lvl in R.LAYERS:
        w = torch.randn((27, coutp, ca + cb), generator=g, dtype=torch.float64) * (2.0 / (27 * (ca + cb))) ** 0.5
        if name == "out":
            w[:, 18:, :] = 0.0
        params[name] = w.requires_grad_(True)
        params[name + ".b"] = (0.1 * torch.randn(cout, generator=g, dtype=torch.float64)).requires_grad_(True)
    rho = torch.rand((1, 8, 8, 8), generator=g, dtype=torch.float64)
    target = torch.randn((1, 8, 8, 8, 18), generator=g, dtype=torch.float64)
    L = R.loss(R.forward(params, rho, bf16=False), target)
    grads = torch.autograd.grad(L, list(params.values()))
    h = 1e-6
    for (k, p), gk in zip(params.items(), grads):
        d = torch.randn(p.shape, generator=g, dtype=torch.float64)
        if k == "out":
            d[:, 18:, :] = 0.0
        with torch.no_grad():
            pp = {kk: (vv + h * d if kk == k else vv) for kk, vv in params.items()}
            pm = {kk: (vv - h * d if kk == k else vv) for kk, vv in params.items()}
            fd = (R.loss(R.forward(pp, rho, bf16=False), target) - R.loss(R.forward(pm, rho, bf16=False), target)) / (2 * h)
        an = float((gk * d).sum())
        assert abs(an - float(fd)) <= 1e-5 * max(abs(an), 1e-12) + 1e-12, (k, an, float(fd))

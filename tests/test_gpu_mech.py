"""Finite-strain cell operators and Newton-CG on the GPU (nlfh_fs_eval, package mech.Cell / mech.Mixed)
against the oracle (oracle/mech_ref.py, pinned in tests/test_oracle_mech.py)."""
import numpy as np
import pytest

from oracle import linear_ref as LO
from oracle import mech_ref as MR

pytestmark = pytest.mark.gpu

CASES = [("neohooke", (500.0, 0.01)), ("neohooke", None), ("linear", (500.0, 0.01)), ("linear", None)]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def M(gpu_pkg):
    import torch
    torch.cuda.set_device(0)
    return gpu_pkg.mech


@pytest.mark.parametrize("shape", [(5, 4, 3), (8, 8, 8), (6, 6, 1)])
@pytest.mark.parametrize("material,wls", CASES)
def test_energy_gradient_and_hessian_product_match_oracle(M, shape, material, wls):
    import torch
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.0, 1.0, shape)
    ref = MR.CellRef(rho, material=material, wls=wls)
    cell = M.Cell(rho, material=material, wls=wls)
    w = 0.05 * rng.standard_normal(shape + (3,))
    Fb = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    wd = torch.as_tensor(w, device="cuda")
    print(f"fs_eval shape={shape} material={material} wls={wls}", flush=True)
    assert cell.energy(wd, Fb) == pytest.approx(ref.energy(w, Fb), rel=1e-12)
    g, P = cell.gradient(wd, Fb)
    gr, Pr = ref.stress(w, Fb)
    assert _rel(g.cpu().numpy(), gr) < 1e-11
    assert _rel(P, Pr) < 1e-12
    v = rng.standard_normal(shape + (3,))
    dF = rng.standard_normal((3, 3))
    hw, hP = cell.hessian_product(wd, Fb, torch.as_tensor(v, device="cuda"), dF)
    hwr, hPr = ref.hvp(ref.tangent(w, Fb), v, dF)
    assert _rel(hw.cpu().numpy(), hwr) < 1e-11
    assert _rel(hP, hPr) < 1e-11
    # a null direction leaves only the macro part
    hw0, hP0 = cell.hessian_product(wd, Fb, None, dF)
    hwr0, hPr0 = ref.hvp(ref.tangent(w, Fb), np.zeros_like(v), dF)
    assert _rel(hw0.cpu().numpy(), hwr0) < 1e-11 and _rel(hP0, hPr0) < 1e-11


def test_precond_is_N_times_the_linear_preconditioner(M):
    import torch
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.0, 1.0, (6, 5, 4))
    cell = M.Cell(rho)
    r = rng.standard_normal((6, 5, 4, 3))
    z = cell.precond_w(torch.as_tensor(r, device="cuda")).cpu().numpy()
    lin = LO.LinearRef((6, 5, 4))
    zr = lin.precond(r[None])[0] * cell.N
    assert _rel(z, zr) < 1e-12


@pytest.mark.parametrize("free", [[], [(1, 1), (2, 2)]])
def test_newton_cg_reaches_the_oracle_equilibrium(M, free):
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.2, 1.0, (4, 4, 4))
    ref = MR.CellRef(rho)
    cell = M.Cell(rho)
    Fb = np.diag([1.03, 1.0, 1.0]) + 0.01 * np.array([[0, 1, 0], [0, 0, 0.5], [0.2, 0, 0]])
    print(f"newton free={free}", flush=True)
    w, F, log = M.Mixed(cell, free=free).solve(Fb, rtol=1e-11, cg_tol=1e-10)
    assert log["converged"] and not log.get("linesearch_failed")
    wr, Fr = MR.newton_dense(ref, Fb, free=free, rtol=1e-12)
    np.testing.assert_allclose(F, Fr, atol=1e-10)
    P = cell.mean_stress(w, F)
    _, Pr = ref.stress(wr, Fr)
    assert _rel(P, Pr) < 1e-8
    for a, b in free:  # zero conjugate stress on the free components
        assert abs(P[a, b]) < 1e-8 * np.abs(P).max()
    assert _rel(w.cpu().numpy(), wr) < 1e-7


def test_small_strain_limit_is_the_linear_C(M, gpu_pkg):
    """Pbar(I + t H) = t C:H + O(t^2) with C from the linear GPU solve (mech.py:1-17)."""
    import torch
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.2, 1.0, (8, 8, 8))
    lh = gpu_pkg.LinearHomogenization((8, 8, 8), tol=1e-10)
    C, _, _ = lh.C_full(torch.as_tensor(rho, device="cuda"))
    C = C.cpu().numpy()
    n = np.array([1.0, -0.3, 0.2, 0.4, -0.2, 0.1])
    H = sum(n[I] * M.voigt_tensor(I) for I in range(6))
    cell = M.Cell(rho)
    full = M.Mixed(cell, free=[])
    errs = []
    for t in (1e-3, 5e-4):
        w, F, log = full.solve(np.eye(3) + t * H, rtol=1e-11, cg_tol=1e-10)
        assert log["converged"]
        Pv = M.to_voigt(cell.mean_stress(w, F))
        Pv[3:] *= 0.5  # engineering shear of the stress vector -> tensor components
        errs.append(_rel(Pv, t * C @ n))
    assert errs[0] < 5e-3 and errs[1] < 0.6 * errs[0]


@pytest.mark.parametrize("shape", [(7, 5, 3), (64, 64, 17), (40, 30, 45)])
def test_projection_removes_the_component_means(M, shape):
    # one block of partial sums, then grid-stride sweeps over more nodes than the partial-sum grid covers
    import torch
    rng = np.random.default_rng(9)
    cell = M.Cell(rng.uniform(0.0, 1.0, shape))
    mix = M.Mixed(cell)
    v = rng.standard_normal(3 * cell.N + 2) + 0.3
    out = mix.proj(torch.as_tensor(v, device="cuda")).cpu().numpy()
    w = v[:-2].reshape(-1, 3)
    np.testing.assert_allclose(out[:-2].reshape(-1, 3), w - w.mean(axis=0), atol=1e-13)
    assert np.abs(out[:-2].reshape(-1, 3).mean(axis=0)).max() < 1e-14
    assert out[-2:].tolist() == v[-2:].tolist()


@pytest.mark.parametrize("shape,k", [((7, 5, 3), 2), ((64, 64, 17), 0), ((40, 30, 45), 3)])
def test_cg_steps_with_folded_dot_products(M, gpu_pkg, shape, k):
    # project_dot + cg_update_parts + cg_direction_parts against the same CG step in numpy
    import ctypes
    import torch
    lib = gpu_pkg.load_library()
    rng = np.random.default_rng(21)
    N = int(np.prod(shape))
    L = 3 * N + k
    vec = lambda: rng.standard_normal(L) + 0.2  # noqa: E731
    x, r, p, Ap, z = vec(), vec(), vec(), vec(), vec()
    rz_old = 0.7
    d = {n: torch.as_tensor(v, device="cuda").clone() for n, v in dict(x=x, r=r, p=p, Ap=Ap, z=z).items()}
    npart = int(lib.nlfh_fs_dot_part_doubles())
    part = torch.empty(npart, dtype=torch.float64, device="cuda")
    shp = (ctypes.c_int * 3)(*shape)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    assert lib.nlfh_fs_project_dot(shp, P(d["Ap"]), P(d["p"]), P(part), None) == 0
    Apw = Ap[:3 * N].reshape(-1, 3)
    Ap_p = np.concatenate([(Apw - Apw.mean(axis=0)).ravel(), Ap[3 * N:]])
    np.testing.assert_allclose(d["Ap"].cpu().numpy(), Ap_p, rtol=0, atol=1e-13)
    pAp = float(p @ Ap_p)
    rz = torch.tensor([rz_old, -1.0], dtype=torch.float64, device="cuda")
    s2 = torch.zeros(2, dtype=torch.float64, device="cuda")
    x1, r1, p1 = (torch.empty(L, dtype=torch.float64, device="cuda") for _ in range(3))
    rn2p = torch.empty((L + 255) // 256, dtype=torch.float64, device="cuda")
    assert lib.nlfh_fs_cg_update_parts(L, k, P(x1), P(r1), P(d["x"]), P(d["r"]), P(d["p"]), P(d["Ap"]), P(rz[0:1]),
                                       P(part), P(s2[0:1]), P(rn2p), None) == 0
    al = rz_old / pAp if pAp > 0 else 0.0
    assert abs(s2[0].item() - pAp) <= 1e-12 * abs(pAp) + 1e-12
    np.testing.assert_allclose(x1.cpu().numpy(), x + al * p, rtol=0, atol=1e-12)
    r1_h = r - al * Ap_p
    np.testing.assert_allclose(r1.cpu().numpy(), r1_h, rtol=0, atol=1e-12)
    assert lib.nlfh_fs_project_dot(shp, P(d["z"]), P(r1), P(part), None) == 0
    zw = z[:3 * N].reshape(-1, 3)
    z_p = np.concatenate([(zw - zw.mean(axis=0)).ravel(), z[3 * N:]])
    rz1 = float(r1_h @ z_p)
    assert lib.nlfh_fs_cg_direction_parts(L, k, P(p1), P(d["z"]), P(d["p"]), P(r1), P(rz[0:1]), P(part), P(rz[1:2]),
                                          P(rn2p), P(s2[1:2]), None) == 0
    assert abs(rz[1].item() - rz1) <= 1e-12 * abs(rz1) + 1e-12
    want = z_p + (rz1 / rz_old) * p
    np.testing.assert_allclose(p1.cpu().numpy(), want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert abs(s2[1].item() - float(r1_h @ r1_h)) <= 1e-12 * float(r1_h @ r1_h)


def test_macro_scatter_and_gather(M, gpu_pkg):
    import ctypes
    import torch
    lib = gpu_pkg.load_library()
    idx = [5, 0, 8]
    ic = (ctypes.c_int * 3)(*idx)
    src = torch.tensor([1.5, -2.0, 3.25], dtype=torch.float64, device="cuda")
    dst = torch.full((9,), 7.0, dtype=torch.float64, device="cuda")
    assert lib.nlfh_fs_macro_scatter(3, ic, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                     None) == 0
    want = np.zeros(9)
    want[idx] = [1.5, -2.0, 3.25]
    assert dst.cpu().numpy().tolist() == want.tolist()
    back = torch.zeros(3, dtype=torch.float64, device="cuda")
    assert lib.nlfh_fs_macro_gather(3, ic, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(back.data_ptr()),
                                    None) == 0
    assert back.cpu().numpy().tolist() == [1.5, -2.0, 3.25]
    assert lib.nlfh_fs_macro_scatter(0, None, None, ctypes.c_void_p(dst.data_ptr()), None) == 0
    assert dst.cpu().numpy().tolist() == [0.0] * 9
    dup = (ctypes.c_int * 2)(4, 4)
    assert lib.nlfh_fs_macro_scatter(2, dup, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                     None) != 0


def test_tiled_cell_has_the_same_energy_density(M):
    """A k^3 supercell of a periodic cell under the same uniform state has the same energy per volume
    (mech.py:118-121)."""
    rng = np.random.default_rng(8)
    rho = rng.uniform(0.0, 1.0, (4, 3, 5))
    cell = M.Cell(rho)
    big = cell.tiled(2)
    assert big.shape == (8, 6, 10)
    Fb = np.eye(3) + 0.03 * rng.standard_normal((3, 3))
    w0 = np.zeros(cell.shape + (3,))
    w1 = np.zeros(big.shape + (3,))
    assert big.energy(w1, Fb) == pytest.approx(cell.energy(w0, Fb), rel=1e-12)
    np.testing.assert_allclose(big.mean_stress(w1, Fb), cell.mean_stress(w0, Fb), rtol=1e-11, atol=1e-15)
    assert cell.Kinv.shape == (4, 3, 3, 3, 3)


@pytest.mark.parametrize("material,wls", CASES)
def test_density_sensitivity_matches_oracle(M, material, wls):
    import torch
    rng = np.random.default_rng(13)
    rho = rng.uniform(0.0, 1.0, (5, 4, 3))
    ref = MR.CellRef(rho, material=material, wls=wls)
    cell = M.Cell(rho, material=material, wls=wls)
    w = 0.05 * rng.standard_normal(rho.shape + (3,))
    v = rng.standard_normal(rho.shape + (3,))
    Fb = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    dF = rng.standard_normal((3, 3))
    g = cell.rho_sensitivity(torch.as_tensor(w, device="cuda"), Fb, torch.as_tensor(v, device="cuda"), dF)
    assert _rel(g.cpu().numpy(), ref.rho_sensitivity(w, Fb, v, dF)) < 1e-11


def _fd_solves(M, rho, free, Fb, d, h, quantity):
    vals = []
    for sgn in (1.0, -1.0):
        c = M.Cell(rho + sgn * h * d)
        w, F, log = M.Mixed(c, free=free).solve(Fb, rtol=1e-12, cg_tol=1e-10)
        assert log["converged"]
        vals.append(quantity(c, w, F))
    return (vals[0] - vals[1]) / (2 * h)


def test_grad_rho_of_poisson_ratio_matches_central_differences(M):
    """The reference's test_grad_rho_of_poisson_ratio (tests/test_mech.py:95-113) on a 4^3 cell."""
    import torch
    rng = np.random.default_rng(6)
    rho = rng.uniform(0.3, 1.0, (4, 4, 4))
    free = [(1, 1), (2, 2)]
    e = 0.04
    Fb = np.diag([1 + e, 1.0, 1.0])
    cell = M.Cell(rho)
    mix = M.Mixed(cell, free=free)
    w, F, log = mix.solve(Fb, rtol=1e-12, cg_tol=1e-10)
    assert log["converged"]
    g, info = mix.grad_rho(log["z"], Fb, lambda z, F_, r: -(z[mix.nw] - 1.0) / e)
    assert info["converged"]
    d = rng.standard_normal(rho.shape) * (rho * (1 - rho))
    fd = _fd_solves(M, rho, free, Fb, d, 1e-5, lambda c, w_, F_: -(F_[1, 1] - 1.0) / e)
    assert float(torch.sum(g * torch.as_tensor(d, device="cuda"))) == pytest.approx(fd, rel=1e-4, abs=1e-9)


def test_grad_rho_of_mean_stress_matches_central_differences(M):
    """q = Pbar_xx at the uniaxial-stress equilibrium (the StressPath loss of nlfiber/steer/losses.py:86-101):
    both the explicit density term and the adjoint coupling are exercised."""
    import torch
    rng = np.random.default_rng(10)
    rho = rng.uniform(0.3, 1.0, (4, 4, 4))
    free = [(1, 1), (2, 2)]
    Fb = np.diag([1.03, 1.0, 1.0])
    cell = M.Cell(rho)
    mix = M.Mixed(cell, free=free)
    w, F, log = mix.solve(Fb, rtol=1e-12, cg_tol=1e-10)

    def q(z, F_, r):
        return cell.mean_stress_t(*mix.unpack_t(z, F_), r)[0, 0]

    g, info = mix.grad_rho(log["z"], Fb, q)
    d = rng.standard_normal(rho.shape) * (rho * (1 - rho))
    fd = _fd_solves(M, rho, free, Fb, d, 1e-5, lambda c, w_, F_: float(c.mean_stress(w_, F_)[0, 0]))
    assert float(torch.sum(g * torch.as_tensor(d, device="cuda"))) == pytest.approx(fd, rel=1e-5)


def test_effective_tangent_at_identity_is_linear_C_and_elliptic(M, gpu_pkg):
    """The reference's test_effective_tangent_at_identity (tests/test_mech.py:43-47) on an 8^3 cell."""
    import torch
    rng = np.random.default_rng(14)
    rho = rng.uniform(0.2, 1.0, (8, 8, 8))
    lh = gpu_pkg.LinearHomogenization((8, 8, 8), tol=1e-11)
    C, _, _ = lh.C_full(torch.as_tensor(rho, device="cuda"))
    C = C.cpu().numpy()
    cell = M.Cell(rho)
    A = M.effective_tangent(cell, torch.zeros(cell.shape + (3,), dtype=torch.float64, device="cuda"), np.eye(3))
    V = np.zeros((6, 6))
    for I, (i, j) in enumerate(LO.VOIGT):
        for J, (k, l) in enumerate(LO.VOIGT):
            V[I, J] = A[i, j, k, l]
    np.testing.assert_allclose(V, C, atol=1e-7 * np.abs(C).max())
    val, n = M.rank_one_min(A, m=200)
    assert val > 0 and abs(np.linalg.norm(n) - 1.0) < 1e-12


def test_effective_tangent_matches_central_differences_of_solves(M):
    import torch
    rng = np.random.default_rng(15)
    rho = rng.uniform(0.3, 1.0, (4, 4, 4))
    cell = M.Cell(rho)
    full = M.Mixed(cell, free=[])
    Fb = np.diag([1.04, 0.99, 1.0]) + 0.01 * np.array([[0, 1, 0], [0.5, 0, 0], [0, 0, 0]])
    w, F, log = full.solve(Fb, rtol=1e-12, cg_tol=1e-10)
    A = M.effective_tangent(cell, w, F)
    h = 1e-5
    for c, d in [(0, 0), (1, 2), (2, 1)]:
        E = np.zeros((3, 3))
        E[c, d] = h
        Pp = cell.mean_stress(*full.solve(F + E, rtol=1e-12, cg_tol=1e-10)[:2])
        Pm = cell.mean_stress(*full.solve(F - E, rtol=1e-12, cg_tol=1e-10)[:2])
        np.testing.assert_allclose(A[:, :, c, d], (Pp - Pm) / (2 * h), rtol=1e-5, atol=1e-8 * np.abs(A).max())


def test_lobpcg_matches_dense_spectrum(M):
    """The reference's test_lobpcg_matches_dense_spectrum (tests/test_mech.py:115-135): the dense Hessian
    here comes from the oracle's stored-tangent products at the GPU solve's state."""
    rng = np.random.default_rng(16)
    rho = rng.uniform(0.2, 1.0, (4, 4, 4))
    cell = M.Cell(rho)
    free = [(1, 1), (2, 2)]
    mix = M.Mixed(cell, free)
    w, F, log = mix.solve(np.diag([0.95, 1.0, 1.0]), rtol=1e-12)
    z = log["z"]
    vals, _, _ = mix.lowest_eigs(z, F, k=3, tol=1e-8)
    ref = MR.CellRef(rho)
    wz = z[: mix.nw].cpu().numpy().reshape(rho.shape + (3,))
    At = ref.tangent(wz, F)
    nz = mix.nw + 2
    H = np.empty((nz, nz))
    for c in range(nz):
        v = np.zeros(nz)
        v[c] = 1.0
        dF = np.zeros((3, 3))
        for j, (a, b) in enumerate(free):
            dF[a, b] = v[mix.nw + j]
        hw, hP = ref.hvp(At, v[: mix.nw].reshape(rho.shape + (3,)), dF)
        H[:, c] = np.concatenate([hw.ravel(), [hP[a, b] for a, b in free]])
    dsc = mix._diag_scale()
    Hs = H * dsc[:, None] * dsc[None, :]
    Tm = np.zeros((nz, 3))
    for i in range(3):
        Tm[i:mix.nw:3, i] = 1.0
    Y, _ = np.linalg.qr(Tm / dsc[:, None])
    Pm = np.eye(nz) - Y @ Y.T
    ev = np.linalg.eigvalsh(Pm @ Hs @ Pm)
    ev = ev[np.abs(ev) > 1e-12 * np.abs(ev).max()]
    np.testing.assert_allclose(vals, np.sort(ev)[:3], rtol=1e-5)


@pytest.mark.parametrize("material,wls", [("neohooke", (500.0, 0.01)), ("neohooke", None), ("linear", None)])
def test_toec_derivatives_match_oracle(M, material, wls):
    import torch
    rng = np.random.default_rng(18)
    rho = rng.uniform(0.1, 1.0, (4, 3, 5))
    chi = 0.3 * rng.standard_normal((6,) + rho.shape + (3,))
    cell = M.Cell(rho, material=material, wls=wls)
    C, T = M.toec(cell, torch.as_tensor(chi, device="cuda"))
    C2, T3 = MR.toec_derivatives(MR.CellRef(rho, material=material, wls=wls), chi)
    assert _rel(C, C2) < 1e-12
    X = np.zeros((6, 6, 6))
    for J in range(6):
        for K in range(6):
            X[:, J, K] = M.to_voigt(M.voigt_tensor(J) @ M.voigt_tensor(K))
    CX = np.einsum("ia,ajk->ijk", C2, X)
    sym = sum(np.transpose(CX, p) for p in [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]) / 6
    assert np.abs(T - (T3 - 3 * sym)).max() < 1e-6 * np.abs(T3 - 3 * sym).max()


def test_toec_against_linear_C_and_nonlinear_solves(M, gpu_pkg):
    """The reference's test_toec_second_order_is_linear_C and test_toec_against_nonlinear_solves
    (tests/test_mech.py:36-68) on a 4^3 cell: C from toec equals the linear C; g(t) = Wbar(I + t n) has
    g''' from central differences of Pbar along the path equal to T(n, n, n) + 3 n.C.voigt(Hn Hn)."""
    import torch
    rng = np.random.default_rng(19)
    rho = rng.uniform(0.3, 1.0, (4, 4, 4))
    lh = gpu_pkg.LinearHomogenization((4, 4, 4), tol=1e-12)
    Cl, chi, _ = lh.C_full(torch.as_tensor(rho, device="cuda"))
    Cl = Cl.cpu().numpy()
    cell = M.Cell(rho)
    C2, T = M.toec(cell, chi.contiguous())
    np.testing.assert_allclose(C2, Cl, atol=1e-9 * np.abs(Cl).max())
    np.testing.assert_allclose(T, np.transpose(T, (1, 0, 2)), atol=1e-12 * np.abs(T).max())
    np.testing.assert_allclose(T, np.transpose(T, (0, 2, 1)), atol=1e-12 * np.abs(T).max())
    n = np.array([1.0, -0.3, 0.2, 0.4, -0.2, 0.1])
    Hn = sum(n[I] * M.voigt_tensor(I) for I in range(6))
    X = M.to_voigt(Hn @ Hn)
    pred = np.einsum("ijk,i,j,k", T, n, n, n) + 3 * n @ C2 @ X
    full = M.Mixed(cell, free=[])
    h = 0.004

    def dg(t):
        w, F, log = full.solve(np.eye(3) + t * Hn, rtol=1e-12, cg_tol=1e-10)
        assert log["converged"]
        return float(np.sum(cell.mean_stress(w, F) * Hn))

    g3 = (dg(h) - 2 * dg(0.0) + dg(-h)) / h**2
    assert g3 == pytest.approx(pred, rel=2e-3)
    c = M.cubic_toec(T)
    assert set(c) == {"C111", "C112", "C123", "C144", "C155", "C456"}


def test_material_functions_match_the_kernel_energy(M):
    """psi_linear / psi_neohooke (torch) against the kernels' energy of a uniform state (w = 0)."""
    import torch
    rng = np.random.default_rng(20)
    rho = np.ones((3, 3, 3))
    Fb = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
    for material in ("linear", "neohooke"):
        cell = M.Cell(rho, material=material, wls=None, emin=0.0)
        W = cell.energy(torch.zeros(cell.shape + (3,), dtype=torch.float64, device="cuda"), Fb)
        psi = float(cell.psi(torch.as_tensor(Fb), cell.lam, cell.mu))
        assert W == pytest.approx(psi, rel=1e-13)


def test_plane_strain_cell_newton_matches_oracle(M):
    """A 2D density (n, n) is one periodic element layer (exact plane strain, mech.py:93-96)."""
    rng = np.random.default_rng(21)
    rho = rng.uniform(0.3, 1.0, (5, 4))
    cell = M.Cell(rho)
    assert cell.shape == (5, 4, 1)
    free = [(1, 1)]
    Fb = np.diag([1.03, 1.0, 1.0])
    w, F, log = M.Mixed(cell, free=free).solve(Fb, rtol=1e-11, cg_tol=1e-10)
    assert log["converged"]
    ref = MR.CellRef(rho)
    wr, Fr = MR.newton_dense(ref, Fb, free=free, rtol=1e-12)
    np.testing.assert_allclose(F, Fr, atol=1e-10)
    assert _rel(cell.mean_stress(w, F), ref.stress(wr, Fr)[1]) < 1e-8


def test_energy_at_another_density(M):
    """Cell.energy(w, Fbar, rho) (mech.py:141-145) equals the energy of a cell built on that density."""
    import torch
    rng = np.random.default_rng(22)
    rho = rng.uniform(0.0, 1.0, (4, 3, 3))
    rho2 = rng.uniform(0.0, 1.0, (4, 3, 3))
    w = torch.as_tensor(0.05 * rng.standard_normal(rho.shape + (3,)), device="cuda")
    Fb = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    a, b = M.Cell(rho), M.Cell(rho2)
    assert a.energy(w, Fb, rho2) == b.energy(w, Fb)
    assert a.energy(w, Fb) == pytest.approx(MR.CellRef(rho).energy(w.cpu().numpy(), Fb), rel=1e-12)
    np.testing.assert_array_equal(a.gradient(w, Fb, rho2)[1], b.gradient(w, Fb)[1])
    mix = M.Mixed(a, free=[])
    assert mix.Ez(w.reshape(-1), Fb, rho2) == b.energy(w, Fb)


@pytest.mark.parametrize("material,wls", [("neohooke", (500.0, 0.01)), ("neohooke", None), ("linear", None)])
def test_toec_closed_form_matches_kernel(M, material, wls):
    """toec's backward differentiates toec_values_t: its 77 values equal the kernel's to round-off."""
    import torch
    rng = np.random.default_rng(21)
    rho = rng.uniform(0.1, 1.0, (4, 5, 3))
    chi = torch.as_tensor(0.3 * rng.standard_normal((6,) + rho.shape + (3,)), device="cuda")
    cell = M.Cell(rho, material=material, wls=wls)
    v = M._toec_kernel(cell, chi.reshape(6, -1).contiguous())
    sw = (cell.wq * cell.scale_of(cell.rho)).reshape(-1)
    gam = cell.gamma_of(cell.rho)
    gam = torch.ones_like(sw) if gam is None else gam.reshape(-1)
    vt = M.toec_values_t(cell, chi.reshape(6, -1), sw, gam)
    assert _rel(vt.cpu(), v.cpu()) < 1e-12


def test_toec_loss_gradient_matches_central_differences(M, gpu_pkg):
    """The reference's test_toec_loss_gradient (tests/test_pipeline.py:112-119) with its ToecTarget loss
    (nlfiber/steer/losses.py:103-124): L(rho) = (C111 + 1)^2 + C456^2 from toec of the cell with the linear
    correctors chi(rho); the autograd gradient (through toec, the corrector adjoint and the stiffness scale)
    against central differences along d = randn * rho (1 - rho)."""
    import torch
    from nlfh_amd.density import DensityMap
    dm = DensityMap(4, r=1.5, v=0.3)
    phi = np.random.default_rng(30).standard_normal((4, 4, 4))
    rho0 = dm.rho(phi, 4.0).cpu().numpy()
    lh = gpu_pkg.LinearHomogenization(rho0.shape, E=1.0, nu=0.3, emin=1e-3, p=3.0, tol=1e-12)
    zeros = torch.zeros((6,) + rho0.shape + (3,), dtype=torch.float64, device="cuda")

    def loss(rho):
        s = lh.stiffness_scale(rho)
        chi = lh.correctors(s, np.eye(6), zeros)
        cell = M.Cell(rho, E=1.0, nu=0.3, wls=(500.0, 0.01), emin=1e-3, p=3.0)
        _, T = M.toec(cell, chi)
        vals = M.cubic_toec_jnp(T)
        return (vals["C111"] + 1.0) ** 2 + vals["C456"] ** 2

    r = torch.tensor(rho0, device="cuda", requires_grad=True)
    L = loss(r)
    (g,) = torch.autograd.grad(L, r)
    d = np.random.default_rng(1).standard_normal(rho0.shape) * rho0 * (1 - rho0)
    h = 1e-5
    with torch.no_grad():
        fd = (float(loss(torch.tensor(rho0 + h * d, device="cuda"))) -
              float(loss(torch.tensor(rho0 - h * d, device="cuda")))) / (2 * h)
    gd = float((g.cpu().numpy() * d).sum())
    assert abs(gd - fd) < 1e-3 * abs(fd) + 1e-8, (gd, fd)
    # without grad the same call returns numpy arrays equal to the tensors' values
    with torch.no_grad():
        cell = M.Cell(rho0, E=1.0, nu=0.3, wls=(500.0, 0.01), emin=1e-3, p=3.0)
        chi = lh.correctors(lh.stiffness_scale(torch.tensor(rho0, device="cuda")), np.eye(6), zeros)
        Cn, Tn = M.toec(cell, chi)
    assert isinstance(Tn, np.ndarray)
    Ct, Tt = M.toec(M.Cell(r, E=1.0, nu=0.3, wls=(500.0, 0.01), emin=1e-3, p=3.0), chi)
    np.testing.assert_allclose(Tt.detach().cpu().numpy(), Tn, rtol=0, atol=1e-13 * np.abs(Tn).max())
    np.testing.assert_allclose(Ct.detach().cpu().numpy(), Cn, rtol=0, atol=1e-13 * np.abs(Cn).max())

"""Batched plans: many cells of uneven sizes in one solve give the same C and iteration counts as
one solve per cell (the reference's per-cell C_full, nlfiber/linear.py:231-236)."""
import numpy as np
import pytest
import torch

from oracle import linear_ref as O

pytestmark = pytest.mark.gpu


def _cells(pkg):
    out = []
    for n, seed in ((8, 1), (12, 2), (8, 3), (16, 4)):
        rng = np.random.default_rng(seed)
        out.append(pkg.cells.exact_symmetrize(pkg.cells.density(rng.standard_normal((n, n, n)), r=1.0, beta=4.0,
                                                                v=0.3, filter_kernel="hat")))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_equals_single_cell_solves(gpu_pkg, precision):
    rhos = _cells(gpu_pkg)
    bh = gpu_pkg.BatchedHomogenization([r.shape for r in rhos], emin=1e-3, tol=1e-10, precision=precision)
    C, u, its = bh.C_full(rhos)
    assert all(bh.last_info["converged"])
    for i, r in enumerate(rhos):
        lh = gpu_pkg.LinearHomogenization(r.shape, emin=1e-3, tol=1e-10, precision=precision)
        Ci, ui, iti = lh.C_full(r)
        assert np.abs(np.asarray(C[i]) - np.asarray(Ci)).max() < 1e-12 * np.abs(np.asarray(Ci)).max()
        if precision == "fp64":
            assert its[i] == iti
        views = bh.field_views(u)
        # mixed: a plan whose cells all take the register-FFT path runs the lean FP32 schedule (spectral r.z,
        # x updated by the matvec), so the batch (with a 12^3 cell) and the lone 16^3 plan round
        # differently in FP32; both converge to tol 1e-10 and C agrees to 1e-12 above
        utol = 1e-9 if precision == "fp64" else 1e-8
        assert np.abs(views[i].cpu().numpy() - ui.cpu().numpy()).max() < utol * np.abs(ui.cpu().numpy()).max()
    # and against the oracle for the largest cell
    Cr, _, _ = O.LinearRef(rhos[3].shape, emin=1e-3, tol=1e-10).C_full(rhos[3])
    assert np.abs(np.asarray(C[3]) - Cr).max() < 1e-9 * np.abs(Cr).max()


def test_mixed_precision_true_residual_and_C(gpu_pkg):
    rho = gpu_pkg.cells.random_cell(16, 7)[0]
    lh = gpu_pkg.LinearHomogenization(rho.shape, emin=1e-4, tol=1e-8, precision="mixed")
    C, u, it = lh.C_full(rho)
    assert lh.last_info["converged"] and max(lh.last_info["relres"]) <= 1e-8
    ref = O.LinearRef(rho.shape, emin=1e-4, tol=1e-8)
    s = ref.stiffness_scale(rho)
    r = ref.residual(s, np.eye(6), u.cpu().numpy())
    b = ref.load(s, np.eye(6))
    true = np.sqrt((r ** 2).sum(axis=(1, 2, 3, 4))) / np.sqrt((b ** 2).sum(axis=(1, 2, 3, 4)))
    assert true.max() <= 1.5e-8
    Cr, _, _ = ref.C_full(rho)
    assert np.abs(np.asarray(C) - Cr).max() <= 1e-6 * np.abs(Cr).max()


def test_carried_directions_match_fp64_and_save_iterations(gpu_pkg):
    """Mixed precision with the search direction carried over the refinement steps (the default) against
    inner solves restarted from p = z and against plain FP64 PCG: every cell converges to the true FP64
    residual without a fallback, C agrees with the FP64 C to 1e-9 of its largest entry, and no cell needs
    more inner iterations than with restarts (measured: 637/701/654/279 carried, 680/727/686/297 restarted,
    585/691/634/272 FP64; the 32^3 cells need at most 5 % more than FP64 PCG)."""
    cells = [gpu_pkg.cells.random_cell(n, seed) for n, seed in ((16, 21), (32, 22), (32, 23), (16, 24))]
    rhos = [c[0] for c in cells]
    shapes = [r.shape for r in rhos]
    res = {}
    for name, kw in (("carry", dict(precision="mixed")), ("restart", dict(precision="mixed", inner_restart=True)),
                     ("fp64", dict(precision="fp64"))):
        bh = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, tol=1e-8, streams=1, **kw)
        C, _, its = bh.C_full(rhos)
        assert all(bh.last_info["converged"]) and bh.last_info["fallback_fp64"] == [], name
        assert np.asarray(bh.last_info["relres"]).max() <= 1e-8, name
        res[name] = (np.asarray(C), np.asarray(its))
        print(name, list(its), flush=True)
    C64, it64 = res["fp64"]
    for name in ("carry", "restart"):
        C, _ = res[name]
        for i in range(len(rhos)):
            assert np.abs(C[i] - C64[i]).max() <= 1e-9 * np.abs(C64[i]).max(), (name, i)
    it_c, it_r = res["carry"][1], res["restart"][1]
    assert np.all(it_c <= it_r), (it_c, it_r)
    assert it_c.sum() <= 0.97 * it_r.sum(), (it_c, it_r)
    for i, r in enumerate(rhos):
        if r.shape[0] == 32:
            assert it_c[i] <= 1.05 * it64[i], (i, it_c[i], it64[i])


def test_mixed_unconverged_cells_fall_back_to_fp64(gpu_pkg):
    """Cells the mixed-precision solve leaves unconverged are solved again by plain FP64 PCG from zero
    (iteration cap forced from the host): the fallback cell's correctors are exactly those of an FP64
    plan of its own with the same cap, the converged cell keeps its mixed-precision correctors."""
    rho_small = gpu_pkg.cells.random_cell(4, 1)[0]
    rho_big = gpu_pkg.cells.random_cell(16, 2)[0]
    maxit = 80  # the 4^3 cell needs 15 FP64 iterations, the 16^3 cell 203 (oracle counts)
    bh = gpu_pkg.BatchedHomogenization([(4, 4, 4), (16, 16, 16)], emin=1e-3, tol=1e-8, maxit=maxit, precision="mixed")
    C, u, it = bh.C_full([rho_small, rho_big])
    info = bh.last_info
    assert info["fallback_fp64"] == [1]
    assert info["converged"][0] and not info["converged"][1]
    one64 = gpu_pkg.BatchedHomogenization([(16, 16, 16)], emin=1e-3, tol=1e-8, maxit=maxit, precision="fp64")
    _, u64, _ = one64.C_full([rho_big])
    f = 3 * 6
    assert torch.equal(u[int(bh.off[1]) * f:int(bh.off[2]) * f], u64)
    alone = gpu_pkg.BatchedHomogenization([(4, 4, 4)], emin=1e-3, tol=1e-8, maxit=maxit, precision="mixed")
    _, u8, _ = alone.C_full([rho_small])
    assert torch.equal(u[:int(bh.off[1]) * f], u8)
    # the single-cell class reports the same fallback
    lh = gpu_pkg.LinearHomogenization((16, 16, 16), emin=1e-3, tol=1e-8, maxit=maxit, precision="mixed")
    lh.C_full(torch.tensor(rho_big))
    assert lh.last_info["fallback_fp64"] and not lh.last_info["converged"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stream_groups_do_not_change_results(gpu_pkg, precision):
    """The solve split over 2 or 3 groups (a plan, a stream and a host thread each) gives every cell the
    iterates of the single-plan solve: C, correctors and iteration counts agree exactly (every cell is of
    a register-FFT size, so all groups run the same FP32 schedule)."""
    rhos = [gpu_pkg.cells.random_cell(n, seed)[0] for n, seed in ((16, 11), (32, 12), (16, 13), (16, 14), (32, 15))]
    res = {}
    for streams in (1, 2, 3):
        bh = gpu_pkg.BatchedHomogenization([r.shape for r in rhos], emin=1e-4, tol=1e-8, precision=precision,
                                           streams=streams)
        C, u, its = bh.C_full(rhos)
        assert all(bh.last_info["converged"])
        res[streams] = (np.asarray(C), u.cpu().numpy(), np.asarray(its), bh._groups)
    assert res[1][3] is None and len(res[2][3]) == 2 and len(res[3][3]) == 3
    for streams in (2, 3):
        assert sorted(i for g in res[streams][3] for i in g) == list(range(len(rhos)))
        np.testing.assert_array_equal(res[streams][2], res[1][2])
        np.testing.assert_array_equal(res[streams][0], res[1][0])
        np.testing.assert_array_equal(res[streams][1], res[1][1])


@pytest.mark.parametrize("cells", [((16, 21), (16, 22), (32, 23), (32, 24), (16, 25)), ((128, 41), (32, 42))])
def test_fp32_direction_iteration_matches_plain_fp32_schedule(gpu_pkg, cells):
    """The FP32 inner iteration of plans whose cells all take the staged register-FFT passes (search
    direction formed in pass C, residual kept as its (z, y) spectrum, iterate stepped every second
    iteration with the owed step of an even last iteration flushed) runs the same PCG recurrences as the
    plain FP32 schedule, which a plan takes as soon as one cell (here 12^3) has no register FFT (128^3 cells
    run the staged passes in place with 512-thread blocks). Per cell,
    the mixed solve needs at most 3 % more inner iterations than in the plain schedule (an inner solve that
    lost a step would need extra refinement steps; measured on MI355X: equal or up to 4 % fewer) and C
    agrees to 1e-12 of max|C|."""
    rhos = [gpu_pkg.cells.random_cell(n, seed)[0] for n, seed in cells]
    extra = gpu_pkg.cells.random_cell(12, 99)[0]
    shapes = [r.shape for r in rhos]
    lean = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, tol=1e-8, precision="mixed", streams=1)
    plain = gpu_pkg.BatchedHomogenization(shapes + [extra.shape], emin=1e-4, tol=1e-8, precision="mixed", streams=1)
    Cl, _, itl = lean.C_full(rhos)
    Cp, _, itp = plain.C_full(rhos + [extra])
    assert all(lean.last_info["converged"]) and lean.last_info["fallback_fp64"] == []
    for i in range(len(rhos)):
        assert itl[i] <= itp[i] + max(8, 0.03 * itp[i]), (i, itl[i], itp[i])
        Ci, Pi = np.asarray(Cl[i]), np.asarray(Cp[i])
        assert np.abs(Ci - Pi).max() < 1e-12 * np.abs(Pi).max()


def test_misaligned_128_cell_runs_the_unstaged_passes(gpu_pkg):
    """A 128^3 cell placed after a 3^3 cell sits at a field offset that is not 16-byte aligned, so its FP32
    passes run the unstaged register-FFT kernels (1024-thread blocks, like the staged ones) in the plain FP32
    schedule (the 3^3 cell has no register FFT). Its C and iteration count match the same cell solved alone,
    aligned, in the staged direction schedule (C to 1e-10 of max|C|, iterations within 3 %)."""
    small = gpu_pkg.cells.random_cell(3, 51)[0]
    big = gpu_pkg.cells.random_cell(128, 52)[0]
    both = gpu_pkg.BatchedHomogenization([small.shape, big.shape], emin=1e-3, tol=1e-8, precision="mixed", streams=1)
    alone = gpu_pkg.BatchedHomogenization([big.shape], emin=1e-3, tol=1e-8, precision="mixed", streams=1)
    assert int(both.off[1]) * 3 * 6 % 4 != 0  # the 128^3 cell's fields start off a 16-byte boundary
    Cb, _, itb = both.C_full([small, big])
    Ca, _, ita = alone.C_full([big])
    assert all(both.last_info["converged"]) and all(alone.last_info["converged"])
    assert abs(int(itb[1]) - int(ita[0])) <= max(8, 0.03 * int(ita[0]))
    c_b, c_a = np.asarray(Cb[1]), np.asarray(Ca[0])
    assert np.abs(c_b - c_a).max() < 1e-10 * np.abs(c_a).max()


def test_mixed_precision_with_two_and_three_right_hand_sides(gpu_pkg):
    """The FP32 direction iteration with fewer right-hand sides than the plan's layout (per-(cell, rhs)
    scalars and partials at stride m): the cubic invariants' two cases (m = 2, as Fiber solves them) and a
    three-strain solve (m = 3) on a 32^3 cell agree with the FP64 solve and converge in at most 25 % more
    inner iterations. Both stop at residual 1e-9 (true and recursive residual), so the correctors may differ
    by the conditioning times that (bound: 1e-4 of their largest entry; a stride mistake gives O(1)); the
    invariants' error is quadratic in the correctors' (bound: 1e-8)."""
    rho = gpu_pkg.cells.random_cell(32, 61)[0]
    mixed = gpu_pkg.LinearHomogenization(rho.shape, emin=1e-4, tol=1e-9, precision="mixed")
    fp64 = gpu_pkg.LinearHomogenization(rho.shape, emin=1e-4, tol=1e-9)
    cm, um, itm = mixed.invariants(rho)
    assert mixed.last_info["converged"] and not mixed.last_info["fallback_fp64"]
    cd, ud, itd = fp64.invariants(rho)
    np.testing.assert_allclose(cm.detach().cpu().numpy(), cd.detach().cpu().numpy(), rtol=1e-8)
    assert (um - ud).abs().max() < 1e-4 * ud.abs().max()
    assert itm <= 1.25 * itd + 10
    Es = np.stack([gpu_pkg.unit_strain(0), gpu_pkg.unit_strain(2), gpu_pkg.unit_strain(4)])
    s = fp64.stiffness_scale(rho)
    u3m, it3m = mixed.solve(s, Es)
    assert mixed.last_info["converged"]
    u3d, it3d = fp64.solve(s, Es)
    assert (u3m - u3d).abs().max() < 1e-4 * u3d.abs().max()
    assert it3m <= 1.25 * it3d + 10


def test_preconditioner_block_shapes_keep_the_solve(gpu_pkg):
    """A mixed-precision batch solved with every FP32 pass B block shape and one or two pass A planes per block
    gives the same C to 1e-12 and the same iteration counts to 2 (the pass B shape regroups the fixed-order r . z
    partial sums, so the scalars round differently); the pass A planes alone change nothing (bitwise)."""
    rhos = [gpu_pkg.cells.random_cell(n, seed)[0] for n, seed in ((32, 31), (32, 32), (16, 33), (64, 34))]
    shapes = [r.shape for r in rhos]
    res = []
    for which, planes in ((0, 1), (0, 2), (1, 1), (2, 2)):
        bh = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, tol=1e-8, streams=1)
        for i in range(3):
            _, d, _ = bh._plan.get_fft_tiling(i)
            bh._plan.set_fft_tiling(i, [d, d // 2, 2 * d][which], planes)
        C, u, its = bh.C_full(rhos)
        assert all(bh.last_info["converged"])
        res.append((np.asarray(C), np.asarray(its), u))
    assert np.array_equal(res[1][0], res[0][0]) and np.array_equal(res[1][1], res[0][1])
    assert torch.equal(res[1][2], res[0][2])
    for C, its, _ in res[2:]:
        assert np.abs(C - res[0][0]).max() <= 1e-12 * np.abs(res[0][0]).max()
        assert np.abs(its - res[0][1]).max() <= 2

"""The benchmark configurations' cells on the GPU, against FP64.

Config 2: a batch that mixes 32^3, 48^3 and 64^3 seeded random cubic-symmetric cells, solved by the
batched mixed-precision solver in two stream groups with autotuned tiling, against the batched plain FP64
PCG of the same cells and, for a 32^3 member, against the oracle. Config 4: a 128^3 high-contrast cell
(emin 1e-6, volume 0.15, beta 16) by the mixed solver against the FP64 solve.

The gate is the reference's normalisation, |dC_ij| <= 1e-6 max|C| (tests/test_core.py:124-133).
"""
import numpy as np
import pytest

from oracle import linear_ref as O

pytestmark = pytest.mark.gpu


def _err(C, C64):
    return float(np.abs(np.asarray(C) - np.asarray(C64)).max() / np.abs(np.asarray(C64)).max())


@pytest.mark.timeout(600)
def test_config2_uneven_batch_mixed_matches_fp64_and_oracle(gpu_pkg):
    cells = [(32, s) for s in range(4)] + [(48, s) for s in range(128, 131)] + [(64, 224)]
    rhos = [gpu_pkg.cells.random_cell(n, seed)[0] for n, seed in cells]
    shapes = [r.shape for r in rhos]
    mixed = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, tol=1e-8, maxit=40000, precision="mixed", streams=2)
    mixed.autotune()
    assert mixed._groups is not None and len(mixed._groups) == 2
    C, u, its = mixed.C_full(rhos)
    info = mixed.last_info
    assert all(info["converged"]) and info["fallback_fp64"] == []
    assert np.asarray(info["relres"]).max() <= 1e-8  # true FP64 residual per (cell, rhs)
    fp64 = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, tol=1e-8, maxit=40000, precision="fp64", streams=2)
    C64, _, its64 = fp64.C_full(rhos)
    assert all(fp64.last_info["converged"])
    for i, (n, seed) in enumerate(cells):
        assert _err(C[i], C64[i]) <= 1e-6, (n, seed, _err(C[i], C64[i]))
        # the refinement restarts cost a few percent of inner iterations, never a multiple
        assert its[i] <= 1.5 * its64[i] + 20, (n, seed, its[i], its64[i])
    ref = O.LinearRef(shapes[0], emin=1e-4, tol=1e-8)
    Cr, _, itr = ref.C_full(rhos[0])
    assert _err(C[0], Cr) <= 1e-6
    assert _err(C64[0], Cr) <= 1e-6
    assert abs(int(its64[0]) - int(itr)) <= 1  # the reference's PCG and stopping rule


@pytest.mark.timeout(900)
def test_config4_stress_cell_mixed_matches_fp64(gpu_pkg):
    rho, meta = gpu_pkg.cells.stress_cell(0, 128)
    assert abs(float(rho.mean()) - 0.15) < 1e-10
    mixed = gpu_pkg.BatchedHomogenization([rho.shape], emin=1e-6, tol=1e-8, maxit=40000, precision="mixed")
    mixed.autotune()
    C, _, its = mixed.C_full([rho])
    info = mixed.last_info
    # converged by the mixed solver itself: no FP64 fallback was needed
    assert info["converged"] == [True] and info["fallback_fp64"] == []
    assert np.asarray(info["relres"]).max() <= 1e-8
    fp64 = gpu_pkg.BatchedHomogenization([rho.shape], emin=1e-6, tol=1e-8, maxit=40000, precision="fp64")
    C64, _, its64 = fp64.C_full([rho])
    assert fp64.last_info["converged"] == [True]
    assert _err(C[0], C64[0]) <= 1e-6, _err(C[0], C64[0])
    # cubic symmetry survives the solve: C11 = C22 = C33, C44 = C55 = C66
    c = np.asarray(C[0])
    d = np.diag(c)
    assert np.ptp(d[:3]) <= 1e-6 * d[:3].max() and np.ptp(d[3:]) <= 1e-6 * d[3:].max()

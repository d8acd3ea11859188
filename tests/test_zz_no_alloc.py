"""The no-allocation finite-strain evaluation (nlfh_fs_eval_noalloc, its workspace a static device array of the
library) against nlfh_fs_eval with a caller-allocated workspace: energy and gradient (mode 1) and the
Hessian-vector product (mode 2) agree bitwise on two calls in a row, the second on a smaller cell (the array
keeps the first call's contents), and a cell larger than the array is refused with an error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _both(pkg, cell, w, F, v, dF, mode):
    """(grad, scalars) from nlfh_fs_eval (own workspace) and from nlfh_fs_eval_noalloc."""
    lib = pkg.load_library()
    Fh = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(9))
    dFd = torch.as_tensor(np.asarray(dF, dtype=np.float64).reshape(9), device="cuda")
    out = []
    for noalloc in (False, True):
        grad = torch.full_like(w, np.nan)
        scal = torch.full((10,), np.nan, dtype=torch.float64, device="cuda")
        common = (cell._shape_c, _ptr(cell._sg), cell._mat.ctypes.data_as(_dp), cell._flags, _ptr(w),
                  Fh.ctypes.data_as(_dp), _ptr(v), _ptr(dFd), mode, _ptr(grad), _ptr(scal))
        if noalloc:
            rc = lib.nlfh_fs_eval_noalloc(*common, _stream())
        else:
            work = torch.empty(int(lib.nlfh_fs_work_doubles(*cell.shape)), dtype=torch.float64, device="cuda")
            rc = lib.nlfh_fs_eval(*common, _ptr(work), _stream())
        pkg._lib.check(rc)
        torch.cuda.synchronize()
        out.append((grad.cpu(), scal.cpu()))
    return out


@pytest.mark.parametrize("material", ["neohooke", "linear"])
def test_noalloc_eval_matches_eval_on_two_calls(gpu_pkg, material):
    from nlfh_amd import mech
    lib = gpu_pkg.load_library()
    nmax = int(lib.nlfh_fs_noalloc_max_voxels())
    assert nmax == 32 ** 3
    rng = np.random.default_rng(3)
    for n in (32, 12):  # the largest cell the array holds, then a smaller one
        rho, _ = gpu_pkg.cells.random_cell(n, 100 + n)
        cell = mech.Cell(rho, material=material)
        w = torch.as_tensor(0.02 * rng.standard_normal(cell.N * 3), device="cuda")
        v = torch.as_tensor(rng.standard_normal(cell.N * 3), device="cuda")
        F = np.eye(3) + 0.03 * rng.standard_normal((3, 3))
        dF = rng.standard_normal((3, 3))
        for mode in (1, 2):
            (g0, s0), (g1, s1) = _both(gpu_pkg, cell, w, F, v, dF, mode)
            assert torch.isfinite(g0).all() and torch.isfinite(s0).all()
            assert torch.equal(g0, g1), (n, mode)
            assert torch.equal(s0, s1), (n, mode)


def test_noalloc_eval_refuses_cells_beyond_its_array(gpu_pkg):
    from nlfh_amd import mech
    lib = gpu_pkg.load_library()
    rho, _ = gpu_pkg.cells.random_cell(40, 7)  # 64000 voxels > 32768
    cell = mech.Cell(rho)
    w = torch.zeros(cell.N * 3, dtype=torch.float64, device="cuda")
    grad = torch.empty_like(w)
    scal = torch.empty(10, dtype=torch.float64, device="cuda")
    F = np.ascontiguousarray(np.eye(3).reshape(9))
    rc = lib.nlfh_fs_eval_noalloc(cell._shape_c, _ptr(cell._sg), cell._mat.ctypes.data_as(_dp), cell._flags, _ptr(w),
                                  F.ctypes.data_as(_dp), None, None, 1, _ptr(grad), _ptr(scal), _stream())
    assert rc != 0
    assert b"static workspace" in lib.nlfh_last_error()

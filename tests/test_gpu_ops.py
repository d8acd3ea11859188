"""Operator parity on the GPU: every C-ABI operator against the FP64 oracle on the same inputs.

Tolerances: float64 results to 1e-12 relative to the largest entry (round-off of different
summation orders); float32 operators to 2e-6 (matvec) and 1e-5 (preconditioner), relative to the
largest entry.
"""
import numpy as np
import pytest
import torch

from oracle import linear_ref as O

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (3, 4, 5), (6, 5, 4), (8, 8, 1), (12, 10, 6), (16, 16, 16), (32, 24, 48), (20, 18, 64),
          (8, 6, 128), (5, 9, 3), (64, 64, 64)]


def _rand(shape, seed, m=2):
    rng = np.random.default_rng(seed)
    s = rng.uniform(1e-3, 1.0, shape)
    u = rng.standard_normal((m,) + shape + (3,))
    return s, u


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("shape", SHAPES)
def test_apply_K_fp64(gpu_pkg, shape):
    s, u = _rand(shape, 1)
    lh = gpu_pkg.LinearHomogenization(shape)
    got = lh.apply_K(s, u).cpu().numpy()
    ref = np.stack([O.LinearRef(shape).apply_K(s, x) for x in u])
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 24, 48), (8, 6, 128)])
def test_apply_K_fp32(gpu_pkg, shape):
    import ctypes
    s, u = _rand(shape, 2, m=3)
    lh = gpu_pkg.LinearHomogenization(shape)
    up = gpu_pkg.to_planar(u, shape).float().contiguous()
    out = torch.empty_like(up)
    st = torch.tensor(s, device="cuda")
    lib = gpu_pkg.load_library()
    rc = lib.nlfh_apply_K(lh._plan.handle, ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(up.data_ptr()),
                          ctypes.c_void_p(out.data_ptr()), 3, 1, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = gpu_pkg.from_planar(out).double().cpu().numpy()
    ref = np.stack([O.LinearRef(shape).apply_K(s, x) for x in u])
    assert _rel(got, ref) < 2e-6


@pytest.mark.parametrize("shape,tiling", [((16, 12, 16), (4, 2, 1)), ((10, 10, 32), (4, 4, 2)),
                                          ((8, 10, 48), (3, 2, 3)), ((6, 9, 64), (6, 4, 4))])
def test_apply_K_sixteen_lane_rows(gpu_pkg, shape, tiling):
    """Tilings whose z rows are 16 lanes (nz / zpt = 16) take their z neighbours by rotations inside a 16-lane DPP
    row: FP64 and FP32 against the oracle, with ragged y tiles and x chunks."""
    import ctypes
    lh = gpu_pkg.LinearHomogenization(shape)
    lh._plan.set_tiling(0, *tiling)
    assert lh._plan.get_tiling(0) == tiling
    s, u = _rand(shape, 3)
    ref = np.stack([O.LinearRef(shape).apply_K(s, x) for x in u])
    assert _rel(lh.apply_K(s, u).cpu().numpy(), ref) < 1e-12
    s, u = _rand(shape, 4, m=3)
    up = gpu_pkg.to_planar(u, shape).float().contiguous()
    out = torch.empty_like(up)
    st = torch.tensor(s, device="cuda")
    lib = gpu_pkg.load_library()
    rc = lib.nlfh_apply_K(lh._plan.handle, ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(up.data_ptr()),
                          ctypes.c_void_p(out.data_ptr()), 3, 1, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = gpu_pkg.from_planar(out).double().cpu().numpy()
    ref = np.stack([O.LinearRef(shape).apply_K(s, x) for x in u])
    assert _rel(got, ref) < 2e-6


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 4, 5), (8, 8, 1), (12, 10, 6)])
def test_load_and_residual(gpu_pkg, shape):
    s, u = _rand(shape, 3, m=6)
    lh = gpu_pkg.LinearHomogenization(shape)
    ref = O.LinearRef(shape)
    Es = np.random.default_rng(4).standard_normal((6, 6))
    assert _rel(lh.load(s, Es).cpu().numpy(), ref.load(s, Es)) < 1e-12
    assert _rel(lh.residual(s, Es, u).cpu().numpy(), ref.residual(s, Es, u)) < 1e-12


# cubic sides 16, 32, 48, 64 and 128 take the register-FFT kernels, the others the generic ones
@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 4, 5), (8, 8, 1), (16, 12, 10), (24, 48, 20), (8, 8, 128),
                                   (16, 16, 16), (32, 32, 32), (48, 48, 48), (64, 64, 64), (128, 128, 128)])
def test_precond_fp64(gpu_pkg, shape):
    rng = np.random.default_rng(5)
    r = rng.standard_normal((2,) + shape + (3,))
    lh = gpu_pkg.LinearHomogenization(shape)
    got = lh.precond(r).cpu().numpy()
    ref = O.LinearRef(shape).precond(r)
    assert _rel(got, ref) < 1e-11


@pytest.mark.parametrize("shape", [(16, 12, 10), (32, 32, 32), (48, 48, 48), (64, 64, 64), (128, 128, 128)])
def test_precond_fp32(gpu_pkg, shape):
    import ctypes
    rng = np.random.default_rng(6)
    r = rng.standard_normal((2,) + shape + (3,))
    lh = gpu_pkg.LinearHomogenization(shape)
    rp = gpu_pkg.to_planar(r, shape).float().contiguous()
    z = torch.empty_like(rp)
    lib = gpu_pkg.load_library()
    rc = lib.nlfh_precond(lh._plan.handle, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(z.data_ptr()), 2, 1,
                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = gpu_pkg.from_planar(z).double().cpu().numpy()
    ref = O.LinearRef(shape).precond(r)
    assert _rel(got, ref) < 1e-5


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 5, 4), (8, 8, 1)])
def test_symbol_matches_reference_construction(gpu_pkg, shape):
    lh = gpu_pkg.LinearHomogenization(shape)
    got = lh.Kinv.cpu().numpy()
    ref = O.LinearRef(shape).Kinv
    assert np.abs(ref.imag).max() < 1e-12 * np.abs(ref.real).max()
    assert _rel(got.real, ref.real) < 1e-12


def test_stiffness_scale_and_projection(gpu_pkg):
    shape = (6, 5, 4)
    rng = np.random.default_rng(7)
    rho = rng.uniform(-0.2, 1.2, shape)
    lh = gpu_pkg.LinearHomogenization(shape, emin=1e-4, p=3.0)
    ref = O.LinearRef(shape, emin=1e-4, p=3.0)
    assert _rel(lh.stiffness_scale(rho).cpu().numpy(), ref.stiffness_scale(rho)) < 1e-15
    r = torch.tensor(rho, device="cuda", requires_grad=True)
    s = lh.stiffness_scale(r)
    s.sum().backward()
    assert _rel(r.grad.cpu().numpy(), ref.dstiffness_drho(rho)) < 1e-14
    v = gpu_pkg.to_planar(rng.standard_normal((3,) + shape + (3,)) + 5.0, shape)
    vref = gpu_pkg.from_planar(v).cpu().numpy().copy()
    lh._project_raw(v)
    assert _rel(gpu_pkg.from_planar(v).cpu().numpy(), O.proj(vref)) < 1e-13


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 4, 5), (8, 8, 1), (16, 12, 10)])
def test_bilinear_and_element_energy(gpu_pkg, shape):
    s, u = _rand(shape, 8, m=6)
    lh = gpu_pkg.LinearHomogenization(shape)
    ref = O.LinearRef(shape)
    Es = np.eye(6)
    C = np.asarray(lh.bilinear(s, Es, u).detach())
    assert _rel(C, ref.bilinear(s, Es, u)) < 1e-12
    st = torch.tensor(s, device="cuda")
    _, dcds = lh._bilinear_raw(st, Es, gpu_pkg.to_planar(u, shape), True)
    b = ref.element_energy(s, Es, u) / np.prod(shape)
    assert _rel(dcds.cpu().numpy(), b) < 1e-12


def test_host_parameters_are_captured_at_call_time(gpu_pkg):
    """The strain matrix is read from host memory during the call: overwriting the caller's buffer
    right after the (stream-ordered, unsynchronised) calls must not change what the kernels use."""
    import ctypes
    n = 8
    lib = gpu_pkg.load_library()
    lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3)
    rng = np.random.default_rng(3)
    s = torch.tensor(rng.uniform(0.1, 1.0, (n, n, n)), device="cuda")
    u = torch.tensor(rng.standard_normal((6, 3, n, n, n)), device="cuda")
    C_ref, _ = lh._bilinear_raw(s, np.eye(6), u, False)
    f_ref = gpu_pkg.to_planar(lh.load(s, np.eye(6)), (n, n, n))
    torch.cuda.synchronize()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dp = ctypes.POINTER(ctypes.c_double)
    for _ in range(20):
        Es = np.eye(6)
        C = torch.empty((6, 6), dtype=torch.float64, device="cuda")
        f = torch.empty_like(u)
        gpu_pkg._lib.check(lib.nlfh_bilinear(lh._plan.handle, ctypes.c_void_p(s.data_ptr()), Es.ctypes.data_as(dp),
                                             ctypes.c_void_p(u.data_ptr()), 6, ctypes.c_void_p(C.data_ptr()), None, st))
        gpu_pkg._lib.check(lib.nlfh_load(lh._plan.handle, ctypes.c_void_p(s.data_ptr()), Es.ctypes.data_as(dp), 6,
                                         ctypes.c_void_p(f.data_ptr()), st))
        Es[:] = 1e30
        torch.cuda.synchronize()
        assert torch.equal(C.cpu(), C_ref)
        assert torch.equal(f, f_ref)


@pytest.mark.parametrize("n", [16, 32, 48, 64, 128])
def test_precond_fp32_pass_b_block_shapes(gpu_pkg, n):
    """Every FP32 pass B block shape the autotuner may pick (the default, half and double of its (ky, kz)
    positions per block) gives the same preconditioner, against the oracle; unsupported values are refused (two
    pass A planes per block too at 128, whose staged passes work in place with one plane per block)."""
    import ctypes
    shape = (n, n, n)
    rng = np.random.default_rng(60 + n)
    m = 2 if n < 128 else 1
    r = rng.standard_normal((m,) + shape + (3,))
    lh = gpu_pkg.LinearHomogenization(shape)
    plan = lh._plan
    cur, default, planes = plan.get_fft_tiling(0)
    assert cur == default > 0 and planes == 1
    rp = gpu_pkg.to_planar(r, shape).float().contiguous()
    lib = gpu_pkg.load_library()
    ref = O.LinearRef(shape).precond(r) if n <= 64 else None
    outs = []
    for lpb in (default, default // 2, 2 * default):
        plan.set_fft_tiling(0, lpb)
        assert plan.get_fft_tiling(0) == (lpb, default, 1)
        z = torch.empty_like(rp)
        assert lib.nlfh_precond(plan.handle, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(z.data_ptr()), m, 1,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)) == 0
        got = gpu_pkg.from_planar(z).double().cpu().numpy()
        if ref is not None:
            assert _rel(got, ref) < 1e-5
        outs.append(got)
    # the x transforms of a position do not depend on which block holds it: identical results
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    with pytest.raises(gpu_pkg.NlfhError):
        plan.set_fft_tiling(0, 3)
    with pytest.raises(gpu_pkg.NlfhError):
        plan.set_fft_tiling(0, default, 3)
    if n < 128:
        plan.set_fft_tiling(0, 0, 2)
        assert plan.get_fft_tiling(0) == (default, default, 2)
    else:  # the in-place 128-point pass A holds one plane per block
        with pytest.raises(gpu_pkg.NlfhError):
            plan.set_fft_tiling(0, 0, 2)
        assert plan.get_fft_tiling(0) == (2 * default, default, 1)

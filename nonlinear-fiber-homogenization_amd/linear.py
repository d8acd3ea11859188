"""Drop-in `LinearHomogenization` (nlfiber/linear.py:99-264) on the gfx950 library.

Same constructor, attributes, methods, return values and conventions as the reference class; the
arrays are torch tensors on the current HIP device (float64), except the small homogenized
stiffness C, returned on the host so `np.asarray(C)` works as in the reference's callers
(nlfiber/fiber.py:75, scripts/width_pilot.py:43). Every operator runs in libnlfh.so; there is no
CPU path.

Layouts: nodal fields are (m, nx, ny, nz, 3) as in the reference. The library stores them
component-planar (m, 3, nx, ny, nz); the fields returned here are views of that storage through the
axis permutation (0, 2, 3, 4, 1), so handing them back (warm starts, nlfiber/fiber.py:41-56)
costs no copy.

Differentiability (the reference uses jax.vjp, nlfiber/fiber.py:34-36):
  * `stiffness_scale` and `bilinear` are torch autograd functions: dC_IJ/ds_e is the element energy
    divided by N with the correctors held fixed (nlfiber/linear.py:10-13), which the library
    returns from the same kernel that assembles C, and ds/drho comes from the scale kernel;
  * `invariants` is therefore differentiable in rho (correctors detached, nlfiber/linear.py:263);
  * `correctors` has the reference's custom VJP (nlfiber/linear.py:118-134): one adjoint PCG
    solve K lam = -proj(ubar) and the element cross energies of (lam, u).
"""
import ctypes

import numpy as np
import torch

from . import _lib

# --- element constants (nlfiber/linear.py:21-77), shared with the finite-strain code ----------
VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
NODES = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
_g = 0.5 + np.array([-1.0, 1.0]) / (2 * np.sqrt(3.0))
GPTS = np.array([[x, y, z] for z in _g for y in _g for x in _g])


def clip01(x):
    """rho limited to [0, 1]; its derivative is taken as 1 on the closed interval, bounds included
    (nlfiber/linear.py:27-30)."""
    if isinstance(x, torch.Tensor):
        return torch.where(x > 1.0, torch.ones_like(x), torch.where(x < 0.0, torch.zeros_like(x), x))
    x = np.asarray(x)
    return np.where(x > 1.0, 1.0, np.where(x < 0.0, 0.0, x))


def dN_ref():
    """Gradients of the 8 trilinear shape functions of the unit cube, evaluated at the 2x2x2 Gauss points;
    indexed [point, node, axis]."""
    d = np.zeros((8, 8, 3))
    for gi, (x, y, z) in enumerate(GPTS):
        for a, (i, j, k) in enumerate(NODES):
            f = [(x if i else 1 - x), (y if j else 1 - y), (z if k else 1 - z)]
            sg = [(1 if i else -1), (1 if j else -1), (1 if k else -1)]
            d[gi, a] = [sg[0] * f[1] * f[2], f[0] * sg[1] * f[2], f[0] * f[1] * sg[2]]
    return d


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def iso_D(E, nu):
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] = lam + 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    return D


def B_matrices():
    """B matrices of the unit Q1 element at its 8 Gauss points, shape (8, 6, 24): Voigt strain (shears as
    engineering strains) from the 24 nodal displacements."""
    dN = dN_ref()
    B = np.zeros((8, 6, 24))
    rows = {0: [(0, 0)], 1: [(1, 1)], 2: [(2, 2)], 3: [(1, 2), (2, 1)], 4: [(0, 2), (2, 0)], 5: [(0, 1), (1, 0)]}
    for g in range(8):
        for a in range(8):
            for r, pairs in rows.items():
                for comp, der in pairs:  # strain row r takes d u_comp / d x_der
                    B[g, r, 3 * a + comp] = dN[g, a, der]
    return B


def unit_strain(I):
    e = np.zeros(6)
    e[I] = 1.0
    return e


# --- device helpers ------------------------------------------------------------------------
def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("nlfh needs a HIP device: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _as_f64(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def _es_host(Es):
    if isinstance(Es, torch.Tensor):
        Es = Es.detach().cpu().numpy()
    Es = np.ascontiguousarray(np.asarray(Es, dtype=np.float64).reshape(-1, 6))
    return Es, Es.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check_dtype(dtype):
    """The reference's `dtype` (nlfiber/linear.py:102,107-111) sets the precision of its operators. This
    class computes and returns float64 (FP32 enters only as the inner solve of precision="mixed"), so any
    other dtype is refused rather than silently replaced."""
    name = str(getattr(dtype, "name", None) or getattr(dtype, "__name__", None) or dtype)
    if "float64" not in name and name not in ("double", "f8", "<f8"):
        raise ValueError(f"dtype {dtype!r} is not supported: LinearHomogenization computes in float64 "
                         "(precision='mixed' runs FP32 inner solves inside FP64 refinement)")


# Grids the FFT preconditioner and the operator kernels take (the plan refuses others)
MAX_NX = 512
MAX_NY = 512
MAX_NZ = 256


def check_grid(shape):
    """Raise ValueError for a cell the library cannot take: every side must be a product of the radices 2,
    3 and 5, with nx, ny <= 512 and nz <= 256 (the preconditioner's line transforms are mixed-radix
    2/3/4/5/8 codelets), and nz must split into at most 64 lanes of 1-4 consecutive z values (the operator
    keeps a row of nz nodes in one wavefront)."""
    if len(shape) != 3:
        raise ValueError(f"cell shape {tuple(shape)} must have three sides")
    for n, lim, ax in zip(shape, (MAX_NX, MAX_NY, MAX_NZ), "xyz"):
        m = int(n)
        if m < 1 or m > lim:
            raise ValueError(f"cell side n{ax} = {n} is outside [1, {lim}]")
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m != 1:
            raise ValueError(f"cell side n{ax} = {n} has a prime factor other than 2, 3, 5, which the FFT "
                             f"preconditioner does not take (the nearest supported sides are "
                             f"{_nearest_smooth(int(n))})")
    nz = int(shape[2])
    if not any(nz % zpt == 0 and nz // zpt <= 64 for zpt in (1, 2, 3, 4)):
        raise ValueError(f"cell side nz = {nz} must be a multiple of 1, 2, 3 or 4 with at most 64 lanes "
                         "(nz = lanes x z values per lane, lanes <= 64)")


def _nearest_smooth(n):
    def smooth(k):
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        return k == 1
    lo = next(k for k in range(n, 0, -1) if smooth(k))
    hi = next(k for k in range(n, 4 * n + 2) if smooth(k))
    return f"{lo} and {hi}"


class Plan:
    """Owner of one nlfh_plan (cells of given shapes, material E, nu)."""

    def __init__(self, shapes, E=1.0, nu=0.3, m_layout=6):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        for sh in dict.fromkeys(self.shapes):
            check_grid(sh)
        lib = _lib.load()
        self.device = _device()
        arr = (ctypes.c_int * (3 * len(self.shapes)))(*[v for s in self.shapes for v in s])
        h = ctypes.c_void_p()
        _lib.check(lib.nlfh_plan_create(arr, len(self.shapes), float(E), float(nu), int(m_layout), ctypes.byref(h)))
        self.handle = h
        self._lib = lib

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.nlfh_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def set_tiling(self, shape_index, ty, xc, zpt):
        _lib.check(self._lib.nlfh_plan_set_tiling(self.handle, shape_index, ty, xc, zpt))

    def set_fft_tiling(self, shape_index, pass_b_positions, pass_a_planes=1):
        """FP32 preconditioner block shapes of a cubic 16-128 shape: pass B (ky, kz) positions per block (0 or the
        default, half or double of it) and x-planes per block of the spectral pass A (1 or 2)."""
        _lib.check(self._lib.nlfh_plan_set_fft_tiling(self.handle, shape_index, int(pass_b_positions),
                                                      int(pass_a_planes)))

    def get_fft_tiling(self, shape_index):
        """(pass B positions per block in use, its default, pass A planes per block); the pass B values are 0 for
        the generic passes."""
        v, d, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.nlfh_plan_get_fft_tiling(self.handle, shape_index, ctypes.byref(v), ctypes.byref(d),
                                                      ctypes.byref(q)))
        return v.value, d.value, q.value

    def set_inner_restart(self, restart):
        """Mixed precision: restart each FP32 inner PCG from p = z (True) or carry the last search direction
        over the refinement steps (False, the default; cubic 16-128 cells only, other plans restart)."""
        _lib.check(self._lib.nlfh_plan_set_inner_restart(self.handle, 1 if restart else 0))

    def get_tiling(self, shape_index):
        ty, xc, zpt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.nlfh_plan_get_tiling(self.handle, shape_index, ctypes.byref(ty), ctypes.byref(xc),
                                                  ctypes.byref(zpt)))
        return ty.value, xc.value, zpt.value

    def nbytes(self):
        return int(self._lib.nlfh_plan_bytes(self.handle))


def to_planar(u, shape):
    """(m, nx, ny, nz, 3) or (nx, ny, nz, 3) -> contiguous (m, 3, nx, ny, nz) float64 on the device."""
    dev = _device()
    u = _as_f64(u, dev)
    if u.dim() == 4:
        u = u.unsqueeze(0)
    if tuple(u.shape[1:4]) != tuple(shape) or u.shape[4] != 3:
        raise ValueError(f"field of shape {tuple(u.shape)} does not match the cell {tuple(shape)} x 3")
    return u.permute(0, 4, 1, 2, 3).contiguous()


def from_planar(x):
    """(m, 3, nx, ny, nz) storage -> (m, nx, ny, nz, 3) view."""
    return x.permute(0, 2, 3, 4, 1)


class _ScaleFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, lh):
        s = torch.empty_like(rho)
        ds = torch.empty_like(rho) if rho.requires_grad else None
        lh._scale_raw(rho, s, ds)
        ctx.save_for_backward(ds)
        return s

    @staticmethod
    def backward(ctx, gs):
        (ds,) = ctx.saved_tensors
        return gs * ds, None


class _BilinearFn(torch.autograd.Function):
    """C (host, m x m) from s (device) with the correctors fixed; grad_s = sum_IJ gC_IJ dC_IJ/ds."""

    @staticmethod
    def forward(ctx, s, lh, Es, u_planar):
        need = s.requires_grad
        C, dcds = lh._bilinear_raw(s, Es, u_planar, need)
        ctx.save_for_backward(dcds if need else torch.empty(0))
        ctx.need = need
        return C

    @staticmethod
    def backward(ctx, gC):
        if not ctx.need:
            return None, None, None, None
        (dcds,) = ctx.saved_tensors
        g = gC.to(device=dcds.device, dtype=torch.float64)
        m = g.shape[0]
        gs = torch.tensordot(g.reshape(m * m), dcds.reshape(m * m, -1), dims=1).reshape(dcds.shape[2:])
        return gs, None, None, None


class _CorrectorsFn(torch.autograd.Function):
    """u(s) with the adjoint VJP of nlfiber/linear.py:118-134."""

    @staticmethod
    def forward(ctx, s, lh, Es, u0):
        u, _ = lh.solve(s.detach(), Es, u0)
        up = u.permute(0, 4, 1, 2, 3).contiguous()
        ctx.save_for_backward(s.detach(), up)
        ctx.lh, ctx.Es = lh, Es
        return from_planar(up.clone())

    @staticmethod
    def backward(ctx, ubar):
        s, up = ctx.saved_tensors
        lh, Es = ctx.lh, ctx.Es
        rhs = -to_planar(ubar, lh.shape)
        lh._project_raw(rhs)
        lam = lh._solve_rhs_raw(s, rhs)  # K lam = -proj(ubar)
        lh._project_raw(lam)
        Es_np, _ = _es_host(Es)
        g = torch.zeros(lh.shape, dtype=torch.float64, device=s.device)
        n = float(np.prod(lh.shape))
        for I in range(Es_np.shape[0]):
            pair = torch.stack([lam[I], up[I]]).contiguous()
            E2 = np.stack([np.zeros(6), Es_np[I]])
            _, dcds = lh._bilinear_raw(s, E2, pair, True)
            g += dcds[0, 1] * n  # b_e(lam_I, E_I + u_I) = lam_e . Ke u_e + lam_e . G^T E_I
        return g, None, None, None


class LinearHomogenization:
    """Small-strain homogenization of voxel cells of one grid size; `shape` is (nx, ny, nz), or (n, n) for a
    plane-strain cell.

    Extra keywords (not in the reference): precision = "fp64" (the reference's PCG, default) or
    "mixed" (FP32 PCG inside FP64 iterative refinement, true-residual stopping rule; inner_restart=True
    restarts each inner PCG from p = z instead of carrying the search direction over); warm_start = a
    `learned.WarmStartNet` (or the path of its weights) whose Ritz-projected prediction starts
    `C_full` when no u0 is given, with a plain zero-start solve as the fallback if that solve does
    not converge."""

    def __init__(self, shape, E=1.0, nu=0.3, emin=1e-3, p=3.0, dtype=torch.float64, tol=1e-9, maxit=4000,
                 precision="fp64", warm_start=None, warm_smoothing=(128, 900.0), inner_restart=False):
        shape = tuple(int(n) for n in shape)
        self.plane = len(shape) == 2
        self.shape = shape + (1,) if self.plane else shape
        self.E, self.nu, self.emin, self.p = E, nu, emin, p
        _check_dtype(dtype)
        check_grid(self.shape)
        self.dtype, self.tol, self.maxit = torch.float64, tol, maxit
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self._dev = _device()
        B = B_matrices()
        D = iso_D(E, nu)
        self.Dm = torch.as_tensor(D, device=self._dev)
        self.Ke = torch.as_tensor(sum(B[g].T @ D @ B[g] for g in range(8)) / 8.0, device=self._dev)
        self.G = torch.as_tensor(sum(D @ B[g] for g in range(8)) / 8.0, device=self._dev)
        self._plan = Plan([self.shape], E, nu)
        self._plan.set_inner_restart(inner_restart)
        self._lib = _lib.load()
        self.last_info = {}
        self.correctors = lambda s, Es, u0: _CorrectorsFn.apply(_as_f64(s, self._dev), self, Es, u0)
        if isinstance(warm_start, str):
            from .learned import WarmStartNet
            warm_start = WarmStartNet.load(warm_start)
        self.warm_start = warm_start
        self.warm_smoothing = warm_smoothing

    # --- internals ------------------------------------------------------------------------
    @property
    def _N(self):
        return int(np.prod(self.shape))

    def _scale_raw(self, rho, s, ds):
        _lib.check(self._lib.nlfh_stiffness_scale(self._plan.handle, _ptr(rho), float(self.emin), float(self.p),
                                                  _ptr(s), _ptr(ds), _stream()))

    def _bilinear_raw(self, s, Es, u_planar, want_grad):
        Es_np, Es_p = _es_host(Es)
        m = Es_np.shape[0]
        s = s.detach().contiguous()
        if m > 6:
            # the library forms at most 6 x 6 at once: every pair (I, J) from a 2-strain call
            C = torch.empty((m, m), dtype=torch.float64)
            dcds = torch.empty((m, m) + self.shape, dtype=torch.float64, device=self._dev) if want_grad else None
            for I in range(m):
                for J in range(I, m):
                    idx = [I, J] if I != J else [I]
                    Cp, dp = self._bilinear_raw(s, Es_np[idx], u_planar[idx].contiguous(), want_grad)
                    C[I, J] = C[J, I] = Cp[0, -1]
                    if want_grad:
                        dcds[I, J] = dp[0, -1]
                        dcds[J, I] = dp[-1, 0]
            return C, dcds
        C = torch.empty((m, m), dtype=torch.float64, device=self._dev)
        dcds = torch.empty((m, m) + self.shape, dtype=torch.float64, device=self._dev) if want_grad else None
        _lib.check(self._lib.nlfh_bilinear(self._plan.handle, _ptr(s), Es_p, _ptr(u_planar), m, _ptr(C), _ptr(dcds),
                                           _stream()))
        return C.cpu(), dcds

    def _project_raw(self, v):
        for i in range(0, v.shape[0], 6):
            blk = v[i:i + 6]
            _lib.check(self._lib.nlfh_project_mean(self._plan.handle, _ptr(blk), blk.shape[0], _stream()))

    def _solve_rhs_raw(self, s, b_planar):
        """x with K x = b for every field of b (m, 3, nx, ny, nz), in blocks of at most 6 right-hand sides
        (FP64 PCG from zero, the reference's _solve_rhs); last_info holds the largest iteration count."""
        m = b_planar.shape[0]
        x = torch.empty_like(b_planar)
        its, conv, rel = [], [], []
        for i in range(0, m, 6):
            k = min(6, m - i)
            it = (ctypes.c_int * 1)()
            cv = (ctypes.c_int * 1)()
            rr = (ctypes.c_double * k)()
            _lib.check(self._lib.nlfh_solve_rhs(self._plan.handle, _ptr(s.contiguous()), _ptr(b_planar[i:i + k]), k,
                                                _ptr(x[i:i + k]), float(self.tol), int(self.maxit), 0, it, cv, rr,
                                                _stream()))
            its.append(it[0])
            conv.append(bool(cv[0]))
            rel += list(rr)
        self.last_info = dict(it=max(its), converged=all(conv), relres=rel)
        return x

    def _s(self, s):
        s = _as_f64(s, self._dev)
        if self.plane and s.dim() == 2:
            s = s[..., None]
        return s.contiguous()

    # --- operators (nlfiber/linear.py:137-170) -----------------------------------------------
    def stiffness_scale(self, rho):
        rho = _as_f64(rho, self._dev)
        if self.plane and rho.dim() == 2:
            rho = rho[..., None]
        rho = rho.contiguous()
        if rho.numel() != self._N:
            flat = rho.reshape(-1, self._N)
            return torch.stack([self.stiffness_scale(r.reshape(self.shape)) for r in flat]).reshape(rho.shape)
        if rho.requires_grad:
            return _ScaleFn.apply(rho, self)
        s = torch.empty_like(rho)
        self._scale_raw(rho, s, None)
        return s

    def apply_K(self, s, u):
        single = u.dim() == 4 if isinstance(u, torch.Tensor) else np.ndim(u) == 4
        up = to_planar(u, self.shape)
        out = torch.empty_like(up)
        s = self._s(s).detach()
        for i in range(0, up.shape[0], 6):
            blk = up[i:i + 6]
            _lib.check(self._lib.nlfh_apply_K(self._plan.handle, _ptr(s), _ptr(blk), _ptr(out[i:i + 6]),
                                              blk.shape[0], 0, _stream()))
        out = from_planar(out)
        return out[0] if single else out

    def load(self, s, Es):
        Es_np, _ = _es_host(Es)
        m = Es_np.shape[0]
        f = torch.empty((m, 3) + self.shape, dtype=torch.float64, device=self._dev)
        sd = self._s(s).detach()
        for i in range(0, m, 6):
            blk, blk_p = _es_host(Es_np[i:i + 6])
            _lib.check(self._lib.nlfh_load(self._plan.handle, _ptr(sd), blk_p, blk.shape[0], _ptr(f[i:i + 6]),
                                           _stream()))
        return from_planar(f)

    def residual(self, s, Es, u):
        Es_np, _ = _es_host(Es)
        up = to_planar(u, self.shape)
        if up.shape[0] != Es_np.shape[0]:
            raise ValueError(f"{up.shape[0]} fields for {Es_np.shape[0]} strains")
        r = torch.empty_like(up)
        sd = self._s(s).detach()
        for i in range(0, up.shape[0], 6):
            blk, blk_p = _es_host(Es_np[i:i + 6])
            _lib.check(self._lib.nlfh_residual(self._plan.handle, _ptr(sd), blk_p, _ptr(up[i:i + 6]), blk.shape[0],
                                               _ptr(r[i:i + 6]), _stream()))
        self._project_raw(r)
        return from_planar(r)

    @property
    def Kinv(self):
        """Preconditioner symbol (nx, ny, nz//2+1, 3, 3), complex128 with zero imaginary part
        (nlfiber/linear.py:154-165)."""
        nx, ny, nz = self.shape
        buf = np.empty((nx, ny, nz // 2 + 1, 6))
        _lib.check(self._lib.nlfh_symbol(self._plan.handle, 0, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        K = np.empty(buf.shape[:3] + (3, 3))
        idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (2, 1): 3, (0, 2): 4, (2, 0): 4, (0, 1): 5, (1, 0): 5}
        for (a, c), k in idx.items():
            K[..., a, c] = buf[..., k]
        return torch.as_tensor(K.astype(np.complex128), device=self._dev)

    def precond(self, r):
        single = r.dim() == 4 if isinstance(r, torch.Tensor) else np.ndim(r) == 4
        rp = to_planar(r, self.shape)
        z = torch.empty_like(rp)
        for i in range(0, rp.shape[0], 6):
            blk = rp[i:i + 6]
            _lib.check(self._lib.nlfh_precond(self._plan.handle, _ptr(blk), _ptr(z[i:i + 6]), blk.shape[0], 0,
                                              _stream()))
        z = from_planar(z)
        return z[0] if single else z

    # --- solver (nlfiber/linear.py:173-217) -------------------------------------------------
    def solve(self, s, Es, u0=None):
        """Correctors for the macroscopic strains Es (m, 6): returns (u (m, nx, ny, nz, 3), it)."""
        Es_np, Es_p = _es_host(Es)
        m = Es_np.shape[0]
        if m > 6:
            us, its = zip(*[self.solve(s, Es_np[i:i + 6], None if u0 is None else u0[i:i + 6])
                            for i in range(0, m, 6)])
            return torch.cat(us), max(its)
        s = self._s(s).detach()
        u0p = None if u0 is None else to_planar(u0.detach() if isinstance(u0, torch.Tensor) else u0, self.shape)
        u = torch.empty((m, 3) + self.shape, dtype=torch.float64, device=self._dev)
        it = (ctypes.c_int * 1)()
        cv = (ctypes.c_int * 1)()
        rr = (ctypes.c_double * m)()
        prec = 0 if self.precision == "fp64" else 1
        _lib.check(self._lib.nlfh_solve(self._plan.handle, _ptr(s), Es_p, m, _ptr(u0p), _ptr(u), float(self.tol),
                                        int(self.maxit), prec, it, cv, rr, _stream()))
        self.last_info = dict(it=it[0], converged=bool(cv[0]), relres=list(rr), precision=self.precision,
                              fallback_fp64=False)
        if prec == 1 and not cv[0]:
            # mixed precision stopped unconverged (iteration cap or refinement stall): plain FP64 PCG
            # from zero, the reference's method
            it_mixed = it[0]
            _lib.check(self._lib.nlfh_solve(self._plan.handle, _ptr(s), Es_p, m, None, _ptr(u), float(self.tol),
                                            int(self.maxit), 0, it, cv, rr, _stream()))
            self.last_info = dict(it=it_mixed + it[0], converged=bool(cv[0]), relres=list(rr),
                                  precision=self.precision, fallback_fp64=True)
            return from_planar(u), it_mixed + it[0]
        return from_planar(u), it[0]

    # --- homogenized stiffness (nlfiber/linear.py:220-264) ------------------------------------
    def bilinear(self, s, Es, u):
        """The stiffness entries C_IJ: the element energy form of the total strains E_I + grad u_I and
        E_J + grad u_J, weighted by s_e and averaged over the elements (a host tensor; torch can differentiate it
        in s)."""
        s = self._s(s) if not isinstance(s, torch.Tensor) or s.device != self._dev else s
        if self.plane and s.dim() == 2:
            s = s[..., None]
        up = to_planar(u.detach() if isinstance(u, torch.Tensor) else u, self.shape)
        return _BilinearFn.apply(s.contiguous(), self, Es, up)

    def cheb_smooth(self, s, Es, u, degree=128, ratio=900.0, hi=None):
        """`degree` Chebyshev-Jacobi steps for K u = -load(s, Es) from u (m, nx, ny, nz, 3) on the spectral interval
        [hi / ratio, hi] of D^-1 K (nlfh_cheb_smooth; hi defaults to the bound lambda_max(Ke) / Ke_00)."""
        Es_np, Es_p = _es_host(Es)
        m = Es_np.shape[0]
        up = to_planar(u, self.shape)
        out = torch.empty_like(up)
        if hi is None:
            hi = float(self._lib.nlfh_jacobi_bound(float(self.E), float(self.nu)))
        _lib.check(self._lib.nlfh_cheb_smooth(self._plan.handle, _ptr(self._s(s).detach()), Es_p, m, _ptr(up),
                                              _ptr(out), int(degree), float(hi) / float(ratio), float(hi), None,
                                              _stream()))
        return from_planar(out)

    def ritz_start(self, s, Es, P):
        """Energy-optimal combination of the predicted fields P (m, nx, ny, nz, 3) for each right-hand
        side (nlfh_ritz_start); returns the start in the reference layout and the (m, m) weights."""
        Es_np, Es_p = _es_host(Es)
        m = Es_np.shape[0]
        Pp = to_planar(P, self.shape)
        out = torch.empty_like(Pp)
        coef = np.zeros(m * m)
        _lib.check(self._lib.nlfh_ritz_start(self._plan.handle, _ptr(self._s(s).detach()), Es_p, m, _ptr(Pp),
                                             _ptr(out), coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             _stream()))
        return from_planar(out), coef.reshape(m, m)

    def C_full(self, rho, u0=None):
        """The full homogenized stiffness: one solve per unit strain, six in all. Returns (C, u, iterations)."""
        s = self.stiffness_scale(rho)
        Es = np.eye(6)
        if u0 is None and self.warm_start is not None and not self.plane:
            r = rho.detach() if isinstance(rho, torch.Tensor) else torch.as_tensor(np.asarray(rho))
            P = self.warm_start.predict(r.to(self._dev, torch.float64))[0]
            if self.warm_smoothing:
                P = self.cheb_smooth(s.detach(), Es, P, *self.warm_smoothing)
            u_w, _ = self.ritz_start(s.detach(), Es, P)
            u, it = self.solve(s.detach(), Es, u_w)
            info = dict(self.last_info, warm_start=True, fallback=False)
            if not info["converged"] or not np.all(np.isfinite(info["relres"])):
                u, it2 = self.solve(s.detach(), Es, None)
                info = dict(self.last_info, warm_start=True, fallback=True, it_warm=it)
                it = it2
            self.last_info = info
        else:
            u, it = self.solve(s.detach(), Es, u0)
        return self.bilinear(s, Es, u), u, it

    def _shear(self):
        return 5 if self.plane else 3

    def sym_cases(self):
        return torch.as_tensor(np.stack([unit_strain(0), unit_strain(self._shear())]), device=self._dev)

    @staticmethod
    def swap_xy(u):
        """For a cell invariant under the exchange of x and y: the yy-strain corrector obtained by mapping the
        xx-strain corrector through that exchange (field and components)."""
        return torch.swapaxes(u, -4, -3)[..., [1, 0, 2]]

    def invariants_from(self, s, u2):
        """C11, C12 and C44 over E (the in-plane shear C66 for a plane cell), assembled from the two correctors u2
        of the xx and shear strains."""
        uyy = self.swap_xy(u2[0])
        u3 = torch.stack([u2[0], uyy, u2[1]])
        Es = np.stack([unit_strain(0), unit_strain(1), unit_strain(self._shear())])
        C = self.bilinear(s, Es, u3)
        return torch.stack([C[0, 0], C[1, 0], C[2, 2]]) / self.E

    def invariants(self, rho, u0=None):
        """Normalised cubic invariants of a cell with the x<->y symmetry, from two solves (xx and shear strain).
        Returns (c, u2, iterations); torch can differentiate c in rho with the correctors held fixed, which is
        exact because the correctors minimise the energy."""
        s = self.stiffness_scale(rho)
        u2, it = self.solve(s.detach(), self.sym_cases(), u0)
        return self.invariants_from(s, u2.detach()), u2, it


# --- host post-processing (nlfiber/linear.py:267-291) ------------------------------------------
def cubic_invariants(C):
    """The three cubic constants of a 6x6 stiffness, each the mean of the entries that cubic symmetry makes equal:
    C11 over the normal diagonal, C12 over the normal off-diagonal pairs, C44 over the shear diagonal."""
    C = np.asarray(C.detach().cpu() if isinstance(C, torch.Tensor) else C)
    return np.array([np.mean([C[0, 0], C[1, 1], C[2, 2]]), np.mean([C[0, 1], C[0, 2], C[1, 2]]),
                     np.mean([C[3, 3], C[4, 4], C[5, 5]])])


def cubic_leakage(C):
    """How far C is from cubic symmetry: the norm of C minus the cubic matrix built from its averaged constants,
    over the norm of C, taken over the whole 6x6 matrix."""
    C = np.asarray(C.detach().cpu() if isinstance(C, torch.Tensor) else C)
    c11, c12, c44 = cubic_invariants(C)
    Cc = np.zeros((6, 6))
    Cc[:3, :3] = c12
    Cc[[0, 1, 2], [0, 1, 2]] = c11
    Cc[[3, 4, 5], [3, 4, 5]] = c44
    return float(np.linalg.norm(C - Cc) / np.linalg.norm(C))


def hashin_shtrikman(v, E=1.0, nu=0.3):
    """Hashin-Shtrikman upper bounds on bulk and shear modulus, (K_HS, G_HS), for an isotropic solid of volume
    fraction v with empty pores."""
    lam, mu = lame(E, nu)
    K = lam + 2 * mu / 3
    K_hs = K + (1 - v) / (-1 / K + 3 * v / (3 * K + 4 * mu)) if v < 1 else K
    G_hs = mu + (1 - v) / (-1 / mu + 6 * v * (K + 2 * mu) / (5 * mu * (3 * K + 4 * mu))) if v < 1 else mu
    return K_hs, G_hs

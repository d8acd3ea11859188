"""Periodic finite-strain homogenization of voxel cells on the GPU (drop-in for nlfiber/mech.py).

The energy, its gradients, the Hessian-vector product, the density sensitivity and the third-order energy
derivatives run in libnlfh.so (`nlfh_fs_*`, FP64); the preconditioner is the linear path's FFT Green operator of
the homogeneous solid (`nlfh_precond`); the CG vectors stay on the device and the Newton / line-search / forcing
logic runs on the host as in the reference.

Kept from the reference (nlfiber/mech.py):
  * Cell(rho, E, nu, material, wls, emin, p): stiffness scale s = emin + (1 - emin) rho^p, the void
    interpolation gamma of Wang et al. (mech.py:91-121), energy(w, Fbar), mean_stress(w, Fbar),
    precond_w(r) (mech.py:141-181), tiled, with_rho, and the material functions (mech.py:31-54);
  * Mixed(cell, free, forcing): unknowns z = (w, free components of Fbar), z0, unpack, proj, tangent,
    hvp_A, macro_diag, precond, solve_linear and the Newton-CG solve with Eisenstat-Walker forcing and the
    Armijo / gradient-decrease line search (mech.py:183-346);
  * Mixed.grad_rho, the adjoint density gradient of a quantity of the equilibrium state (mech.py:348-363), with
    q a torch function (torch autograd stands in for jax.grad; Cell.mean_stress_t is the differentiable Pbar);
  * Mixed.lowest_eigs: scipy's LOBPCG on the Jacobi-scaled Hessian, the block products on the device
    (mech.py:365-409);
  * effective_tangent (9 linear solves with the matrix-free Hessian), fibonacci_sphere and rank_one_min
    (mech.py:411-451);
  * toec and cubic_toec: second and third energy derivatives along the linear correctors from one kernel pass
    (mech.py:454-515).
Differences of form, not of result: the tangent is applied matrix-free (`tangent` returns the state it is
taken at instead of 81 numbers per Gauss point); z, w and Fbar are torch / numpy float64 arrays instead of
JAX arrays; 3x3 results (the returned Fbar, mean_stress) are numpy arrays; toec's C and T are numpy arrays, or
torch tensors differentiable in chi and cell.rho when either requires grad (the reference's is differentiable
in rho through chi and the cell's stiffness; cubic_toec_t is the traceable cubic_toec).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .linear import NODES, LinearHomogenization, VOIGT, dN_ref, lame

STALL_RTOL = 1e-6


def clip01(x):
    """Clamp to [0, 1] (mech.py:31-35); the kernels differentiate it with slope 1 on [0, 1], bounds included."""
    return torch.where(x > 1.0, torch.ones_like(x), torch.where(x < 0.0, torch.zeros_like(x), x))


def psi_linear(F, lam, mu):
    """Linear-elastic energy density of the small strain sym(F - I), for (..., 3, 3) tensors (mech.py:39-43)."""
    H = F - torch.eye(3, dtype=F.dtype, device=F.device)
    eps = 0.5 * (H + H.transpose(-1, -2))
    tr = eps.diagonal(dim1=-2, dim2=-1).sum(-1)
    return 0.5 * lam * tr**2 + mu * (eps * eps).sum(dim=(-2, -1))


def psi_neohooke(F, lam, mu):
    """Compressible neo-Hookean energy density, for (..., 3, 3) tensors (mech.py:46-50)."""
    J = torch.linalg.det(F)
    lnJ = torch.log(J)
    return 0.5 * mu * ((F * F).sum(dim=(-2, -1)) - 3.0) - mu * lnJ + 0.5 * lam * lnJ**2


# the kernels evaluate these materials (and their derivatives) in closed form; the functions above state them
MATERIALS = {"linear": psi_linear, "neohooke": psi_neohooke}

_dp = ctypes.POINTER(ctypes.c_double)


def voigt_tensor(I):
    """3x3 strain tensor whose Voigt vector is the I-th unit vector; a unit engineering shear puts 1/2 in both
    off-diagonal slots (mech.py:57-65)."""
    H = np.zeros((3, 3))
    i, j = VOIGT[I]
    if i == j:
        H[i, i] = 1.0
    else:
        H[i, j] = H[j, i] = 0.5
    return H


def to_voigt(M):
    """The symmetric part of a 3x3 matrix as a Voigt 6-vector in the order xx, yy, zz, yz, xz, xy, the three
    off-diagonal entries counted twice (engineering shears) (mech.py:68-71)."""
    M = np.asarray(M, dtype=np.float64)
    S = 0.5 * (M + M.T)
    return np.array([S[0, 0], S[1, 1], S[2, 2], 2 * S[1, 2], 2 * S[0, 2], 2 * S[0, 1]])


def _linesearch(E, gnorm, x, dx, E0, slope, gn):
    """Step length by halving from 1 (mech.py:74-88): accept the first t with a sufficient energy decrease
    (Armijo, c = 1e-4), or, when the energy change is below its round-off level, with a smaller gradient
    norm than at t = 0. Gives up (None) once t would drop below 1e-6."""
    tol = 64 * np.finfo(np.float64).eps * max(abs(E0), 1e-300)
    t = 1.0
    while t > 1e-6:
        Et = E(x + t * dx)
        if math.isfinite(Et):
            if Et <= E0 + 1e-4 * t * slope:
                return t
            if Et <= E0 + tol and gnorm(x + t * dx) < gn:
                return t
        t *= 0.5
    return None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class _MeanStressFn(torch.autograd.Function):
    """Pbar(w, Fbar; rho) with the backward of the cell Hessian: w-bar = H_wF G, Fbar-bar = H_FF G (one
    Hessian-vector product with direction (0, G)), rho-bar = the density sensitivity of Pbar : G."""

    @staticmethod
    def forward(ctx, w, F, rho, cell):
        Fh = F.detach().to("cpu", torch.float64).numpy().reshape(3, 3)
        _, sc = cell._eval(w.detach(), Fh, 1)
        ctx.cell, ctx.Fh = cell, Fh
        ctx.save_for_backward(w.detach())
        return sc[1:].clone().reshape(3, 3)

    @staticmethod
    def backward(ctx, gP):
        cell, Fh = ctx.cell, ctx.Fh
        (w,) = ctx.saved_tensors
        dF = gP.detach().reshape(-1).to(torch.float64).contiguous()
        gw = gF = gr = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            hw, sc = cell._eval(w, Fh, 2, None, dF)
            gw = hw.reshape(w.shape)
            gF = sc[1:].clone().reshape(3, 3)
        if ctx.needs_input_grad[2]:
            g3, _ = cell._eval(w, Fh, 3, None, dF)
            gr = g3.reshape(cell.shape)
        return gw, gF, gr, None


class Cell:
    """Density field on a periodic voxel grid with its material model (mech.py:91-121): rho holds one value in
    [0, 1] per voxel, shaped (nx, ny, nz), or (n, n) for a plane cell."""

    def __init__(self, rho, E=1.0, nu=0.3, material="neohooke", wls=(500.0, 0.01), emin=1e-3, p=3.0):
        if material not in MATERIALS:
            raise ValueError(f"material must be one of {tuple(MATERIALS)}")
        self._dev = torch.device("cuda", torch.cuda.current_device())
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self._dev)
        if rho.dim() == 2:
            rho = rho[..., None]
        self.rho, self.shape = rho.contiguous(), tuple(int(n) for n in rho.shape)
        self.N = int(np.prod(self.shape))
        self.E, self.nu, self.material, self.wls, self.emin, self.p = E, nu, material, wls, emin, p
        self.lam, self.mu = lame(E, nu)
        self.psi = MATERIALS[material]
        self.wq = 1.0 / (8.0 * self.N)
        self.s = self.scale_of(self.rho)
        self.gamma = self.gamma_of(self.rho)
        self._lib = _lib.load()
        b, r0 = wls if wls is not None else (0.0, 0.0)
        self._mat = np.array([self.lam, self.mu, emin, p, b, r0], dtype=np.float64)
        self._flags = (1 if material == "neohooke" else 0) | (2 if wls is not None else 0)
        self._shape_c = (ctypes.c_int * 3)(*self.shape)
        nwork = int(self._lib.nlfh_fs_work_doubles(*self.shape))
        self._work = torch.empty(nwork, dtype=torch.float64, device=self._dev)
        self._scal = torch.empty(10, dtype=torch.float64, device=self._dev)
        self._sg = torch.empty(4 * self.N, dtype=torch.float64, device=self._dev)
        _lib.check(self._lib.nlfh_fs_prepare(self._shape_c, _ptr(self.rho), self._mat.ctypes.data_as(_dp),
                                             self._flags, _ptr(self._sg), _stream()))
        self._lh = None

    def scale_of(self, rho):
        rho = torch.as_tensor(rho, dtype=torch.float64)
        return self.emin + (1.0 - self.emin) * rho.clamp(0.0, 1.0) ** self.p

    def gamma_of(self, rho):
        if self.wls is None:
            return None
        b, r0 = self.wls
        v = torch.as_tensor(rho, dtype=torch.float64).clamp(0.0, 1.0) ** self.p
        return (math.tanh(b * r0) + torch.tanh(b * (v - r0))) / (math.tanh(b * r0) + math.tanh(b * (1 - r0)))

    def with_rho(self, rho):
        return Cell(rho, self.E, self.nu, self.material, self.wls, self.emin, self.p)

    def tiled(self, k):
        """The cell repeated k times along each axis (along x and y only for a plane cell) (mech.py:118-121)."""
        reps = (k, k, 1) if self.shape[2] == 1 else (k, k, k)
        return Cell(self.rho.repeat(*reps), self.E, self.nu, self.material, self.wls, self.emin, self.p)

    @property
    def Kinv(self):
        """Inverse symbol of the unit-stiffness linear operator of this cell, (nx, ny, nz//2+1, 3, 3) complex
        (mech.py:175-178): N times the linear path's symbol."""
        if self._lh is None:
            self._lh = LinearHomogenization(self.shape, self.E, self.nu, self.emin, self.p)
        return self._lh.Kinv * float(self.N)

    # --- operators ---------------------------------------------------------------------------------------
    def _sg_for(self, rho):
        """Per-element scale / interpolation factors of another density on this cell's grid (the reference's
        energy(w, Fbar, rho), mech.py:141-145, evaluates the cell at a density other than its own)."""
        r = torch.as_tensor(rho, dtype=torch.float64, device=self._dev).reshape(self.shape).contiguous()
        sg = torch.empty(4 * self.N, dtype=torch.float64, device=self._dev)
        _lib.check(self._lib.nlfh_fs_prepare(self._shape_c, _ptr(r), self._mat.ctypes.data_as(_dp), self._flags,
                                             _ptr(sg), _stream()))
        return sg

    def _eval(self, w, Fbar, mode, v=None, dFbar=None, out=None, scalars=None, sg=None):
        """One nlfh_fs_eval call; returns (grad or None, the 10 scalars as a device tensor, unsynchronised).
        scalars: optional bit mask of the scalars to form (bit q: scalar q; the others come back as 0).
        dFbar (modes 2, 3): a 3x3 host array or a device tensor of 9 values. Mode 3 returns the per-element
        density sensitivity (N values) as grad."""
        w = torch.as_tensor(w, dtype=torch.float64, device=self._dev).reshape(-1).contiguous()
        if w.numel() != 3 * self.N:
            raise ValueError(f"w must hold {3 * self.N} values (shape {self.shape + (3,)})")
        F = np.ascontiguousarray(np.asarray(Fbar, dtype=np.float64).reshape(9))
        grad = (torch.empty(self.N, dtype=torch.float64, device=self._dev) if mode == 3 else torch.empty_like(w)) \
            if mode >= 1 and out is None else out
        if v is not None:
            v = torch.as_tensor(v, dtype=torch.float64, device=self._dev).reshape(-1).contiguous()
            if v.numel() != 3 * self.N:
                raise ValueError(f"v must hold {3 * self.N} values (shape {self.shape + (3,)})")
        dF = None
        if mode >= 2:
            dF = torch.as_tensor(np.zeros(9) if dFbar is None else dFbar, dtype=torch.float64,
                                 device=self._dev).reshape(-1).contiguous()
            if dF.numel() != 9:
                raise ValueError("dFbar must hold 9 values")
        flags = self._flags if scalars is None else self._flags | 4 | (int(scalars) << 8)
        _lib.check(self._lib.nlfh_fs_eval(self._shape_c, _ptr(self._sg if sg is None else sg),
                                          self._mat.ctypes.data_as(_dp), flags,
                                          _ptr(w), F.ctypes.data_as(_dp), _ptr(v), _ptr(dF), mode, _ptr(grad),
                                          _ptr(self._scal), _ptr(self._work), _stream()))
        return grad, self._scal

    def energy(self, w, Fbar, rho=None):
        """W(w, Fbar) = wq sum_e s_e sum_g psi_eff(Fbar + grad w at g) (mech.py:141-154), a Python float; rho:
        another density on this grid (default: the cell's)."""
        return float(self._eval(w, Fbar, 0, sg=None if rho is None else self._sg_for(rho))[1][0])

    def gradient(self, w, Fbar, rho=None):
        """(dW/dw with w's shape, Pbar = dW/dFbar (3, 3)) (the reference's Cell._g and Cell._P); rho as in
        energy."""
        g, sc = self._eval(w, Fbar, 1, sg=None if rho is None else self._sg_for(rho))
        return g.reshape(self.shape + (3,)), sc[1:].cpu().numpy().reshape(3, 3)

    def mean_stress(self, w, Fbar):
        """Pbar = dW/dFbar (mech.py:180-181)."""
        return self.gradient(w, Fbar)[1]

    def hessian_product(self, w, Fbar, v, dFbar):
        """Cell Hessian at (w, Fbar) applied to (v, dFbar): (w part, dW/dFbar part (3, 3))."""
        g, sc = self._eval(w, Fbar, 2, v, dFbar)
        return g.reshape(self.shape + (3,)), sc[1:].cpu().numpy().reshape(3, 3)

    def rho_sensitivity(self, w, Fbar, v, dFbar):
        """d/drho_e of sum_e wq s_e sum_g P_eff(F_g; gamma_e) : (dFbar + grad v at g), per element (nx, ny, nz):
        the density term of Mixed.grad_rho's coupling (mech.py:360-362)."""
        g, _ = self._eval(w, Fbar, 3, v, dFbar)
        return g.reshape(self.shape)

    def mean_stress_t(self, w, Fbar, rho=None):
        """Pbar (3, 3) as a device tensor differentiable (torch autograd) in w, Fbar (a torch 3x3 or host array)
        and this cell's density rho (pass the tensor whose gradient you want; its values must be the cell's): the
        counterpart of differentiating the reference's Cell._P / grad(energy, argnums=1) with jax.grad."""
        F = Fbar if isinstance(Fbar, torch.Tensor) else torch.as_tensor(Fbar, dtype=torch.float64, device=self._dev)
        r = self.rho if rho is None else rho
        return _MeanStressFn.apply(torch.as_tensor(w, dtype=torch.float64, device=self._dev), F, r, self)

    def precond_w(self, r, scale=1.0):
        """The preconditioner's Fourier multiplier: the inverse of the homogeneous (s = 1) operator (mech.py:175-181). That
        operator is the linear path's K(1) divided by N (its energy carries wq = 1/(8N) instead of the
        element matrix's 1/8), so this is N times the linear preconditioner."""
        if self._lh is None:
            self._lh = LinearHomogenization(self.shape, self.E, self.nu, self.emin, self.p)
        r = torch.as_tensor(r, dtype=torch.float64, device=self._dev).reshape(self.shape + (3,))
        return self._lh.precond(r) * (float(self.N) * scale)


class Mixed:
    """Equilibrium under partly prescribed macroscopic deformation (mech.py:183-330). The entries of Fbar listed
    in `free` are solved for, with their average stress driven to zero; every other entry keeps its given value
    (an empty `free` prescribes all nine). The unknown vector z stacks the periodic displacement w and the free
    entries, and the solver minimises W over it."""

    def __init__(self, cell, free=((1, 1), (2, 2)), forcing="ew"):
        self.c = cell
        self.free = [tuple(int(i) for i in f) for f in free]
        self.forcing = forcing
        self.shape = cell.shape
        self.nw = 3 * cell.N
        self.k = len(self.free)
        self.idx = tuple(np.array(self.free).T) if self.k else None
        self._dev = cell._dev
        # row-major flat index of each free component, for device-side gathers and scatters
        self._flat = torch.tensor([3 * a + b for a, b in self.free], dtype=torch.long, device=self._dev)
        # the Hessian products need only the stress sums of the free components (scalar 1 + 3a + b)
        self._smask = sum(1 << (1 + 3 * a + b) for a, b in set(self.free))
        # host copy of the flat indices for the macro scatter / gather kernels, and the dFbar of the products
        self._flat_c = (ctypes.c_int * max(self.k, 1))(*[3 * a + b for a, b in self.free])
        self._dF = torch.zeros(9, dtype=torch.float64, device=self._dev)

    # --- unknowns ----------------------------------------------------------------------------------------
    def z0(self, Fb):
        Fb = np.asarray(Fb, dtype=np.float64)
        mac = torch.as_tensor(Fb[self.idx] if self.k else np.zeros(0), dtype=torch.float64, device=self._dev)
        return torch.cat([torch.zeros(self.nw, dtype=torch.float64, device=self._dev), mac])

    def unpack(self, z, Fb):
        w = z[: self.nw].reshape(self.shape + (3,))
        F = np.array(Fb, dtype=np.float64)
        if self.k:
            F[self.idx] = z[self.nw:].cpu().numpy()
        return w, F

    def unpack_t(self, z, Fb):
        """unpack as torch tensors (w, Fbar 3x3 on the device), differentiable in z (for quantities q(z, Fb, rho)
        handed to grad_rho)."""
        w = z[: self.nw].reshape(self.shape + (3,))
        F = torch.as_tensor(np.asarray(Fb, dtype=np.float64), device=self._dev)
        if self.k:
            F = F.reshape(-1).index_put((self._flat,), z[self.nw:]).reshape(3, 3)
        return w, F

    def proj(self, v):
        """Copy of v with the per-component mean of its w part removed (mech.py:236-239)."""
        out = v.clone()
        _lib.check(self.c._lib.nlfh_fs_project(self.c._shape_c, _ptr(out), _ptr(self.c._work), _stream()))
        return out

    def _E(self, z, Fb):
        return self.c.energy(*self.unpack(z, Fb))

    def Ez(self, z, Fb, rho=None):
        """W at the unknowns z (mech.py:201-203); rho as in Cell.energy."""
        w, F = self.unpack(z, Fb)
        return self.c.energy(w, F, rho)

    def _g(self, z, Fb):
        w, F = self.unpack(z, Fb)
        g, P = self.c.gradient(w, F)
        mac = torch.as_tensor(P[self.idx] if self.k else np.zeros(0), dtype=torch.float64, device=self._dev)
        return torch.cat([g.reshape(-1), mac])

    def tangent(self, z, Fb):
        """The state (w, Fbar) the Hessian is taken at: hvp_A applies its tangent matrix-free."""
        w, F = self.unpack(z, Fb)
        return (w.clone(), F)

    def hvp_A(self, A, v, out=None):
        """Hessian at the state A = (w, Fbar) applied to v (into `out` when given, a contiguous vector of
        nw + k values)."""
        w, F = A
        lib = self.c._lib
        if self.k:  # dFbar from the direction's macro part, on the device
            _lib.check(lib.nlfh_fs_macro_scatter(self.k, self._flat_c, _ptr(v[self.nw:]), _ptr(self._dF), _stream()))
        if out is None:
            out = torch.empty(self.nw + self.k, dtype=torch.float64, device=self._dev)
        _, sc = self.c._eval(w, F, 2, v[: self.nw], self._dF, out=out[: self.nw], scalars=self._smask)
        if self.k:  # the free components of the stress increment sum
            _lib.check(lib.nlfh_fs_macro_gather(self.k, self._flat_c, _ptr(sc[1:]), _ptr(out[self.nw:]), _stream()))
        return out

    def macro_diag(self, A, Fb):
        """Hessian entries H_jj of the free macro entries j, one Hessian product per entry; the preconditioner
        divides their residual components by them (mech.py:240-246)."""
        d = []
        for j in range(self.k):
            e = torch.zeros(self.nw + self.k, dtype=torch.float64, device=self._dev)
            e[self.nw + j] = 1.0
            d.append(float(self.hvp_A(A, e)[self.nw + j]))
        return np.abs(np.array(d)) + 1e-12

    def precond(self, r, dmac, swscale, out=None):
        """The reference's preconditioner (mech.py:248-251): the FFT Green operator of the unit-stiffness solid,
        scaled by 1/swscale, on w, the inverse macro diagonal on the macro unknowns (into `out` when given)."""
        c = self.c
        if c._lh is None:
            c._lh = LinearHomogenization(c.shape, c.E, c.nu, c.emin, c.p)
        if out is None:
            out = torch.empty(self.nw + self.k, dtype=torch.float64, device=self._dev)
        rp = r[: self.nw].reshape(c.N, 3).t().contiguous()  # component-planar (3, N) for the FFT passes
        zp = torch.empty_like(rp)
        _lib.check(c._lib.nlfh_precond(c._lh._plan.handle, _ptr(rp), _ptr(zp), 1, 0, _stream()))
        # back to (N, 3) with the scale N / swscale in one pass
        torch.mul(zp.t(), float(c.N) / swscale, out=out[: self.nw].view(c.N, 3))
        if self.k:
            torch.div(r[self.nw:], torch.as_tensor(dmac, dtype=torch.float64, device=self._dev), out=out[self.nw:])
        return out

    def _proj_(self, v):
        """proj in place on a vector this class owns."""
        _lib.check(self.c._lib.nlfh_fs_project(self.c._shape_c, _ptr(v), _ptr(self.c._work), _stream()))
        return v

    # --- linear solves (mech.py:253-288) ----------------------------------------------------------------
    def _cg(self, A, b, dmac, sws, tol, maxit):
        """Preconditioned CG on the mean-free space with the reference's negative-curvature exit: a
        non-positive p.Ap stops the loop, keeping x (or z0 = M b on the first iteration). The scalars stay on
        the device: the projections of Ap and z leave the partial sums of p.Ap and r.z, which the x, r update
        and the direction update (two fused kernels) sum themselves. The loop test of step k needs p.Ap and
        |r|^2 on the host: they are copied into pinned memory behind an event, and step k + 1 is queued before
        the host waits for them, so the GPU does not idle while the host decides and launches. Three buffer sets
        rotate, so a step queued past the exit never touches the iterate the exit keeps; its results are
        dropped."""
        proj = self.proj
        lib, c = self.c._lib, self.c

        dmac = torch.as_tensor(dmac, dtype=torch.float64, device=self._dev)  # one copy, not one per iteration

        def M(r, out=None):
            return self._proj_(self.precond(r, dmac, sws, out))

        Apb = torch.empty_like(b)
        zb = torch.empty_like(b)
        b = proj(b)
        bn = float(torch.linalg.norm(b)) + 1e-300
        z0 = M(b)
        L = b.numel()
        sets = [[torch.empty_like(b) for _ in range(3)] for _ in range(3)]  # (x, r, p) of step k in set k % 3
        npart = int(lib.nlfh_fs_dot_part_doubles())
        part_pa = torch.empty(npart, dtype=torch.float64, device=self._dev)
        part_rz = torch.empty(npart, dtype=torch.float64, device=self._dev)
        rn2_part = torch.empty((L + 255) // 256, dtype=torch.float64, device=self._dev)
        rzs = torch.empty(2, dtype=torch.float64, device=self._dev)  # r.z read by step k in rzs[k % 2]
        rzs[0] = torch.dot(b, z0)
        s2 = torch.empty((2, 2), dtype=torch.float64, device=self._dev)  # (p.Ap, |r|^2) of step k in s2[k % 2]
        h2 = torch.empty((2, 2), dtype=torch.float64, pin_memory=True)
        events = [torch.cuda.Event(), torch.cuda.Event()]
        stream = torch.cuda.current_stream()
        inputs = {0: (torch.zeros_like(b), b, z0)}  # (x, r, p) each step starts from

        def launch(k):
            x, r, p = inputs[k]
            x1, r1, p1 = sets[k % 3]
            rz, rz1 = rzs[k % 2:k % 2 + 1], rzs[(k + 1) % 2:(k + 1) % 2 + 1]
            sk = s2[k % 2]
            Ap = self.hvp_A(A, p, Apb)
            _lib.check(lib.nlfh_fs_project_dot(c._shape_c, _ptr(Ap), _ptr(p), _ptr(part_pa), _stream()))
            _lib.check(lib.nlfh_fs_cg_update_parts(L, self.k, _ptr(x1), _ptr(r1), _ptr(x), _ptr(r), _ptr(p), _ptr(Ap),
                                                   _ptr(rz), _ptr(part_pa), _ptr(sk[0:1]), _ptr(rn2_part), _stream()))
            z1 = self.precond(r1, dmac, sws, zb)
            _lib.check(lib.nlfh_fs_project_dot(c._shape_c, _ptr(z1), _ptr(r1), _ptr(part_rz), _stream()))
            _lib.check(lib.nlfh_fs_cg_direction_parts(L, self.k, _ptr(p1), _ptr(z1), _ptr(p), _ptr(r1), _ptr(rz),
                                                      _ptr(part_rz), _ptr(rz1), _ptr(rn2_part), _ptr(sk[1:2]),
                                                      _stream()))
            h2[k % 2].copy_(sk, non_blocking=True)
            events[k % 2].record(stream)
            inputs[k + 1] = (x1, r1, p1)

        it, done, conv, x = 0, False, False, None
        if maxit > 0:
            launch(0)
        while not done and it < maxit:
            k = it
            if k + 1 < maxit:
                launch(k + 1)  # queued ahead; dropped if step k ends the loop
            events[k % 2].synchronize()
            pAp_h, rn2_h = h2[k % 2].tolist()
            neg = pAp_h <= 0
            conv1 = math.sqrt(max(rn2_h, 0.0)) <= tol * bn
            x = (z0 if k == 0 else inputs[k][0]) if neg else inputs[k + 1][0]
            inputs.pop(k - 1, None)
            it = k + 1
            done, conv = neg or conv1, conv1 and not neg
        if x is None:
            x = torch.zeros_like(b)
        return x.clone(), it, conv, done and not conv

    def _sws(self):
        # the FFT preconditioner is the homogeneous operator times one number: exp(mean log s), the log-average
        # element stiffness; the spread of s around it is left to the Krylov iteration
        return float(torch.exp(torch.mean(torch.log(self.c.s))))

    def solve_linear(self, A, b, tol=1e-10, maxit=20000):
        """Preconditioned CG on the Hessian at state A, the w part of every iterate kept at zero mean; returns the
        solution, the iteration count and whether the tolerance was met."""
        dmac = self.macro_diag(A, None) if self.k else np.zeros(0)
        x, it, conv, _ = self._cg(A, b, dmac, self._sws(), tol, int(maxit))
        return x, int(it), bool(conv)

    # --- adjoint (mech.py:348-363) -------------------------------------------------------------------------
    def grad_rho(self, z, Fb, q, cg_tol=1e-10, cg_maxit=20000):
        """d/drho of q(z(rho), Fb, rho), with z(rho) the equilibrium for the density rho (mech.py:348-363).
        One adjoint solve with the Hessian, H lam = dq/dz, then the explicit part dq/drho at fixed z minus the
        density derivative of lam . grad_z W. q is a torch function of the device unknowns z and of the cell's
        density, for instance of the macro unknowns z[nw:] or of cell.mean_stress_t(*mix.unpack_t(z, Fb),
        rho); torch autograd differentiates it."""
        Fb = np.asarray(Fb, dtype=np.float64)
        z = torch.as_tensor(z, dtype=torch.float64, device=self._dev)
        zz = z.detach().clone().requires_grad_(True)
        rr = self.c.rho.detach().clone().requires_grad_(True)
        with torch.enable_grad():
            qv = q(zz, Fb, rr)
            gz, gr = torch.autograd.grad(qv, (zz, rr), allow_unused=True)
        gz = torch.zeros_like(z) if gz is None else gz.detach()
        gr = torch.zeros_like(self.c.rho) if gr is None else gr.detach()
        dq_dz = self.proj(gz)
        A = self.tangent(z, Fb)
        lam, its, ok = self.solve_linear(A, dq_dz, cg_tol, cg_maxit)
        w, F = A
        dF = torch.zeros(9, dtype=torch.float64, device=self._dev)
        if self.k:
            dF.index_add_(0, self._flat, lam[self.nw:])
        coupling = self.c.rho_sensitivity(w, F, lam[: self.nw], dF)
        return gr - coupling, dict(cg=its, converged=ok)

    # --- stability (mech.py:365-409) ------------------------------------------------------------------------
    def lowest_eigs(self, z, Fb, k=4, X=None, tol=1e-5, maxiter=2000):
        """The smallest eigenvalues of the Hessian after a symmetric diagonal scaling, D^-1/2 H D^-1/2, with D the
        Hessian's diagonal at F = I over all unknowns including the macro entries, on the space orthogonal to
        rigid translations. A congruence keeps the signs of the eigenvalues (Sylvester's law of inertia), so a
        negative value means H is indefinite. scipy's LOBPCG runs on the host and calls one device Hessian
        product per block column. Returns the eigenvalues, their eigenvectors mapped back to z, and LOBPCG's
        iteration count."""
        from scipy.sparse.linalg import LinearOperator, lobpcg
        Fb = np.asarray(Fb, dtype=np.float64)
        A = self.tangent(torch.as_tensor(z, dtype=torch.float64, device=self._dev), Fb)
        dsc = self._diag_scale()
        nz = self.nw + self.k
        T = np.zeros((nz, 3))
        for i in range(3):
            T[i:self.nw:3, i] = 1.0
        Y, _ = np.linalg.qr(T / dsc[:, None])
        dsc_t = torch.as_tensor(dsc, device=self._dev)

        def mm(v):
            v = np.asarray(v, dtype=np.float64)
            V = v.reshape(nz, -1)
            out = np.empty_like(V)
            for c in range(V.shape[1]):
                x = torch.as_tensor(np.ascontiguousarray(V[:, c]), device=self._dev) * dsc_t
                out[:, c] = (self.hvp_A(A, x) * dsc_t).cpu().numpy()
            return out.reshape(v.shape)

        op = LinearOperator((nz, nz), matvec=mm, matmat=mm, dtype=np.float64)
        if X is None:
            X = np.random.default_rng(0).standard_normal((nz, k))
        vals, vecs, hist = lobpcg(op, X, Y=Y, largest=False, tol=tol, maxiter=maxiter, retLambdaHistory=True)
        o = np.argsort(vals)
        return vals[o], vecs[:, o] * dsc[:, None], len(hist)

    def _diag_scale(self):
        """1 / sqrt|diag H| at the undeformed state: the w block probed with the 8 colours of the 2x2x2
        colouring x 3 components (nodes of one colour do not share an element), the macro entries directly."""
        if not hasattr(self, "_dsc"):
            Fb = np.eye(3)
            A = self.tangent(self.z0(Fb), Fb)
            d = np.zeros(self.nw + self.k)
            dw = d[: self.nw].reshape(self.shape + (3,))
            col = np.indices(self.shape) % 2
            col = col[0] + 2 * col[1] + 4 * col[2]
            for cc in range(8):
                mask = col == cc
                if not mask.any():
                    continue
                for comp in range(3):
                    e = np.zeros(self.shape + (3,))
                    e[..., comp] = mask
                    v = torch.cat([torch.as_tensor(e.ravel(), device=self._dev),
                                   torch.zeros(self.k, dtype=torch.float64, device=self._dev)])
                    hv = self.hvp_A(A, v)[: self.nw].reshape(self.shape + (3,)).cpu().numpy()
                    dw[..., comp][mask] = hv[..., comp][mask]
            for j in range(self.k):
                e = torch.zeros(self.nw + self.k, dtype=torch.float64, device=self._dev)
                e[self.nw + j] = 1.0
                d[self.nw + j] = float(self.hvp_A(A, e)[self.nw + j])
            self._dsc = 1.0 / np.sqrt(np.abs(d) + 1e-30)
        return self._dsc

    # --- Newton-CG (mech.py:290-330) --------------------------------------------------------------------
    def solve(self, Fbar, z0=None, rtol=1e-9, maxnewton=30, cg_tol=1e-6, cg_maxit=20000, g_ref=None):
        """Equilibrium by Newton's method with CG inner solves (mech.py:302-346). Fbar gives the fixed macro
        components, and its free components seed the iteration when no z0 is passed. Returns the fluctuation
        w, the full macro gradient and a log dict (Newton and CG counts, gradient norms, flags; 'z' holds
        the final unknown vector)."""
        Fb = np.asarray(Fbar, dtype=np.float64)
        z = self.z0(Fb) if z0 is None else torch.as_tensor(z0, dtype=torch.float64, device=self._dev)
        g0, log = g_ref, dict(newton=0, cg=0, converged=False, gnorm=[], negative_curvature=False)
        eta, gprev = None, None
        sws = self._sws()
        for k in range(maxnewton):
            g = self.proj(self._g(z, Fb))
            gn = float(torch.linalg.norm(g))
            g0 = gn if g0 is None else g0
            log["gnorm"].append(gn)
            if gn <= rtol * max(g0, 1e-300) or gn < 1e-14:
                log["converged"] = True
                break
            tol = cg_tol
            if self.forcing == "ew":
                e_new = 0.1 if gprev is None else 0.9 * (gn / gprev) ** 2
                if eta is not None and 0.9 * eta**2 > 0.1:
                    e_new = max(e_new, 0.9 * eta**2)
                eta = min(e_new, 0.1)
                tol = max(eta, cg_tol, 0.5 * rtol * g0 / gn) if gprev is not None else max(eta, cg_tol)
            gprev = gn
            A = self.tangent(z, Fb)
            dmac = self.macro_diag(A, Fb) if self.k else np.zeros(0)
            dz, its, ok, neg = self._cg(A, -g, dmac, sws, tol, int(cg_maxit))
            log["cg"] += int(its)
            log["negative_curvature"] |= bool(neg)
            E0, slope = self._E(z, Fb), float(torch.dot(g, dz))
            if slope >= 0:
                dz = -self.proj(self.precond(g, dmac, sws))
                slope = float(torch.dot(g, dz))
            t = _linesearch(lambda x: self._E(x, Fb), lambda x: float(torch.linalg.norm(self.proj(self._g(x, Fb)))),
                            z, dz, E0, slope, gn)
            log["newton"] = k + 1
            if t is None:
                log["linesearch_failed"] = True
                if gn <= STALL_RTOL * g0:
                    log["converged"] = log["roundoff_floor"] = True
                break
            z = z + t * dz
        w, F = self.unpack(z, Fb)
        log["z"], log["g0"] = z, g0
        return w, F, log


# --- effective tangent and strong ellipticity (mech.py:411-451) ---------------------------------------------------
def effective_tangent(cell, w, Fbar, cg_tol=1e-10):
    """Homogenized tangent A_abcd = d Pbar_ab / d Fbar_cd of the equilibrium (w, Fbar), all of Fbar prescribed
    (mech.py:411-427): for each
    unit direction E_cd, the displacement response dw solves H_ww dw = -H_wF E_cd, and A[:, :, c, d] is the stress
    part of the Hessian product with (dw, E_cd). 9 linear solves; a numpy (3, 3, 3, 3) array."""
    full = Mixed(cell, free=[])
    Fb = np.asarray(Fbar, dtype=np.float64)
    z = torch.as_tensor(w, dtype=torch.float64, device=cell._dev).reshape(-1)
    Atan = full.tangent(z, Fb)
    A = np.zeros((3, 3, 3, 3))
    for c in range(3):
        for d in range(3):
            E = np.zeros((3, 3))
            E[c, d] = 1.0
            hw, _ = cell.hessian_product(Atan[0], Fb, None, E)
            dz, _, _ = full.solve_linear(Atan, -hw.reshape(-1), cg_tol)
            _, dP = cell.hessian_product(Atan[0], Fb, dz, E)
            A[:, :, c, d] = dP
    return A


def fibonacci_sphere(m):
    """m nearly uniform unit vectors (golden-angle spiral; mech.py:430-434)."""
    t = (np.arange(m) + 0.5) / m
    polar = np.arccos(1.0 - 2.0 * t)
    azim = np.pi * (1.0 + np.sqrt(5.0)) * np.arange(0.5, m, 1.0)
    return np.stack([np.cos(azim) * np.sin(polar), np.sin(azim) * np.sin(polar), np.cos(polar)], axis=1)


def rank_one_min(A, m=1000, plane=False):
    """Ellipticity margin of a tangent A (mech.py:437-451): over unit directions n (in the x-y plane for a plane
    cell), the least eigenvalue of the symmetrised Q_ik = A_ijkl n_j n_l, and the n that attains it. A value
    <= 0 marks a direction in which the material can form a localised band."""
    if plane:
        th = np.linspace(0.0, np.pi, m, endpoint=False)
        ns = np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], axis=1)
    else:
        ns = fibonacci_sphere(m)
    Q = np.einsum("ijkl,nj,nl->nik", np.asarray(A, dtype=np.float64), ns, ns)
    Q = 0.5 * (Q + np.swapaxes(Q, 1, 2))
    if plane:
        Q = Q[:, :2, :2]
    lo = np.linalg.eigvalsh(Q)[:, 0]
    i = int(np.argmin(lo))
    return float(lo[i]), ns[i]


# --- third-order elastic constants (mech.py:454-515) -----------------------------------------------------------
_PAIRS = [(I, J) for I in range(6) for J in range(I, 6)]
_TRIPLES = [(I, J, K) for I in range(6) for J in range(I, 6) for K in range(J, 6)]


def _toec_kernel(cell, chi_t):
    """The 77 energy derivatives at F = I from `nlfh_fs_toec`: d2f_IJ over the pairs I <= J, then d3f_IJK over the
    triples I <= J <= K, with the cell's own stiffness scale and interpolation factor."""
    out = torch.empty(77, dtype=torch.float64, device=cell._dev)
    _lib.check(cell._lib.nlfh_fs_toec(cell._shape_c, _ptr(cell._sg), cell._mat.ctypes.data_as(_dp), cell._flags,
                                      _ptr(chi_t), _ptr(out), _ptr(cell._work), _stream()))
    return out


def toec_values_t(cell, chi, sw, gam):
    """The kernel's 77 values as a torch function of chi (6 fields (N, 3)), sw = wq s_e and gamma_e (N each): per
    Gauss point H_I = e_I + grad chi_I, d2f_IJ = sum_e sw_e sum_g D2[H_I, H_J] with the linear tangent
    D2[a, b] = lam tr a tr b + mu (a : b + a : b^T) (the interpolation drops out at F = I), and, neo-Hookean only,
    d3f_IJK = sum_e sw_e gamma_e^3 sum_g D3[H_I, H_J, H_K] with
    D3[a, b, c] = -mu (tr abc + tr acb) - lam (tr a tr bc + tr b tr ac + tr c tr ab). The same closed forms as the
    kernel (whose values it reproduces to round-off); toec differentiates it for its backward pass."""
    dev = chi.device
    x = chi.reshape((6,) + cell.shape + (3,))
    nod = torch.stack([torch.roll(x, shifts=tuple(-int(o) for o in off), dims=(1, 2, 3)) for off in NODES], dim=1)
    nod = nod.reshape(6, 8, -1, 3)
    dn = torch.as_tensor(dN_ref(), dtype=torch.float64, device=dev)
    E = torch.stack([torch.as_tensor(voigt_tensor(I), dtype=torch.float64, device=dev) for I in range(6)])
    lam, mu = cell.lam, cell.mu
    neo = cell.material == "neohooke"
    w3 = sw * gam ** 3 if neo else None
    d2 = torch.zeros((6, 6), dtype=torch.float64, device=dev)
    d3 = torch.zeros((6, 6, 6), dtype=torch.float64, device=dev)
    for g in range(8):
        H = torch.einsum("ianc,ad->incd", nod, dn[g]) + E[:, None]
        tr = H.diagonal(dim1=-2, dim2=-1).sum(-1)
        inner = torch.einsum("incd,jncd->ijn", H, H) + torch.einsum("incd,jndc->ijn", H, H)
        d2 = d2 + torch.einsum("ijn,n->ij", lam * tr[:, None] * tr[None] + mu * inner, sw)
        if neo:
            prod = torch.einsum("incd,jnde->ijnce", H, H)
            trp = prod.diagonal(dim1=-2, dim2=-1).sum(-1)
            t3 = torch.einsum("ijncd,kndc->ijkn", prod, H)
            D3 = -mu * (t3 + t3.transpose(1, 2)) - lam * (tr[:, None, None] * trp[None] + tr[None, :, None] *
                                                         trp[:, None, :] + tr[None, None, :] * trp[:, :, None])
            d3 = d3 + torch.einsum("ijkn,n->ijk", D3, w3)
    pi = torch.tensor(_PAIRS, device=dev)
    ti = torch.tensor(_TRIPLES, device=dev)
    return torch.cat([d2[pi[:, 0], pi[:, 1]], d3[ti[:, 0], ti[:, 1], ti[:, 2]]])


class _ToecFn(torch.autograd.Function):
    """The 77 derivatives from the kernel; the backward differentiates toec_values_t (the same closed forms) in
    chi, sw and gamma."""

    @staticmethod
    def forward(ctx, chi_flat, sw, gam, cell):
        ctx.cell = cell
        ctx.save_for_backward(chi_flat, sw, gam)
        return _toec_kernel(cell, chi_flat.detach().contiguous())

    @staticmethod
    def backward(ctx, gv):
        chi_flat, sw, gam = ctx.saved_tensors
        with torch.enable_grad():
            xs = [t.detach().requires_grad_(True) for t in (chi_flat, sw, gam)]
            v = toec_values_t(ctx.cell, *xs)
            grads = torch.autograd.grad(v, xs, gv, allow_unused=True)
        return (*grads, None)


def _geometric_X():
    """X[:, J, K] = voigt(e_J e_K): the Voigt vector of the product of two unit Voigt strains."""
    X = np.zeros((6, 6, 6))
    for J in range(6):
        for K in range(6):
            X[:, J, K] = to_voigt(voigt_tensor(J) @ voigt_tensor(K))
    return X


_SYM3 = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _toec_assemble_t(v):
    """C (6x6) and T = f''' - 3 Sym C X (6x6x6) from the 77 values, in torch (differentiable)."""
    dev = v.device
    idx2 = np.zeros((6, 6), dtype=np.int64)
    for q, (I, J) in enumerate(_PAIRS):
        idx2[I, J] = idx2[J, I] = q
    idx3 = np.zeros((6, 6, 6), dtype=np.int64)
    for q, (I, J, K) in enumerate(_TRIPLES):
        for a, b, c in {(I, J, K), (I, K, J), (J, I, K), (J, K, I), (K, I, J), (K, J, I)}:
            idx3[a, b, c] = 21 + q
    C = v[torch.as_tensor(idx2, device=dev)]
    T3 = v[torch.as_tensor(idx3, device=dev)]
    CX = torch.einsum("ia,ajk->ijk", C, torch.as_tensor(_geometric_X(), dtype=torch.float64, device=dev))
    sym = sum(CX.permute(*p) for p in _SYM3) / 6.0
    return C, T3 - 3.0 * sym


def toec(cell, chi):
    """Second- and third-order elastic constants of the cell, C (6x6) and T (6x6x6) in Lagrangian strain, from the
    linear correctors chi
    (6, nx, ny, nz, 3) of unit Voigt strains (same stiffness scale and E, nu) (mech.py:454-490): with
    f(s) = <psi(I + sum_I s_I (e_I + grad chi_I))>, which equals Wbar to O(s^4), C_IJ = f''_IJ and
    T_IJK = f'''_IJK - 3 Sym C[e_I, e_J e_K]. The second and third derivatives come from one kernel pass
    (`nlfh_fs_toec`, closed-form derivatives of the energy at F = I). numpy arrays; when chi or cell.rho requires
    grad (and grad mode is on), torch tensors on the cell's device, differentiable in both (the reference's toec
    is differentiable in rho through chi and the cell's stiffness, mech.py:458-462)."""
    chi_t = torch.as_tensor(chi, dtype=torch.float64, device=cell._dev).reshape(6, -1)
    if chi_t.shape[1] != 3 * cell.N:
        raise ValueError(f"chi must have shape (6,) + {cell.shape + (3,)}")
    if torch.is_grad_enabled() and (chi_t.requires_grad or cell.rho.requires_grad):
        sw = (cell.wq * cell.scale_of(cell.rho)).reshape(-1)
        gam = cell.gamma_of(cell.rho)
        gam = torch.ones_like(sw) if gam is None else gam.reshape(-1)
        return _toec_assemble_t(_ToecFn.apply(chi_t, sw, gam, cell))
    C, T = _toec_assemble_t(_toec_kernel(cell, chi_t.detach().contiguous()).cpu())
    return C.numpy(), T.numpy()


CUBIC_TOEC = {
    "C111": [(0, 0, 0), (1, 1, 1), (2, 2, 2)],
    "C112": [(0, 0, 1), (0, 0, 2), (1, 1, 0), (1, 1, 2), (2, 2, 0), (2, 2, 1)],
    "C123": [(0, 1, 2)],
    "C144": [(0, 3, 3), (1, 4, 4), (2, 5, 5)],
    "C155": [(0, 4, 4), (0, 5, 5), (1, 3, 3), (1, 5, 5), (2, 3, 3), (2, 4, 4)],
    "C456": [(3, 4, 5)],
}


def cubic_toec(T):
    """Reduce T (6x6x6, Voigt order xx yy zz yz xz xy) to the six independent constants of cubic symmetry,
    C111, C112, C123, C144, C155 and C456, averaging the entries that cubic symmetry makes equal
    (mech.py:493-515)."""
    T = np.asarray(T.detach().cpu() if isinstance(T, torch.Tensor) else T)
    return {k: float(np.mean([T[i] for i in idx])) for k, idx in CUBIC_TOEC.items()}


def cubic_toec_t(T):
    """cubic_toec for a torch T, differentiable: dict of 0-d tensors (the reference's cubic_toec_jnp,
    mech.py:504-506)."""
    T = torch.as_tensor(T)
    return {k: torch.stack([T[i] for i in idx]).mean() for k, idx in CUBIC_TOEC.items()}


cubic_toec_jnp = cubic_toec_t  # the reference's name for the traceable form

"""FP64 CPU restatement of the reference's finite-strain cell operators and Newton solve (nlfiber/mech.py).

TEST INFRASTRUCTURE ONLY. The parity tests call this module to check the HIP path
(`nlfh_fs_eval` and the package's `mech.Cell` / `mech.Mixed`); the product never imports it.

Restated, in numpy float64:
  * the materials psi_linear / psi_neohooke and the void interpolation psi_eff (mech.py:39-50, 107-112,
    134-139), their first Piola stresses, and the tangent d2psi/dF2 as an explicit 3x3x3x3 tensor
    (mech.py:156-167 stores it per Gauss point);
  * the cell energy W(w, Fbar) = wq sum_e s_e sum_g psi_eff(Fbar + H_g), wq = 1/(8N) (mech.py:124-154),
    its gradients in w and Fbar (Cell._g / Cell._P, mech.py:169-174) and the stored-tangent product of
    Mixed.hvp_A (mech.py:218-233);
  * a dense Newton solve of the mixed-control problem (mech.py:183-210 sets it up) for small cells: the
    Hessian is assembled column by column from the tangent product and each Newton step is a direct
    solve on the mean-free space, so the equilibrium it reaches does not depend on CG details.

Pins (tests/test_oracle_mech.py): the stresses against central differences of the energies; the
tangent against central differences of the stresses; the cell gradient and Hessian product against
central differences of the cell energy and gradient; the density sensitivity against central differences in rho;
the neo-Hookean tangent at F = I against the
linear one; for the linear material, W at the linear corrector u_I and Fbar = I + E_I equals C_II / 2
of the linear oracle, with zero gradient in w (the corrector is the equilibrium).
"""
import numpy as np

from .linear_ref import NODES, dN_ref, lame

DN = dN_ref()
I3 = np.eye(3)


def clip01(x):
    return np.where(x > 1.0, 1.0, np.where(x < 0.0, 0.0, x))


# ---- materials (mech.py:39-50) ------------------------------------------------------------------
def psi_lin(F, lam, mu):
    H = F - I3
    e = 0.5 * (H + np.swapaxes(H, -1, -2))
    tr = np.trace(e, axis1=-2, axis2=-1)
    return 0.5 * lam * tr**2 + mu * np.sum(e * e, axis=(-2, -1))


def P_lin(F, lam, mu):
    H = F - I3
    e = 0.5 * (H + np.swapaxes(H, -1, -2))
    tr = np.trace(e, axis1=-2, axis2=-1)
    return lam * tr[..., None, None] * I3 + 2.0 * mu * e


def A_lin(lam, mu):
    d = I3
    return (lam * np.einsum("ij,kl->ijkl", d, d) + mu * (np.einsum("ik,jl->ijkl", d, d) +
                                                        np.einsum("il,jk->ijkl", d, d)))


def psi_neo(F, lam, mu):
    J = np.linalg.det(F)
    lnJ = np.log(J)
    return 0.5 * mu * (np.sum(F * F, axis=(-2, -1)) - 3.0) - mu * lnJ + 0.5 * lam * lnJ**2


def P_neo(F, lam, mu):
    J = np.linalg.det(F)
    lnJ = np.log(J)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    return mu * F + (lam * lnJ - mu)[..., None, None] * FiT


def A_neo(F, lam, mu):
    """d P_ij / d F_kl = mu d_ik d_jl + lam Fi_ji Fi_lk + (mu - lam ln J) Fi_jk Fi_li, per point (..., 3,3,3,3)."""
    Fi = np.linalg.inv(F)
    lnJ = np.log(np.linalg.det(F))
    d = I3
    return (mu * np.einsum("ik,jl->ijkl", d, d) + lam * np.einsum("...ji,...lk->...ijkl", Fi, Fi) +
            (mu - lam * lnJ)[..., None, None, None, None] * np.einsum("...jk,...li->...ijkl", Fi, Fi))


class MatRef:
    """psi_eff, P_eff, A_eff of mech.py:134-139 for per-point gamma (None: no interpolation)."""

    def __init__(self, material, lam, mu):
        self.neo = material == "neohooke"
        self.lam, self.mu = lam, mu

    def _nl(self, F):
        if self.neo:
            return psi_neo(F, self.lam, self.mu), P_neo(F, self.lam, self.mu), A_neo(F, self.lam, self.mu)
        A = np.broadcast_to(A_lin(self.lam, self.mu), F.shape[:-2] + (3, 3, 3, 3))
        return psi_lin(F, self.lam, self.mu), P_lin(F, self.lam, self.mu), A

    def eval(self, F, gamma):
        lam, mu = self.lam, self.mu
        if gamma is None:
            return self._nl(F)
        g = gamma[..., None, None]
        Fg = I3 + g * (F - I3)
        pn, Pn, An = self._nl(Fg)
        Al = np.broadcast_to(A_lin(lam, mu), F.shape[:-2] + (3, 3, 3, 3))
        psi = pn - psi_lin(Fg, lam, mu) + psi_lin(F, lam, mu)
        P = g * (Pn - P_lin(Fg, lam, mu)) + P_lin(F, lam, mu)
        A = (gamma**2)[..., None, None, None, None] * (An - Al) + Al
        return psi, P, A


# ---- the cell (mech.py:91-174) --------------------------------------------------------------------
class CellRef:
    def __init__(self, rho, E=1.0, nu=0.3, material="neohooke", wls=(500.0, 0.01), emin=1e-3, p=3.0):
        rho = np.asarray(rho, dtype=np.float64)
        if rho.ndim == 2:
            rho = rho[..., None]
        self.rho, self.shape = rho, rho.shape
        self.N = int(np.prod(self.shape))
        self.lam, self.mu = lame(E, nu)
        self.mat = MatRef(material, self.lam, self.mu)
        self.wq = 1.0 / (8.0 * self.N)
        v = clip01(rho) ** p
        # derivatives with the reference's clip01 convention (1 on [0, 1], bounds included; mech.py:31-35)
        dv = np.where((rho > 1.0) | (rho < 0.0), 0.0, p * clip01(rho) ** (p - 1.0))
        self.s = emin + (1.0 - emin) * v
        self.ds = (1.0 - emin) * dv
        if wls is None:
            self.gamma = None
            self.dgamma = np.zeros_like(rho)
        else:
            b, r0 = wls
            den = np.tanh(b * r0) + np.tanh(b * (1 - r0))
            self.gamma = (np.tanh(b * r0) + np.tanh(b * (v - r0))) / den
            self.dgamma = b * (1.0 - np.tanh(b * (v - r0)) ** 2) / den * dv

    def nodal(self, w):
        return np.stack([np.roll(w, shift=tuple(-int(o) for o in off), axis=(0, 1, 2)) for off in NODES])

    def grad_w(self, w):
        """(8, nx, ny, nz, 3, 3): H[..., i, j] = d w_i / d x_j at each Gauss point (mech.py:127-129)."""
        return np.einsum("aijkc,gad->gijkcd", self.nodal(w), DN)

    def _scatter_g(self, Pg):
        """Adjoint of grad_w: nodal forces of the Gauss-point stresses Pg (8, nx, ny, nz, 3, 3)."""
        fe = np.einsum("gijkcd,gad->aijkc", Pg, DN)
        out = np.zeros(self.shape + (3,))
        for a, off in enumerate(NODES):
            out += np.roll(fe[a], shift=tuple(int(o) for o in off), axis=(0, 1, 2))
        return out

    def energy(self, w, Fbar):
        F = np.asarray(Fbar) + self.grad_w(w)
        psi, _, _ = self.mat.eval(F, None if self.gamma is None else np.broadcast_to(self.gamma, F.shape[:-2]))
        return self.wq * float(np.sum(self.s * psi))

    def stress(self, w, Fbar):
        """(grad in w (nx, ny, nz, 3), Pbar (3, 3)) = (Cell._g, Cell._P)."""
        F = np.asarray(Fbar) + self.grad_w(w)
        gam = None if self.gamma is None else np.broadcast_to(self.gamma, F.shape[:-2])
        _, P, _ = self.mat.eval(F, gam)
        P = self.wq * self.s[None, ..., None, None] * P
        return self._scatter_g(P), P.sum(axis=(0, 1, 2, 3))

    def tangent(self, w, Fbar):
        """wq s d2 psi / dF2 per Gauss point (8, nx, ny, nz, 3, 3, 3, 3) (Cell.tangent_g for all g)."""
        F = np.asarray(Fbar) + self.grad_w(w)
        gam = None if self.gamma is None else np.broadcast_to(self.gamma, F.shape[:-2])
        _, _, A = self.mat.eval(F, gam)
        return self.wq * self.s[None, ..., None, None, None, None] * A

    def work(self, w, Fbar, v, dFbar):
        """sum_e wq s_e sum_g P_eff(F_g; gamma_e) : dL_g with dL = dFbar + grad v: the pairing whose density
        derivative Mixed.grad_rho's coupling term takes (mech.py:360-362)."""
        F = np.asarray(Fbar) + self.grad_w(w)
        gam = None if self.gamma is None else np.broadcast_to(self.gamma, F.shape[:-2])
        _, P, _ = self.mat.eval(F, gam)
        dL = np.asarray(dFbar) + self.grad_w(v)
        return self.wq * float(np.sum(self.s * np.sum(P * dL, axis=(-2, -1))))

    def rho_sensitivity(self, w, Fbar, v, dFbar):
        """d work / d rho_e per element (nx, ny, nz), from the explicit tangent: s' P : dL + s gamma' dP/dgamma : dL
        with dP/dgamma = P_nl(Fg) - P_lin(Fg) + gamma (A_nl(Fg) - A_lin) : (F - I)."""
        F = np.asarray(Fbar) + self.grad_w(w)
        dL = np.asarray(dFbar) + self.grad_w(v)
        if self.gamma is None:
            _, P, _ = self.mat.eval(F, None)
            return self.wq * self.ds * np.sum(P * dL, axis=(0, -2, -1))
        gam = np.broadcast_to(self.gamma, F.shape[:-2])
        _, P, _ = self.mat.eval(F, gam)
        H = F - I3
        Fg = I3 + gam[..., None, None] * H
        _, Pn, An = self.mat._nl(Fg)
        Pl = P_lin(Fg, self.lam, self.mu)
        Al = A_lin(self.lam, self.mu)
        dPg = Pn - Pl + gam[..., None, None] * np.einsum("...abcd,...cd->...ab", An - Al, H)
        t = np.sum(P * dL, axis=(0, -2, -1))
        u = np.sum(dPg * dL, axis=(0, -2, -1))
        return self.wq * (self.ds * t + self.s * self.dgamma * u)

    def hvp(self, A, v, dFbar):
        """Hessian product (w part, sum of dP) with the stored tangent A (Mixed.hvp_A, mech.py:227-231)."""
        dF = np.asarray(dFbar) + self.grad_w(v)
        dP = np.einsum("gijkabcd,gijkcd->gijkab", A, dF)
        return self._scatter_g(dP), dP.sum(axis=(0, 1, 2, 3))


# ---- mixed control, dense Newton (small cells) -----------------------------------------------------
def newton_dense(cell, Fbar, free=(), rtol=1e-12, maxnewton=30):
    """Equilibrium of W(w, Fbar) with the free components of Fbar as unknowns (zero conjugate stress):
    Newton with the Hessian assembled from hvp columns and a direct solve on the mean-free space."""
    Fb = np.array(Fbar, dtype=np.float64)
    free = [tuple(f) for f in free]
    k = len(free)
    shape = cell.shape
    nw = 3 * cell.N
    w = np.zeros(shape + (3,))

    def grad(w, Fb):
        gw, P = cell.stress(w, Fb)
        gw = gw - gw.mean(axis=(0, 1, 2))
        return np.concatenate([gw.ravel(), np.array([P[a, b] for a, b in free])])

    g0 = None
    for _ in range(maxnewton):
        g = grad(w, Fb)
        gn = np.linalg.norm(g)
        g0 = gn if g0 is None else g0
        if gn <= rtol * max(g0, 1e-300) or gn < 1e-14:
            break
        A = cell.tangent(w, Fb)
        H = np.empty((nw + k, nw + k))
        for c in range(nw + k):
            e = np.zeros(nw + k)
            e[c] = 1.0
            dF = np.zeros((3, 3))
            if c >= nw:
                a, b = free[c - nw]
                dF[a, b] = 1.0
            hw, hP = cell.hvp(A, e[:nw].reshape(shape + (3,)), dF)
            hw = hw - hw.mean(axis=(0, 1, 2))
            H[:, c] = np.concatenate([hw.ravel(), np.array([hP[a, b] for a, b in free])])
        # the mean translation is a null direction: pin it with the mean-free constraint
        Cm = np.zeros((3, nw + k))
        for c in range(3):
            Cm[c, c:nw:3] = 1.0
        Kkt = np.block([[H, Cm.T], [Cm, np.zeros((3, 3))]])
        dz = np.linalg.solve(Kkt, np.concatenate([-g, np.zeros(3)]))[:nw + k]
        w = w + dz[:nw].reshape(shape + (3,))
        for j, (a, b) in enumerate(free):
            Fb[a, b] += dz[nw + j]
    return w, Fb


# ---- third-order constants (mech.py:454-490) ---------------------------------------------------------------
def _voigt_tensor(I):
    from .linear_ref import VOIGT
    H = np.zeros((3, 3))
    i, j = VOIGT[I]
    if i == j:
        H[i, i] = 1.0
    else:
        H[i, j] = H[j, i] = 0.5
    return H


def toec_derivatives(cell, chi, h=1e-4):
    """(d2f (6, 6), d3f (6, 6, 6)) of f(s) = wq sum_e s_e sum_g psi_eff(I + sum_I s_I H_I,g), H_I = e_I + grad chi_I:
    d2f from the explicit tangent at F = I, d3f by central differences of that second derivative along each
    e_K (O(h^2))."""
    Hs = np.stack([_voigt_tensor(I) + cell.grad_w(np.asarray(chi[I])) for I in range(6)])  # (6, 8, nx, ny, nz, 3, 3)
    gam = None if cell.gamma is None else np.broadcast_to(cell.gamma, Hs.shape[1:-2])

    def d2(sv):
        F = I3 + np.tensordot(sv, Hs, axes=1)
        _, _, A = cell.mat.eval(F, gam)
        AH = np.einsum("gxyzabcd,Jgxyzcd->Jgxyzab", A, Hs)
        return cell.wq * np.einsum("xyz,Igxyzab,Jgxyzab->IJ", cell.s, Hs, AH)

    C = d2(np.zeros(6))
    T3 = np.zeros((6, 6, 6))
    for K in range(6):
        e = np.zeros(6)
        e[K] = h
        T3[:, :, K] = (d2(e) - d2(-e)) / (2 * h)
    return C, T3

"""Many cells, of uneven grid sizes, homogenized together: one plan, one launch per kernel and PCG
iteration for all cells, per-cell convergence flags on the device (a converged cell's blocks exit at
once), iterations captured in a HIP graph.

The solve itself runs the cells in `streams` groups (default 2), each a plan of its own on its own
stream and host thread. The matvec is bound by instruction issue and the FFT passes by HBM, so two
groups whose iterations drift out of phase share the GPU better than one launch sequence (8 cells of
64^3: 337 ms against 363 ms per solve on MI355X). Every cell's PCG is independent, so the grouping
does not change any cell's iterates.

The reference homogenizes one cell per `LinearHomogenization` (nlfiber/linear.py:99-236) and
shards tasks over processes (nlfiber/runner.py:55-75); this class is the batched form of
`C_full` for the same algorithm, with identical per-cell results.
"""
import ctypes
import threading

import numpy as np
import torch

from . import _lib
from .linear import Plan, _device, _es_host, _ptr, _stream, check_grid


class BatchedHomogenization:
    """C_full for a list of cells. `shapes`: list of (nx, ny, nz); `m` right-hand sides per cell
    (the unit strains 0..m-1). precision: "fp64" or "mixed"; inner_restart=True restarts each FP32 inner
    PCG of the mixed solver from p = z instead of carrying the search direction over."""

    def __init__(self, shapes, E=1.0, nu=0.3, emin=1e-3, p=3.0, tol=1e-8, maxit=4000, precision="mixed", m=6,
                 tune=False, warm_start=None, streams=2, warm_smoothing=(128, 900.0), inner_restart=False):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.E, self.nu, self.emin, self.p, self.tol, self.maxit = E, nu, emin, p, tol, maxit
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.m = m
        for sh in dict.fromkeys(self.shapes):
            check_grid(sh)
        self._dev = _device()
        self.inner_restart = bool(inner_restart)
        self._plan = Plan(self.shapes, E, nu, m_layout=m)
        self._plan.set_inner_restart(self.inner_restart)
        self._lib = _lib.load()
        self.nvox = [int(np.prod(s)) for s in self.shapes]
        self.off = np.concatenate([[0], np.cumsum(self.nvox)]).astype(np.int64)
        self.tot = int(self.off[-1])
        self.last_info = {}
        if isinstance(warm_start, str):
            from .learned import WarmStartNet
            warm_start = WarmStartNet.load(warm_start)
        self.warm_start = warm_start
        self.warm_smoothing = warm_smoothing
        self.streams = max(1, int(streams))
        self._groups = self._partition(min(self.streams, len(self.shapes)))
        self._subs = None  # (plan, stream) per group, made at the first solve
        if tune:
            self.autotune()

    def _partition(self, G):
        """Cells in G groups of about equal voxel count (largest first onto the lightest group), each
        group in cell order; None for a single group."""
        if G <= 1:
            return None
        load, groups = [0] * G, [[] for _ in range(G)]
        for i in sorted(range(len(self.shapes)), key=lambda i: -self.nvox[i]):
            g = int(np.argmin(load))
            groups[g].append(i)
            load[g] += self.nvox[i]
        return [sorted(g) for g in groups]

    def _sub_plans(self):
        """One plan and stream per group, tiled like this batch's plan for the same shapes."""
        if self._subs is None:
            self._subs = []
            for g in self._groups:
                shapes = [self.shapes[i] for i in g]
                off = np.concatenate([[0], np.cumsum([self.nvox[i] for i in g])]).astype(np.int64)
                sub = Plan(shapes, self.E, self.nu, m_layout=self.m)
                sub.set_inner_restart(self.inner_restart)
                self._subs.append((sub, torch.cuda.Stream(), g, off))
        distinct = list(dict.fromkeys(self.shapes))
        for plan, _, g, _ in self._subs:
            for j, sh in enumerate(dict.fromkeys(self.shapes[i] for i in g)):
                want = self._plan.get_tiling(distinct.index(sh))
                if plan.get_tiling(j) != want:
                    plan.set_tiling(j, *want)
                wb, _, wa = self._plan.get_fft_tiling(distinct.index(sh))
                hb, _, ha = plan.get_fft_tiling(j)
                if (hb, ha) != (wb, wa):
                    plan.set_fft_tiling(j, wb, wa)
        return self._subs

    def autotune(self, precision=None):
        """Tile every distinct shape with the autotuner and apply it: shapes that several cells of one
        stream group share take the batched tiling, the others the single-cell one. Mixed-precision batches also
        take the tuned FP32 pass B block shape of their cubic 16-128 shapes."""
        from . import autotune as at
        distinct = []
        for s in self.shapes:
            if s not in distinct:
                distinct.append(s)
        prec = precision or ("f64" if self.precision == "fp64" else "f32")
        groups = self._groups or [list(range(len(self.shapes)))]
        for i, s in enumerate(distinct):
            per_group = max(sum(1 for c in g if self.shapes[c] == s) for g in groups)
            tmp = Plan([s], self.E, self.nu)
            best, _ = at.tune(tmp, 0, m=self.m, precision=prec, batched=per_group >= 2)
            self._plan.set_tiling(i, *best)
            if prec == "f32":
                bb, _ = at.tune_fft(tmp, 0, m=self.m, batched=per_group >= 2)
                if bb is not None:
                    self._plan.set_fft_tiling(i, *bb)

    def pack(self, rhos):
        """Concatenate the cells' densities into one device array (cell-major)."""
        out = torch.empty(self.tot, dtype=torch.float64, device=self._dev)
        for i, r in enumerate(rhos):
            t = r if isinstance(r, torch.Tensor) else torch.as_tensor(np.asarray(r, dtype=np.float64))
            out[self.off[i]:self.off[i + 1]] = t.to(device=self._dev, dtype=torch.float64).reshape(-1)
        return out

    def field_views(self, buf):
        """Per-cell (m, nx, ny, nz, 3) views of a batched field buffer."""
        views = []
        for i, s in enumerate(self.shapes):
            a = int(self.off[i]) * 3 * self.m
            b = int(self.off[i + 1]) * 3 * self.m
            views.append(buf[a:b].view((self.m, 3) + s).permute(0, 2, 3, 4, 1))
        return views

    def stiffness_scale(self, rho_packed):
        s = torch.empty_like(rho_packed)
        _lib.check(self._lib.nlfh_stiffness_scale(self._plan.handle, _ptr(rho_packed), float(self.emin), float(self.p),
                                                  _ptr(s), None, _stream()))
        return s

    def warm_start_fields(self, rho_packed, s, max_batch=32):
        """Learned start for every cell: the network's prediction per group of equal cubic shapes
        (cells it cannot take start from zero), smoothed by `warm_smoothing` = (degree, hi / lo) steps of
        Chebyshev-Jacobi iteration (nlfh_cheb_smooth, hi the spectral bound of D^-1 K; None: no smoothing),
        then the per-cell energy-optimal combination (nlfh_ritz_start). Returns a packed field buffer."""
        if self.m != 6:
            raise ValueError("the learned start predicts all six unit-strain correctors (m = 6)")
        P = torch.zeros(self.tot * 3 * self.m, dtype=torch.float64, device=self._dev)
        groups = {}
        for i, sh in enumerate(self.shapes):
            if sh[0] == sh[1] == sh[2] and sh[0] % 4 == 0 and sh[0] ** 3 % 16 == 0:
                groups.setdefault(sh, []).append(i)
        for sh, idx in groups.items():
            n = sh[0]
            for k in range(0, len(idx), max_batch):
                part = idx[k:k + max_batch]
                r = torch.stack([rho_packed[self.off[i]:self.off[i + 1]].view(sh) for i in part])
                pred = self.warm_start.predict(r)  # (B, 6, n, n, n, 3)
                planar = pred.permute(0, 1, 5, 2, 3, 4).reshape(len(part), -1)
                for j, i in enumerate(part):
                    a = int(self.off[i]) * 3 * self.m
                    P[a:a + planar.shape[1]] = planar[j]
        Es_np, Es_p = _es_host(np.eye(6))
        if self.warm_smoothing:
            degree, ratio = self.warm_smoothing
            hi = float(self._lib.nlfh_jacobi_bound(float(self.E), float(self.nu)))
            _lib.check(self._lib.nlfh_cheb_smooth(self._plan.handle, _ptr(s), Es_p, self.m, _ptr(P), _ptr(P),
                                                  int(degree), hi / float(ratio), hi, None, _stream()))
        _lib.check(self._lib.nlfh_ritz_start(self._plan.handle, _ptr(s), Es_p, self.m, _ptr(P), _ptr(P), None,
                                             _stream()))
        return P

    def _solve_plan(self, plan, s, Es_p, u0, u, nc, stream):
        it = (ctypes.c_int * nc)()
        cv = (ctypes.c_int * nc)()
        rr = (ctypes.c_double * (nc * self.m))()
        prec = 0 if self.precision == "fp64" else 1
        _lib.check(self._lib.nlfh_solve(plan.handle, _ptr(s), Es_p, self.m, _ptr(u0), _ptr(u), float(self.tol),
                                        int(self.maxit), prec, it, cv, rr, stream))
        return dict(it=list(it), converged=[bool(c) for c in cv], relres=np.asarray(list(rr)).reshape(nc, self.m))

    def _solve(self, s, Es_p, u0):
        nc = len(self.shapes)
        u = torch.empty(self.tot * 3 * self.m, dtype=torch.float64, device=self._dev)
        if self._groups is None:
            return u, self._solve_plan(self._plan, s, Es_p, u0, u, nc, _stream())
        # inputs and outputs of every group are allocated on the caller's stream, which each group's
        # stream waits for before and which waits for every group after
        f = 3 * self.m
        cur = torch.cuda.current_stream()
        jobs = []
        for plan, st, g, off in self._sub_plans():
            s_g = torch.cat([s[self.off[i]:self.off[i + 1]] for i in g])
            u0_g = None if u0 is None else torch.cat([u0[self.off[i] * f:self.off[i + 1] * f] for i in g])
            u_g = torch.empty(int(off[-1]) * f, dtype=torch.float64, device=self._dev)
            st.wait_stream(cur)
            jobs.append([plan, st, g, off, s_g, u0_g, u_g, None, None])

        def work(job):
            plan, st, g, _, s_g, u0_g, u_g = job[:7]
            try:
                job[7] = self._solve_plan(plan, s_g, Es_p, u0_g, u_g, len(g), ctypes.c_void_p(st.cuda_stream))
            except BaseException as e:  # re-raised on the caller's thread
                job[8] = e

        threads = [threading.Thread(target=work, args=(j,)) for j in jobs[1:]]
        for t in threads:
            t.start()
        work(jobs[0])
        for t in threads:
            t.join()
        for j in jobs:
            cur.wait_stream(j[1])
        for j in jobs:
            if j[8] is not None:
                raise j[8]
        info = dict(it=[0] * nc, converged=[False] * nc, relres=np.zeros((nc, self.m)))
        for plan, st, g, off, s_g, u0_g, u_g, inf, _ in jobs:
            for k, i in enumerate(g):
                u[self.off[i] * f:self.off[i + 1] * f] = u_g[off[k] * f:off[k + 1] * f]
                info["it"][i] = inf["it"][k]
                info["converged"][i] = inf["converged"][k]
                info["relres"][i] = inf["relres"][k]
        return u, info

    def C_full(self, rhos, u0=None):
        """Returns (C (ncells, m, m) host float64, u (packed device buffer), iterations (ncells,)).
        `rhos` is a list of arrays or an already packed device array; u0 a packed buffer or None.
        With a warm_start network and no u0, the learned start is used; if any cell then fails to
        converge, the batch is solved again from zero (last_info["fallback"])."""
        rho = rhos if isinstance(rhos, torch.Tensor) and rhos.dim() == 1 else self.pack(rhos)
        s = self.stiffness_scale(rho)
        Es = np.eye(6)[: self.m]
        Es_np, Es_p = _es_host(Es)
        nc = len(self.shapes)
        warm = u0 is None and self.warm_start is not None
        if warm:
            u0 = self.warm_start_fields(rho, s)
        u, info = self._solve(s, Es_p, u0)
        info["warm_start"], info["fallback"] = warm, False
        if warm and (not all(info["converged"]) or not np.all(np.isfinite(info["relres"]))):
            it_warm = info["it"]
            u, info = self._solve(s, Es_p, None)
            info.update(warm_start=True, fallback=True, it_warm=it_warm)
        info["fallback_fp64"] = []
        if self.precision == "mixed" and not all(info["converged"]):
            self._fp64_fallback(rho, u, info)
        C = torch.empty((nc, self.m, self.m), dtype=torch.float64, device=self._dev)
        _lib.check(self._lib.nlfh_bilinear(self._plan.handle, _ptr(s), Es_p, _ptr(u), self.m, _ptr(C), None,
                                           _stream()))
        self.last_info = info
        return C.cpu(), u, np.asarray(info["it"])

    def _fp64_fallback(self, rho, u, info):
        """Cells the mixed-precision solve left unconverged (iteration cap or a refinement stall) are
        solved again by plain FP64 PCG from zero in a plan of their own; their correctors replace
        the mixed ones in `u` and their iteration counts add to the cell's."""
        bad = [i for i, c in enumerate(info["converged"]) if not c]
        sub = BatchedHomogenization([self.shapes[i] for i in bad], E=self.E, nu=self.nu, emin=self.emin, p=self.p,
                                    tol=self.tol, maxit=self.maxit, precision="fp64", m=self.m)
        rho_sub = torch.cat([rho[self.off[i]:self.off[i + 1]] for i in bad])
        _, u_sub, it_sub = sub.C_full(rho_sub)
        f = 3 * self.m
        for j, i in enumerate(bad):
            u[int(self.off[i]) * f:int(self.off[i + 1]) * f] = u_sub[int(sub.off[j]) * f:int(sub.off[j + 1]) * f]
            info["it"][i] += int(it_sub[j])
            info["converged"][i] = sub.last_info["converged"][j]
            info["relres"][i] = sub.last_info["relres"][j]
        info["fallback_fp64"] = bad

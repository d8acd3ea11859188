"""Train the learned warm start on correctors produced by the solver itself, then measure what it
gains over a zero start on held-out cells.

    python nonlinear-fiber-homogenization_amd/train_warmstart.py --n 64 --train 5000 --heldout 512 \
        --out weights.npz --report report.json

Cells are the seeded random cubic-symmetric cells of `cells.random_cell` (training seeds from
`--seed0`, held-out seeds from `--heldout-seed0`; the ranges must not overlap). Their densities
are generated by a pool of
CPU workers while the GPU solves; training targets are the six unit-strain correctors of the batched
mixed-precision solver at `--data-tol`. Evaluation solves every held-out cell at `--tol` (or at each of
`--tols`) from zero, from the learned start (the network's prediction, Chebyshev-Jacobi smoothing and
energy-optimal Ritz weights; inference and smoothing time included in the warm timing) and from the unsmoothed
Ritz start, and reports iterations, wall-clock solves/s and the largest C difference; for
the first `--start-diagnostics` cells it also reports how good a start the prediction is (relative
residual, which the stopping rule measures, and relative energy-norm error, per right-hand side).

The reference has no learned warm start: it restarts each solve from the previous corrector
(nlfiber/fiber.py:41-56) through the u0 argument (nlfiber/linear.py:212-216) that this uses.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np


def _cell(args):
    n, seed = args
    from nlfh_amd import cells
    rho, meta = cells.random_cell(n, seed)
    return seed, rho.astype(np.float64), meta


def _load_pkg():
    import importlib.util
    root = os.path.dirname(os.path.abspath(__file__))
    if "nlfh_amd" in sys.modules:
        return sys.modules["nlfh_amd"]
    spec = importlib.util.spec_from_file_location("nlfh_amd", os.path.join(root, "__init__.py"),
                                                  submodule_search_locations=[root])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["nlfh_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cell_stream(pool, n, seeds, chunk=8):
    for item in pool.imap(_cell, [(n, int(s)) for s in seeds], chunksize=chunk):
        yield item


def _log(*a):
    print(*a, flush=True)


def pack_prediction(u):
    """(B, 6, n, n, n, 3) -> packed planar field buffer [cell][rhs][comp][x][y][z]."""
    return u.permute(0, 1, 5, 2, 3, 4).contiguous().reshape(-1)


def to_target(u_packed, B, n):
    """packed planar (B, 6, 3, n, n, n) float64 -> channels-last (B, n, n, n, 18) float32, divided by n."""
    u = u_packed.view(B, 6, 3, n, n, n).permute(0, 3, 4, 5, 1, 2).reshape(B, n, n, n, 18)
    return (u / n).float()


def evaluate(net, held, n, bs, emin, tol):
    """Zero start against the learned start (prediction + Chebyshev-Jacobi smoothing + Ritz weights, the
    default of the warm_start option) and, for comparison, the unsmoothed Ritz start, on the held-out cells
    in batches of `bs`, mixed precision to relative tolerance `tol`. The learned solve's time includes the
    network inference, the smoothing and the Ritz step."""
    import torch
    from nlfh_amd import BatchedHomogenization

    def plans(B):
        kw = dict(emin=emin, tol=tol, precision="mixed")
        return (BatchedHomogenization([(n, n, n)] * B, **kw),
                BatchedHomogenization([(n, n, n)] * B, warm_start=net, **kw),
                BatchedHomogenization([(n, n, n)] * B, warm_start=net, warm_smoothing=None, **kw))

    cold_it, warm_it, raw_it, cold_t, warm_t, dC, fallbacks = [], [], [], 0.0, 0.0, [], 0
    bh, bw, bu = plans(bs)
    for k in range(0, len(held), bs):
        buf = [rho for _, rho in held[k:k + bs]]
        B = len(buf)
        if B != bs:
            bh, bw, bu = plans(B)
        rho_d = bh.pack(buf)
        torch.cuda.synchronize()
        t = time.time()
        C0, _, it0 = bh.C_full(rho_d)
        torch.cuda.synchronize()
        cold_t += time.time() - t
        t = time.time()
        C1, _, it1 = bw.C_full(rho_d)  # prediction + smoothing + Ritz weights + solve
        torch.cuda.synchronize()
        warm_t += time.time() - t
        fallbacks += int(bw.last_info["fallback"]) * B
        # the unsmoothed Ritz start, for comparison
        _, _, it2 = bu.C_full(rho_d)
        cold_it.extend(it0.tolist())
        warm_it.extend(it1.tolist())
        raw_it.extend(it2.tolist())
        c0, c1 = np.asarray(C0), np.asarray(C1)
        dC.append(float(np.max(np.abs(c1 - c0) / np.abs(c0).max(axis=(1, 2), keepdims=True))))
        _log(f"tol {tol:g} eval {k + B}/{len(held)}: mean it cold {np.mean(cold_it):.1f} learned "
             f"{np.mean(warm_it):.1f} learned unsmoothed {np.mean(raw_it):.1f}")
    return dict(cold_mean_it=float(np.mean(cold_it)), warm_mean_it=float(np.mean(warm_it)),
                unsmoothed_mean_it=float(np.mean(raw_it)), cold_solves_per_s=len(held) / cold_t,
                warm_solves_per_s=len(held) / warm_t, it_ratio=float(np.mean(warm_it) / np.mean(cold_it)),
                max_rel_dC=float(max(dC)), fallback_cells=fallbacks, cold_seconds=cold_t, warm_seconds=warm_t,
                warm_better_frac=float(np.mean(np.asarray(warm_it) < np.asarray(cold_it))))


def start_quality(net, rho, emin):
    """How good a start the prediction is, per right-hand side: the relative residual ||b - K x0|| / ||b|| (a
    zero start has 1, and the stopping rule measures this norm) and the relative energy-norm error
    sqrt(e.K e / u.K u), e = x0 - u with u the converged corrector (a zero start has 1), for the Ritz-combined
    start and for the raw prediction."""
    import torch
    from nlfh_amd import BatchedHomogenization, LinearHomogenization
    n = rho.shape[0]
    lh = LinearHomogenization((n, n, n), emin=emin, tol=1e-10, precision="fp64")
    bw = BatchedHomogenization([(n, n, n)], emin=emin, tol=1e-10, precision="fp64", warm_start=net)
    packed = bw.pack([rho])
    ritz = bw.field_views(bw.warm_start_fields(packed, bw.stiffness_scale(packed)))[0]
    raw = net.predict(rho.unsqueeze(0))[0]
    s = lh.stiffness_scale(rho)
    eye = np.eye(6)
    bnorm = lh.load(s, eye).reshape(6, -1).norm(dim=1)
    _, u, _ = lh.C_full(rho)
    uKu = (u.reshape(6, -1) * lh.apply_K(s, u).reshape(6, -1)).sum(1)
    out = {}
    for name, x in (("ritz", ritz), ("raw", raw)):
        res = lh.residual(s, eye, x).reshape(6, -1).norm(dim=1) / bnorm
        e = x - u
        eKe = (e.reshape(6, -1) * lh.apply_K(s, e).reshape(6, -1)).sum(1)
        out[name] = dict(residual=[round(float(v), 4) for v in res],
                         energy_error=[round(float(v), 4) for v in torch.sqrt(eKe / uKu)])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--train", type=int, default=5000)
    ap.add_argument("--heldout", type=int, default=512)
    ap.add_argument("--seed0", type=int, default=10_000, help="first training seed")
    ap.add_argument("--heldout-seed0", type=int, default=20_000, help="first held-out seed")
    ap.add_argument("--bs-solve", type=int, default=32)
    ap.add_argument("--bs-train", type=int, default=4)
    ap.add_argument("--epochs", type=float, default=3.0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--data-tol", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--emin", type=float, default=1e-4)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--init", default=None, help="start from these weights (npz)")
    ap.add_argument("--out", default=None, help="write the trained weights here (npz)")
    ap.add_argument("--report", default=None, help="write the evaluation here (json)")
    ap.add_argument("--skip-train", action="store_true")
    ap.add_argument("--tols", type=float, nargs="+", default=None,
                    help="evaluate the held-out cells at each of these tolerances (default: --tol)")
    ap.add_argument("--start-diagnostics", type=int, default=4,
                    help="held-out cells whose starting residual and energy error are reported")
    args = ap.parse_args(argv)

    # the package (no GPU call yet), then the worker pool: the workers only run numpy and must not
    # inherit an initialised GPU runtime
    _load_pkg()
    import torch
    ctx = mp.get_context("fork")
    pool = ctx.Pool(args.workers)
    from nlfh_amd.batch import BatchedHomogenization
    from nlfh_amd.learned import WarmStartNet

    n, bs = args.n, args.bs_solve
    dev = torch.device("cuda", 0)
    report = dict(n=n, train=args.train, heldout=args.heldout, data_tol=args.data_tol, tol=args.tol,
                  emin=args.emin, epochs=args.epochs, bs_train=args.bs_train)

    net = WarmStartNet.load(args.init) if args.init else WarmStartNet(seed=0)

    if not args.skip_train and args.train > 0:
        # ---- data: densities (bf16 net input) and solver correctors, resident in HBM --------------
        t0 = time.time()
        bh = BatchedHomogenization([(n, n, n)] * bs, emin=args.emin, tol=args.data_tol, precision="mixed")
        X = torch.empty((args.train, n, n, n, 16), dtype=torch.int16, device=dev)
        Y = torch.empty((args.train, n, n, n, 18), dtype=torch.float32, device=dev)
        seeds = np.arange(args.seed0, args.seed0 + args.train)
        buf, k, its = [], 0, []
        for seed, rho, _ in cell_stream(pool, n, seeds):
            buf.append(rho)
            if len(buf) == bs or k + len(buf) == args.train:
                B = len(buf)
                if B != bs:
                    bh = BatchedHomogenization([(n, n, n)] * B, emin=args.emin, tol=args.data_tol, precision="mixed")
                _, u, it = bh.C_full(buf)
                its.extend(it.tolist())
                Y[k:k + B] = to_target(u, B, n)
                X[k:k + B] = net.input(torch.as_tensor(np.stack(buf)))
                k += B
                buf = []
                if (k // bs) % 8 == 0 or k == args.train:
                    _log(f"data {k}/{args.train} cells, {time.time() - t0:.0f} s, mean it {np.mean(its):.0f}")
        report["data_seconds"] = time.time() - t0
        report["data_mean_it"] = float(np.mean(its))
        # ---- training: Adam with cosine decay over shuffled mini-batches -----------------------------
        t0 = time.time()
        steps = int(args.epochs * args.train / args.bs_train)
        g = torch.Generator().manual_seed(0)
        perm = torch.randperm(args.train, generator=g)
        pos, losses = 0, []
        for step in range(steps):
            if pos + args.bs_train > args.train:
                perm = torch.randperm(args.train, generator=g)
                pos = 0
            idx = perm[pos:pos + args.bs_train].to(dev)
            pos += args.bs_train
            lr = args.lr_min + 0.5 * (args.lr - args.lr_min) * (1 + np.cos(np.pi * step / max(steps, 1)))
            out = net.forward(X.index_select(0, idx), keep=True)
            L, dout = net.loss_and_grad(out, Y.index_select(0, idx))
            net.adam(net.backward(dout), lr=float(lr))
            losses.append(L)
            if step % 250 == 0 or step == steps - 1:
                _log(f"train step {step}/{steps} loss {np.mean(losses[-50:]):.3e} lr {lr:.1e} "
                     f"{time.time() - t0:.0f} s")
        torch.cuda.synchronize()
        report["train_seconds"] = time.time() - t0
        report["train_steps"] = steps
        report["loss_first"] = float(np.mean(losses[:20]))
        report["loss_last"] = float(np.mean(losses[-50:]))
        report["target_mean_square"] = float(Y[:256].pow(2).mean())
        del X, Y
        torch.cuda.empty_cache()
        if args.out:
            net.save(args.out, n=n, train=args.train, epochs=args.epochs, data_tol=args.data_tol)
            _log(f"weights written to {args.out}")

    # ---- evaluation on held-out cells: zero start vs learned start --------------------------------
    if args.heldout > 0:
        seeds = np.arange(args.heldout_seed0, args.heldout_seed0 + args.heldout)
        held = [(seed, rho) for seed, rho, _ in cell_stream(pool, n, seeds)]
        report["by_tol"] = {}
        for tol in (args.tols or [args.tol]):
            r = evaluate(net, held, n, bs, args.emin, tol)
            report["by_tol"][f"{tol:g}"] = r
            if tol == args.tol:
                report.update(r)
            _log(f"tol {tol:g}: " + json.dumps(r))
        report["start_quality"] = [start_quality(net, torch.tensor(rho, device="cuda"), args.emin)
                                   for _, rho in held[:args.start_diagnostics]]
        _log(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    pool.close()
    pool.join()
    return report


if __name__ == "__main__":
    main()

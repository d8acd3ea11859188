"""Benchmark: converged 6-load-case homogenizations per second on one MI355X.

Workloads (`--workload`):
  cell64     (default) seeded random cubic-symmetric 64^3 cells (Fourier modes |k| <= 2, 48-fold Oh
             averaging, Gaussian filter r = 1.5, beta in {2, 4, 8, 16}, volume in [0.2, 0.5],
             emin = 1e-4, p = 3), `--cells` of them homogenized together per step by the batched
             solver (one launch per kernel and PCG iteration for all cells).
  batch256   256 such cells of uneven sizes per step: 128 at 32^3, 96 at 48^3, 32 at 64^3.
  stress128  16 cells at 128^3 with emin = 1e-6, volume 0.15, beta = 16.
For every cell: the 6 unit-strain correctors by PCG to a true FP64 relative residual of 1e-8 and the
6x6 C. Inputs (rho) are resident in HBM before the timed region; the timed region is rho -> C on the
device (plan, symbol and tiling selection excluded, as the reference's __init__ is). The learned start
(the default of the cell64 workload: the packaged network's prediction, smoothed by Chebyshev-Jacobi steps
and combined by Ritz weights; inference and smoothing included in the timed region) starts the solve;
`--no-learned` starts from zero, `--learned WEIGHTS` picks other weights.

Prints one JSON line. `roofline` is for the matvec kernel of the timed solver precision over the
step's cells, timed live with HIP events on the stream it runs on (algorithmic bytes
N (2 * 3 * 6 * w + w_s) per application). `cpu_baseline` times the FP64 oracle on this host's
cores over a bounded number of PCG iterations of cell 0 and scales by the GPU's iteration count.
`plain_fp64_gpu` is the reference algorithm (plain FP64 PCG, one cell at a time) on the same cells.
`fp64_check` (after the timed region, mixed precision only) solves the timed cells again with the batched
plain FP64 PCG and reports the error of every C entry against it and that solver's rate; the run exits
with status 3 if any entry differs by more than 1e-6 of max|C|.
"""
import argparse
import ctypes
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged 6-load-case homogenizations/sec @64^3, rel.resid 1e-8; HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cell64", choices=["cell64", "batch256", "stress128", "fiber", "finite", "net"])
    ap.add_argument("--fiber-n-o", type=int, default=16, help="octant side of the fiber workload (cell 2 n_o)")
    ap.add_argument("--fiber-precision", default="fp64", choices=["fp64", "mixed"],
                    help="linear solver of the fiber workload (fp64: the reference's PCG)")
    ap.add_argument("--finite-n", type=int, default=64, help="cell side of the finite-strain workload")
    ap.add_argument("--finite-strain", type=float, default=0.05, help="uniaxial stretch - 1 of the finite workload")
    ap.add_argument("--cells", type=int, default=8, help="64^3 cells per step (cell64)")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--learned", default=None,
                    help="warm-start weights (npz); the cell64 workload uses the packaged weights unless --no-learned")
    ap.add_argument("--no-learned", action="store_true", help="zero start (the cell64 default is the learned start)")
    ap.add_argument("--smoothing", default="128,900",
                    help="DEGREE,RATIO of the learned start's Chebyshev-Jacobi smoothing (0,1: none)")
    ap.add_argument("--no-tune", action="store_true", help="default tiling instead of the autotuner")
    ap.add_argument("--tiling", default=None, help="TY,XC,ZPT matvec tiling for every shape (skips the autotuner)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 PCG comparison")
    ap.add_argument("--streams", type=int, default=2, help="stream groups of the batched solver")
    ap.add_argument("--inner-restart", action="store_true",
                    help="mixed precision: restart every FP32 inner PCG from p = z (default: carry the direction over)")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64 re-solve of the timed cells (by default C is compared against it, outside "
                         "the timed region, and the run fails above 1e-6 of max|C|)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the oracle's full FP64 solve of the 32^3 Schwarz cell (about 1-2 minutes)")
    ap.add_argument("--cpu-iters", type=int, default=6, help="PCG iterations timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workers", type=int, default=16,
                    help="host processes generating the cells (0: in this process, e.g. under a profiler)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="skip the timed solve: only the matvec roofline launches (for counter collection)")
    ap.add_argument("--precond-only", action="store_true",
                    help="skip the solve: time the preconditioner (FP64 and FP32, 6 RHS) on the workload's cells")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM-counter calibration)")
    return ap.parse_args()


def _cell(job):
    n, seed, kind = job
    import __graft_entry__ as ge
    pkg = ge.load_package()
    if kind == "stress":
        rho, meta = pkg.cells.stress_cell(seed, n)
    else:
        rho, meta = pkg.cells.random_cell(n, seed)
    return rho, meta


def workload(args):
    if args.workload == "cell64":
        jobs = [(64, s, "random") for s in range(args.cells)]
        return jobs, 1e-4, "cell64"
    if args.workload == "batch256":
        jobs = ([(32, s, "random") for s in range(128)] + [(48, s, "random") for s in range(128, 224)] +
                [(64, s, "random") for s in range(224, 256)])
        return jobs, 1e-4, "batch256"
    jobs = [(128, s, "stress") for s in range(16)]
    return jobs, 1e-6, "stress128"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "fiber":
        return fiber_bench(args)
    if args.workload == "finite":
        return finite_bench(args)
    if args.workload == "net":
        return net_bench(args)
    jobs, emin, wname = workload(args)
    tol = 1e-8
    if wname == "stress128":
        os.environ.setdefault("NLFH_PROGRESS", "1")  # solver progress on stderr during minutes-long steps
    # densities on host worker processes before the GPU runtime starts in this process
    import __graft_entry__ as ge
    pkg = ge.load_package()
    if args.workers > 0:
        with mp.get_context("fork").Pool(min(args.workers, len(jobs))) as pool:
            made = pool.map(_cell, jobs)
    else:
        made = [_cell(j) for j in jobs]
    rhos = [r for r, _ in made]
    metas = [m for _, m in made]
    shapes = [r.shape for r in rhos]

    import torch
    pkg.load_library()
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    net = None
    learned = args.learned
    if learned is None and args.workload == "cell64" and not args.no_learned:
        learned = os.path.join(ROOT, "nonlinear-fiber-homogenization_amd", "data", "warmstart64.npz")
    if learned and not args.no_learned:
        from nlfh_amd.learned import WarmStartNet
        net = WarmStartNet.load(learned)
    deg, ratio = (float(v) for v in args.smoothing.split(","))
    smoothing = (int(deg), ratio) if deg > 0 else None
    bh = pkg.BatchedHomogenization(shapes, emin=emin, tol=tol, maxit=40000, precision=args.precision,
                                   warm_start=net, streams=args.streams, warm_smoothing=smoothing,
                                   inner_restart=args.inner_restart)
    t_tune = time.perf_counter()
    if args.tiling:
        ty, xc, zpt = (int(v) for v in args.tiling.split(","))
        for i in range(len(set(shapes))):
            bh._plan.set_tiling(i, ty, xc, zpt)
    elif not args.no_tune:
        bh.autotune()
    t_tune = time.perf_counter() - t_tune
    tiling = {f"{sh[0]}x{sh[1]}x{sh[2]}": bh._plan.get_tiling(i)
              for i, sh in enumerate(dict.fromkeys(shapes))}
    fft_tiling = {f"{sh[0]}x{sh[1]}x{sh[2]}": list(bh._plan.get_fft_tiling(i))
                  for i, sh in enumerate(dict.fromkeys(shapes))}
    rho_d = bh.pack(rhos)

    if args.precond_only:
        print(json.dumps({"precond": precond_timing(pkg, bh, shapes)}), flush=True)
        return
    if args.kernels_only:
        for prec in ("mixed", "fp64"):
            print(json.dumps({"roofline_" + prec: matvec_roofline(pkg, bh, rho_d, prec)}), flush=True)
        copy_probe(args.copy_probe_mb)
        return

    fallback_fp64 = []

    def step():
        C, u, it = bh.C_full(rho_d)
        if not all(bh.last_info["converged"]):
            raise RuntimeError(f"cells did not converge: {bh.last_info['converged']}")
        fallback_fp64.append(list(bh.last_info.get("fallback_fp64", [])))
        return C, it, bh.last_info.get("fallback", False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its_all, fallbacks = [], 0
    for k in range(args.steps):
        C, its, fb = step()
        its_all.append(its.tolist())
        fallbacks += int(fb)
        print(f"step {k + 1}/{args.steps} done, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ncell = len(jobs)
    ms_step = 1e3 * dt / args.steps
    value = ncell * args.steps / dt

    roofline = matvec_roofline(pkg, bh, rho_d, args.precision)
    # solver level: the reference algorithm's bytes per PCG iteration (13 field passes of 6 right-hand
    # sides x 3 components, s once, the 3x3 symmetric symbol once) times the iterations every cell took,
    # over the step's wall time; fusion below 13 passes shows as time saved, not as fewer bytes
    w_it = 8 if args.precision == "fp64" else 4
    b_it = [int(np.prod(sh)) * (13 * 3 * 6 * w_it + w_it) + sh[0] * sh[1] * (sh[2] // 2 + 1) * 6 * w_it
            for sh in shapes]
    iters = its_all[-1] if its_all else []
    achieved_solve = sum(b * it for b, it in zip(b_it, iters)) / (ms_step * 1e-3) / 1e9
    solver_bw = {"achieved": round(achieved_solve, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved_solve / HBM_PEAK_GBS, 4), "bytes_per_iteration_all_cells": sum(b_it),
                 "iterations_summed": int(sum(iters)),
                 "basis": "13-pass PCG iteration bytes of the inner precision x PCG iterations per cell / wall time"}

    extra = {}
    if args.precision == "mixed" and not args.no_fp64_check:
        # after the timed region: every timed cell solved again by the batched plain FP64 PCG, C compared
        extra["fp64_check"] = fp64_check(pkg, shapes, rho_d, emin, tol, np.asarray(C), args.streams)
    zero_start = None
    if net is not None:
        # the same batched mixed solver from zero on the same cells (after the timed region), for the gain
        bz = pkg.BatchedHomogenization(shapes, emin=emin, tol=tol, maxit=40000, precision=args.precision,
                                       streams=args.streams, inner_restart=args.inner_restart)
        for i in range(len(set(shapes))):
            bz._plan.set_tiling(i, *bh._plan.get_tiling(i))
            fb, _, fa = bh._plan.get_fft_tiling(i)
            bz._plan.set_fft_tiling(i, fb, fa)
        bz.C_full(rho_d)
        torch.cuda.synchronize()
        tz = time.perf_counter()
        for _ in range(args.steps):
            _, _, itz = bz.C_full(rho_d)
        torch.cuda.synchronize()
        tz = time.perf_counter() - tz
        zero_start = {"value": round(ncell * args.steps / tz, 3), "unit": "homogenizations/s",
                      "pcg_iterations": itz.tolist(), "learned_speedup": round(value / (ncell * args.steps / tz), 4)}
    plain = None
    if not args.no_plain and args.workload == "cell64":
        plain = plain_fp64(pkg, rhos, emin, tol, dev)
        plain["speedup"] = round(value / plain["value"], 3)

    cpu = None
    if not args.no_cpu:
        # the oracle runs the plain FP64 algorithm: scale by that algorithm's iteration count
        it_fp64 = plain["pcg_iterations"][0] if plain else extra.get("fp64_check", {}).get("fp64_iterations",
                                                                                        its_all[-1])[0]
        cpu = cpu_baseline(rhos[0], emin, tol, it_fp64, args.cpu_iters)

    dtype = "f64" if args.precision == "fp64" else "f32-inner/f64-refinement"
    desc = {"cell64": f"{ncell} seeded random cubic-symmetric 64^3 cells per step (seeds 0-{ncell - 1})",
            "batch256": "256 seeded random cubic-symmetric cells per step: 128 x 32^3, 96 x 48^3, 32 x 64^3",
            "stress128": "16 seeded high-contrast cells at 128^3 (emin 1e-6, volume 0.15, beta 16)"}[wname]
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "homogenizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": f"{desc}, 6 unit strains, PCG to true rel. residual {tol:g}, emin {emin:g}, C 6x6, "
                        f"batched solver ({args.precision})" + (", learned warm start" if net else ""),
            "cells_per_step": ncell, "rhs": 6, "tol": tol, "precision": args.precision,
            "streams": args.streams, "inner_restart": bool(args.inner_restart),
            "tiling": {"source": "given" if args.tiling else ("default" if args.no_tune else "autotuned"),
                       "ty_xc_zpt": tiling,
                       "fft_pass_b_positions_default_pass_a_planes": fft_tiling},
            "tune_seconds": round(t_tune, 2),
            "learned_start": bool(net), "learned_start_fallback_steps": fallbacks,
            "learned_start_smoothing": list(smoothing) if (net and smoothing) else None,
            "fp64_fallback_cells": fallback_fp64[-1] if fallback_fp64 else [],
            "pcg_iterations": its_all[-1] if its_all else [],
            "cells": metas if ncell <= 16 else f"{ncell} cells (seeds in the workload description)",
        },
        "roofline": roofline,
        "solver_bandwidth": solver_bw,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": plain,
        "mixed_zero_start": zero_start,
    }
    res.update(extra)
    if args.cpu_full and not args.no_cpu:
        res["cpu_config0"] = cpu_config0(pkg)
    print(json.dumps(res), flush=True)
    chk = extra.get("fp64_check")
    if chk is not None and not (chk["max_rel_err_C"] <= 1e-6 and chk["fp64_converged"]):
        print(f"C differs from the FP64 solve by {chk['max_rel_err_C']:.3g} of max|C| (gate 1e-6)", file=sys.stderr)
        sys.exit(3)


def matvec_roofline(pkg, bh, rho_d, precision):
    """The matvec over all cells of the step (6 right-hand sides), in the precision of the inner
    solve, timed with HIP events over repeated launches on the current stream."""
    import torch
    lib = pkg.load_library()
    f32 = precision == "mixed"
    s = bh.stiffness_scale(rho_d)
    nf = bh.tot * 3 * 6
    dt = torch.float32 if f32 else torch.float64
    u = torch.randn(nf, dtype=dt, device=rho_d.device)
    out = torch.empty_like(u)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    args = (bh._plan.handle, ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(u.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), 6, 1 if f32 else 0, sp)

    def mv():
        pkg._lib.check(lib.nlfh_apply_K(*args))

    # 20 untimed launches first and up to 200 timed ones (about 20 ms at 64^3): the launch's duration at the clock
    # the GPU holds under load; a short window straight after idle catches the clock ramp (95-103 us against
    # 88-90 us for the same 8 x 64^3 launch)
    for _ in range(20):
        mv()
    reps = max(20, min(200, int(4e8 // max(bh.tot, 1))))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        mv()
    e1.record(stream)
    torch.cuda.synchronize()
    t_mv = e0.elapsed_time(e1) / reps * 1e-3
    w = 4 if f32 else 8
    # the field values in the operator's precision; s is read in FP64 by both (the FP32 operator rounds it
    # as it loads it)
    bytes_mv = bh.tot * (2 * 3 * 6 * w + 8)
    achieved = bytes_mv / t_mv / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "matvec_traffic.json")
    kname = f"matvec_kernel<{'float' if f32 else 'double'}>"
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        tiling = ",".join(str(v) for v in bh._plan.get_tiling(0))
        key = f"{kname} voxels={bh.tot} tiling={tiling}"  # counters of the same launch shape only
        if key in tr:
            traffic = tr[key]["hbm_bytes"]
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"{kname} (K(s) u, 6 RHS, {bh.tot} voxels = {len(bh.shapes)} cells)",
            "algorithmic_bytes": bytes_mv, "avg_us": round(t_mv * 1e6, 2)}


def precond_timing(pkg, bh, shapes, reps=5):
    """nlfh_precond on all cells of the plan, 6 right-hand sides, FP64 then FP32, random residuals: HIP events over
    `reps` applications after a warm-up; algorithmic bytes N 2 3 m w + Nh 6 w per distinct shape."""
    import torch
    lib = pkg.load_library()
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    out = {}
    for dt, tdt, w in ((0, torch.float64, 8), (1, torch.float32, 4)):
        r = torch.randn(bh.tot * 18, dtype=tdt, device="cuda")
        z = torch.empty_like(r)
        args = (bh._plan.handle, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(z.data_ptr()), 6, dt, sp)
        pkg._lib.check(lib.nlfh_precond(*args))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            pkg._lib.check(lib.nlfh_precond(*args))
        e1.record(st)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps * 1e-3
        b = bh.tot * 2 * 3 * 6 * w + sum(sh[0] * sh[1] * (sh[2] // 2 + 1) * 6 * w for sh in set(shapes))
        out["fp64" if dt == 0 else "fp32"] = {"ms": round(t * 1e3, 3), "algorithmic_bytes": b,
                                              "GBs": round(b / t / 1e9, 1)}
        del r, z
    return out


def fiber_bench(args):
    """Fiber.value_jac (density map + warm-started 2-RHS PCG + invariant gradient + density-map VJP) along a
    path of pre-densities phi_k = phi_0 + 0.02 k d, as a projection loop evaluates them."""
    import torch

    import __graft_entry__ as ge
    pkg = ge.load_package()
    pkg.load_library()
    torch.cuda.set_device(0)
    from nlfh_amd.fiber import Fiber, random_start
    n_o, beta, v, r = args.fiber_n_o, 8.0, 0.3, 1.5
    phi0 = random_start(n_o, seed=0)
    d = random_start(n_o, seed=1)
    fib = Fiber(n_o, v=v, r=r, precision=args.fiber_precision)  # the reference's defaults: emin 1e-3, tol 1e-9
    phis = [torch.tensor(phi0 + 0.02 * k * d, device="cuda") for k in range(args.warmup + args.steps)]
    for k in range(args.warmup):
        fib.value_jac(phis[k], beta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for k in range(args.warmup, args.warmup + args.steps):
        c, J = fib.value_jac(phis[k], beta)
        its.append(fib.lh.last_info["it"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = args.steps / dt
    # roofline: the FP64 matvec on the fiber's 2 right-hand sides
    lib = pkg.load_library()
    s = fib.lh.stiffness_scale(fib.rho(phis[-1], beta))
    n = 2 * n_o
    u = torch.randn((2, 3, n, n, n), dtype=torch.float64, device="cuda")
    out = torch.empty_like(u)
    st = torch.cuda.current_stream()
    a = (fib.lh._plan.handle, ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(u.data_ptr()),
         ctypes.c_void_p(out.data_ptr()), 2, 0, ctypes.c_void_p(st.cuda_stream))
    for _ in range(3):
        pkg._lib.check(lib.nlfh_apply_K(*a))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(50):
        pkg._lib.check(lib.nlfh_apply_K(*a))
    e1.record(st)
    torch.cuda.synchronize()
    t_mv = e0.elapsed_time(e1) / 50 * 1e-3
    bytes_mv = n ** 3 * (2 * 3 * 2 * 8 + 8)
    roof = {"bound": "hbm", "achieved": round(bytes_mv / t_mv / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bytes_mv / t_mv / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"matvec_kernel<double> (K(s) u, 2 RHS, {n}^3)", "algorithmic_bytes": bytes_mv,
            "avg_us": round(t_mv * 1e6, 2)}
    cpu = None
    if not args.no_cpu:
        cpu = fiber_cpu(phis[args.warmup].cpu().numpy(), beta, v, r)
    print(json.dumps({
        "metric": "Fiber.value_jac evaluations/sec (density map + 2-RHS PCG + Jacobian rows)",
        "value": round(value, 3), "unit": "value_jac/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.fiber_precision == "fp64" else "f32-inner/f64-refinement",
        "data": "synthetic",
        "config": {"solver_precision": args.fiber_precision,
                   "workload": f"Fiber(n_o={n_o}) on a {n}^3 cell, beta {beta}, v {v}, r {r}, emin 1e-3, tol 1e-9, "
                               f"warm-started along phi_k = phi_0 + 0.02 k d", "pcg_iterations": its},
        "roofline": roof, "cpu_baseline": cpu}), flush=True)


def net_flops_per_voxel():
    """Multiply-add flops per input voxel of one forward pass of the warm-start network (2 * 27 * cin * coutp per
    output voxel of every convolution, padded channel counts, at its level's resolution)."""
    from nlfh_amd.learned import LAYERS
    return sum(2 * 27 * (ca + cb) * coutp / 8 ** lvl for _, ca, _, cb, _, coutp, _, lvl in LAYERS)


def net_bench(args):
    """The learned start's network on 64^3 cells: inference of 32 cells (the learned start's batch) and training
    steps of 4 cells (forward, MSE loss, backward, Adam), synthetic densities and targets, HIP events around
    repeated calls on the current stream. The roofline is MFMA-bound: algorithmic flops (forward 1x, training
    3x: forward, data and weight gradients) over time against the 2.5 PFLOP/s dense BF16 peak."""
    import torch

    import __graft_entry__ as ge
    pkg = ge.load_package()
    pkg.load_library()
    torch.cuda.set_device(0)
    from nlfh_amd.learned import WarmStartNet
    net = WarmStartNet.load(os.path.join(ROOT, "nonlinear-fiber-homogenization_amd", "data", "warmstart64.npz"))
    n, bi, bt = 64, 32, 4
    g = torch.Generator(device="cuda").manual_seed(0)
    rho = torch.rand((bi, n, n, n), dtype=torch.float64, device="cuda", generator=g)
    X = net.input(rho[:bt])
    Y = 0.01 * torch.randn((bt, n, n, n, 18), device="cuda", generator=g)

    def infer():
        net.predict(rho)

    def train():
        out = net.forward(X, keep=True)
        _, dout = net.loss_and_grad(out, Y)
        net.adam(net.backward(dout), lr=1e-6)

    st = torch.cuda.current_stream()
    res = {}
    for name, f, reps in (("inference", infer, max(args.steps, 5)), ("training_step", train, max(args.steps, 5))):
        for _ in range(max(args.warmup, 1)):
            f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            f()
        e1.record(st)
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / reps * 1e-3
    fv = net_flops_per_voxel()
    f_inf = fv * n ** 3 * bi
    f_tr = 3 * fv * n ** 3 * bt
    tf_inf, tf_tr = f_inf / res["inference"] / 1e12, f_tr / res["training_step"] / 1e12
    print(json.dumps({
        "metric": "warm-start network: 64^3 cells predicted per second (inference)", "value": round(bi / res["inference"], 1),
        "unit": "cells/s", "n_gpus": 1, "steps": max(args.steps, 5), "warmup": max(args.warmup, 1),
        "ms_per_step": round(1e3 * res["inference"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": f"inference of {bi} cells and training steps of {bt} cells at 64^3 (3 levels, widths "
                               f"16/32/64, 18 outputs)", "flops_per_voxel_forward": fv,
                   "training_step_ms": round(1e3 * res["training_step"], 3),
                   "training_tflops": round(tf_tr, 1)},
        "roofline": {"bound": "mfma", "achieved": round(tf_inf, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tf_inf / 2500.0, 4), "traffic": None,
                     "kernel": "conv3d_lds_kernel (all convolutions of one inference, 32 x 64^3)"},
        "cpu_baseline": None}), flush=True)


FS_FLOP_PER_VOXEL = 6016  # Hessian-vector product, neo-Hookean with void interpolation (8 Gauss points x 752)
FS_HVP_BYTES_PER_VOXEL = 80  # read w, v (2 x 24), rho (8); write the product (24)


def finite_bench(args):
    """Finite-strain uniaxial-stress solve (Mixed with the lateral stretches free, neo-Hookean with the void
    interpolation, the reference's defaults: rtol 1e-9, cg_tol 1e-6, Eisenstat-Walker forcing) of one n^3 cell,
    from a zero start each step."""
    import torch

    import __graft_entry__ as ge
    pkg = ge.load_package()
    pkg.load_library()
    torch.cuda.set_device(0)
    from nlfh_amd import mech
    n, e = args.finite_n, args.finite_strain
    rho, _ = pkg.cells.random_cell(n, 0)
    cell = mech.Cell(rho)
    mix = mech.Mixed(cell, free=[(1, 1), (2, 2)])
    Fb = np.diag([1.0 + e, 1.0, 1.0])
    logs = []
    for _ in range(args.warmup):
        mix.solve(Fb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w, F, log = mix.solve(Fb)
        logs.append(log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the Hessian-vector product kernels (element pass + node gather) at the final state, HIP events on
    # the stream they run on
    # (the direction's macro part and the output on the device beforehand, as in the Newton-CG loop, so that the
    # timed calls hold the three kernels only)
    st = torch.cuda.current_stream()
    v = torch.randn(cell.N * 3, dtype=torch.float64, device="cuda")
    dF = torch.as_tensor(np.diag([0.0, 1.0, 1.0]).ravel(), dtype=torch.float64, device="cuda")
    wd = torch.as_tensor(w, dtype=torch.float64, device="cuda").reshape(-1).contiguous()
    hv = torch.empty_like(wd)
    for _ in range(3):
        cell._eval(wd, F, 2, v, dF, out=hv)
    reps = 20
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        cell._eval(wd, F, 2, v, dF, out=hv)
    e1.record(st)
    torch.cuda.synchronize()
    t_h = e0.elapsed_time(e1) / reps * 1e-3
    flops = FS_FLOP_PER_VOXEL * cell.N
    traffic = None
    prof = os.path.join(ROOT, "profiles", "finite_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        key = f"hvp voxels={cell.N}"  # counters of the same launch shape only
        if key in tr:
            traffic = tr[key]["hbm_bytes"]
    roof = {"bound": "fp64-valu", "achieved": round(flops / t_h / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
            "frac": round(flops / t_h / 1e12 / 78.6, 4), "traffic": traffic,
            "kernel": f"fs_element_kernel<2> + fs_gather_kernel + fs_reduce_kernel (one Hessian-vector product, "
                      f"{n}^3)", "flops_per_voxel": FS_FLOP_PER_VOXEL, "avg_us": round(t_h * 1e6, 2),
            "hbm_algorithmic_GBs": round(FS_HVP_BYTES_PER_VOXEL * cell.N / t_h / 1e9, 1)}
    cpu = None
    if not args.no_cpu:
        cpu = finite_cpu(rho, Fb, logs[0], n)
    print(json.dumps({
        "metric": "finite-strain uniaxial-stress solves/sec (Newton-CG to rtol 1e-9, neo-Hookean, void interpolation)",
        "value": round(args.steps / dt, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Mixed(Cell(random cell {n}^3 seed 0), free=[(1,1),(2,2)]).solve(diag(1+{e},1,1))",
                   "newton": logs[0]["newton"], "cg": logs[0]["cg"], "converged": logs[0]["converged"],
                   "Pbar_xx": float(cell.mean_stress(w, F)[0, 0]), "F_lateral": [float(F[1, 1]), float(F[2, 2])]},
        "roofline": roof, "cpu_baseline": cpu}), flush=True)


def finite_cpu(rho, Fb, log, n):
    """The oracle's operators for the same solve, timed on a bounded sample and scaled: per Newton step one
    gradient, one tangent, the line-search energies; per CG iteration one stored-tangent product and one FFT
    preconditioner application (on the same cell), times the GPU solve's Newton and CG counts."""
    from oracle import linear_ref as O
    from oracle import mech_ref as MR
    cores = os.cpu_count() or 1
    O.set_fft_workers(cores)
    ref = MR.CellRef(rho)
    lin = O.LinearRef(rho.shape)
    w = np.zeros(rho.shape + (3,))
    t0 = time.perf_counter()
    ref.stress(w, Fb)
    t1 = time.perf_counter()
    ref.energy(w, Fb)
    t2 = time.perf_counter()
    A = ref.tangent(w, Fb)
    t3 = time.perf_counter()
    v = np.random.default_rng(0).standard_normal(w.shape)
    k = 3
    for _ in range(k):
        hw, _ = ref.hvp(A, v, np.zeros((3, 3)))
        lin.precond(hw[None])
    t4 = time.perf_counter()
    t_g, t_e, t_a, t_cg = t1 - t0, t2 - t1, t3 - t2, (t4 - t3) / k
    # per Newton step: gradient, tangent, 2 macro-diagonal products, energy + one line-search energy
    total = log["newton"] * (t_g + t_a + 2 * t_cg + 2 * t_e) + log["cg"] * t_cg
    t_newton = t_g + t_e + t_a
    return {"value": round(1.0 / total, 6), "unit": "solves/s", "cores": cores, "cpu": cpu_model(), "kind": "port",
            "sample": f"numpy oracle on the {n}^3 cell: one gradient + energy + tangent ({t_newton:.2f} s), {k} "
                      f"tangent products + FFT preconditioner ({t_cg:.2f} s each); scaled to the GPU solve's "
                      f"{log['newton']} Newton steps and {log['cg']} CG iterations (extrapolated)",
            "seconds_sampled": round(t4 - t0, 2)}


def fiber_cpu(phi, beta, v, r):
    """The oracle chain for one cold value_jac on the host cores."""
    from oracle import density_ref as DR
    from oracle import linear_ref as O
    cores = os.cpu_count() or 1
    O.set_fft_workers(cores)
    t0 = time.perf_counter()
    rho = DR.rho(phi, beta, r, v)
    ref = O.LinearRef(rho.shape, emin=1e-3, tol=1e-9)
    c, u2, it = ref.invariants(rho)
    Jrho = ref.invariants_grad(rho, u2)
    DR.rho_vjp(phi, beta, r, v, Jrho)
    t = time.perf_counter() - t0
    return {"value": round(1.0 / t, 5), "unit": "value_jac/s", "cores": cores, "cpu": cpu_model(), "kind": "port",
            "sample": f"numpy oracle: one cold value_jac ({it} PCG iterations, 2 RHS) on the host cores",
            "seconds": round(t, 2)}


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB: a kernel with known HBM traffic (read + write of that
    size) whose FETCH_SIZE / WRITE_SIZE counters calibrate the counters of the other kernels."""
    if mib <= 0:
        return
    import torch
    src = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": src.numel() * 4}), flush=True)


def plain_fp64(pkg, rhos, emin, tol, dev, steps=1):
    """The reference algorithm on the GPU: plain FP64 PCG (default tiling), one cell at a time."""
    import torch
    shapes = sorted(set(r.shape for r in rhos))
    lhs = {sh: pkg.LinearHomogenization(sh, emin=emin, tol=tol, maxit=40000, precision="fp64") for sh in shapes}
    rd = [torch.tensor(r, device=dev) for r in rhos]
    lhs[rhos[0].shape].C_full(rd[0])  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        for r in rd:
            lh = lhs[tuple(r.shape)]
            _, _, it = lh.C_full(r)
            its.append(it)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": round(len(rd) * steps / dt, 3), "unit": "homogenizations/s", "pcg_iterations": its[:len(rd)]}


def fp64_check(pkg, shapes, rho_d, emin, tol, C, streams):
    """Every cell solved again by the batched plain FP64 PCG (the reference's algorithm and stopping rule,
    same stream groups as the timed solver), one timed solve after a warm-up. Reports max over cells and
    entries of |dC_ij| / max|C| (the reference's normalisation, tests/test_core.py:124-133), the per-entry
    relative error |dC_ij| / |C_ij| over the entries above 1e-3 max|C| (entries that vanish by symmetry
    have no meaningful relative error), and the batched FP64 rate."""
    import torch
    bh = pkg.BatchedHomogenization(shapes, emin=emin, tol=tol, maxit=40000, precision="fp64", streams=streams)
    bh.autotune()
    warm = sum(int(np.prod(sh)) for sh in shapes) <= (1 << 23)
    if warm:
        bh.C_full(rho_d)  # warm-up (graphs, tiling); skipped for the minutes-long 128^3 batch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C64, _, it = bh.C_full(rho_d)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    C64 = np.asarray(C64)
    scale = np.abs(C64).max(axis=(1, 2), keepdims=True)
    err = np.abs(C - C64) / scale
    big = np.abs(C64) > 1e-3 * scale
    rel = np.where(big, np.abs(C - C64) / np.where(big, np.abs(C64), 1.0), 0.0)
    return {"max_rel_err_C": float(err.max()), "per_cell_max_rel_err_C": [float(e.max()) for e in err],
            "max_entry_rel_err_C": float(rel.max()), "fp64_iterations": it.tolist(),
            "fp64_converged": all(bh.last_info["converged"]),
            "batched_fp64_gpu": {"value": round(len(shapes) / dt, 3), "unit": "homogenizations/s",
                                 "note": "the same cells by the batched plain FP64 PCG (reference algorithm), "
                                         f"{streams} stream groups, one timed solve"
                                         + (" after a warm-up" if warm else " (no warm-up)")}}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_config0(pkg):
    """The oracle's full FP64 solve (PCG to 1e-8, C assembly) of the 32^3 Schwarz P cell (phi = -(cos X +
    cos Y + cos Z), r 1.5, beta 16, volume 0.3, emin 1e-4), timed whole on the host cores."""
    from oracle import linear_ref as O
    cores = os.cpu_count() or 1
    O.set_fft_workers(cores)
    rho = pkg.cells.schwarz_cell(32)
    t0 = time.perf_counter()
    ref = O.LinearRef(rho.shape, emin=1e-4, tol=1e-8)
    C, _, it = ref.C_full(rho)
    t = time.perf_counter() - t0
    return {"value": round(1.0 / t, 6), "unit": "homogenizations/s", "cores": cores, "cpu": cpu_model(),
            "kind": "port", "sample": f"numpy FP64 oracle, the whole 32^3 Schwarz cell solve ({it} PCG iterations) "
                                      "and C, not extrapolated", "seconds": round(t, 2), "iterations": int(it)}


def cpu_baseline(rho, emin, tol, it0, iters):
    """Oracle (numpy FP64, FFTs on all host cores) on cell 0: setup + `iters` PCG iterations + C,
    scaled to the GPU's iteration count for that cell."""
    from oracle import linear_ref as O

    cores = os.cpu_count() or 1
    O.set_fft_workers(cores)
    t0 = time.perf_counter()
    lr = O.LinearRef(rho.shape, emin=emin, tol=tol)
    s = lr.stiffness_scale(rho)
    Es = np.eye(6)
    t1 = time.perf_counter()
    u, it = lr.solve(s, Es, maxit=iters)
    t2 = time.perf_counter()
    lr.bilinear(s, Es, u)
    t3 = time.perf_counter()
    per_it = (t2 - t1) / max(iters, 1)
    t_cell = (t1 - t0) + per_it * it0 + (t3 - t2)
    return {"value": round(1.0 / t_cell, 6), "unit": "homogenizations/s", "cores": cores, "cpu": cpu_model(),
            "kind": "port",
            "sample": f"numpy FP64 oracle on cell 0 ({rho.shape[0]}^3): setup + {iters} timed PCG iterations "
                      f"({per_it:.3f} s/iteration) + C assembly, scaled to the {it0} iterations the GPU "
                      f"solve of that cell took",
            "seconds_per_homogenization": round(t_cell, 2)}


if __name__ == "__main__":
    main()

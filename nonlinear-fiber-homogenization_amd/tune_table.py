"""Measure the matvec tiling table shipped with the package (data/tiling_gfx950.json): for every
shape and precision, the autotuner's sweep (no cache) repeated `--repeat` times; the tiling with the
lowest median time wins. `--batched` measures the entries of batched plans (a plan of
autotune.batch_cells(shape) cells per sweep). `--fft` measures the FP32 pass B block shape of the
preconditioner instead (autotune.tune_fft; keys "...|fftb"). Entries are merged into the existing table.

    python nonlinear-fiber-homogenization_amd/tune_table.py --shapes 32 48 64 128
    python nonlinear-fiber-homogenization_amd/tune_table.py --shapes 32 48 64 --precisions f32 --batched
    python nonlinear-fiber-homogenization_amd/tune_table.py --shapes 16 32 48 64 128 --fft [--batched]
"""
import argparse
import json
import os
import sys

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, nargs="+", default=[16, 32, 48, 64, 128])
    ap.add_argument("--precisions", nargs="+", default=["f32", "f64"])
    ap.add_argument("--m", type=int, default=6)
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--fft", action="store_true", help="the FP32 pass B block shape instead of the matvec tiling")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                   "tiling_gfx950.json"))
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import __graft_entry__ as ge
    pkg = ge.load_package()
    from nlfh_amd import autotune as at
    from nlfh_amd.linear import Plan
    try:
        with open(args.out) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    for n in args.shapes:
        shape = (n, n, n)
        if args.fft:
            if args.batched and at.batch_cells(shape) <= 1:
                continue  # one cell fills the GPU: the single-cell entry serves batches too
            bests, pb, pa = [], {}, {}
            for _ in range(args.repeat):
                plan = Plan([shape], 1.0, 0.3)
                best, tab = at.tune_fft(plan, 0, m=args.m, reps=10, use_cache=False, batched=args.batched)
                bests.append(tuple(best))
                for (b, q), t in tab[:-2]:
                    pb.setdefault(b, []).append(t)
                for (b, q), t in tab[-2:]:
                    pa.setdefault(q, []).append(t)
            best = max(set(bests), key=bests.count)
            key = at._key(shape, "f32", args.m, args.batched, "fftb")
            table[key] = {"best": list(best),
                          "pass_b_us": {str(b): round(float(np.median(v)), 2) for b, v in sorted(pb.items())},
                          "pass_a_us_per_iteration": {str(q): round(float(np.median(v)), 2) for q, v in sorted(pa.items())}}
            if args.batched:
                table[key]["cells"] = at.batch_cells(shape)
            print(key, best, table[key]["pass_b_us"], table[key]["pass_a_us_per_iteration"], flush=True)
            continue
        for prec in args.precisions:
            runs = {}
            for _ in range(args.repeat):
                plan = Plan([shape], 1.0, 0.3)
                _, tab = at.tune(plan, 0, m=args.m, precision=prec, reps=10, use_cache=False, batched=args.batched)
                for cand, t in tab:
                    runs.setdefault(cand, []).append(t)
            med = sorted((float(np.median(v)), c) for c, v in runs.items())
            best_t, best = med[0]
            key = at._key(shape, prec, args.m, args.batched)
            table[key] = {"best": list(best), "us": round(best_t, 2), "runner_up": [list(c) for _, c in med[1:4]]}
            if args.batched:
                table[key]["cells"] = at.batch_cells(shape)
            print(key, best, f"{best_t:.1f} us", flush=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()

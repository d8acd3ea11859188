"""Host-side logic that needs no GPU: the packaged gfx950 tiling table and the batched-tuning keys."""
import ast
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "nonlinear-fiber-homogenization_amd", "data", "tiling_gfx950.json")


def _entries():
    with open(TABLE) as f:
        table = json.load(f)
    for key, e in table.items():
        parts = key.split("|")
        shape = ast.literal_eval(parts[0])
        kind = parts[3] if len(parts) > 3 and parts[3] != "batch" else ""
        yield key, tuple(shape), parts[1], int(parts[2]), parts[-1] == "batch", kind, e


def test_packaged_tilings_are_valid_candidates(pkg):
    """Every tiling the package ships is one the autotuner could have picked for that shape (tile rows,
    x-chunk and z values per lane that fit the 512-thread block and the LDS of the FP64 kernels, no three z values per
    lane for FP64 entries; for the FP32
    pass B entries, the three positions per block of autotune.fft_candidates around the default)."""
    from nlfh_amd import autotune as at
    n = nb = 0
    for key, shape, prec, m, batched, kind, e in _entries():
        assert prec in ("f32", "f64") and m == 6, key
        if kind == "fftb":
            assert prec == "f32" and shape[0] == shape[1] == shape[2] and shape[0] in (16, 32, 48, 64, 128), key
            tried = sorted(int(b) for b in e["pass_b_us"])
            d = tried[1]
            assert d in (16, 32, 64) and tried == sorted({c[0] for c in at.fft_candidates(d)}), key
            assert tuple(e["best"]) in at.fft_candidates(d), key
            assert sorted(int(q) for q in e["pass_a_us_per_iteration"]) == [1, 2], key
            nb += 1
            continue
        assert kind == "", key
        assert tuple(e["best"]) in at.candidates(shape, 8, prec), key
        for ru in e.get("runner_up", []):
            assert tuple(ru) in at.candidates(shape, 8, prec), key
        n += 1
    assert n >= 10 and nb >= 9


def test_batched_keys_only_for_shapes_that_batch(pkg):
    """Batched entries exist only for shapes whose representative batch has several cells (a 128^3 cell fills
    the GPU alone and keeps its single-cell entry), and the key format is the autotuner's."""
    from nlfh_amd import autotune as at
    for key, shape, prec, m, batched, kind, e in _entries():
        assert key == at._key(shape, prec, m, batched, kind)
        if batched:
            assert at.batch_cells(shape) > 1, key
            assert e.get("cells") in (None, at.batch_cells(shape)), key
    assert at.batch_cells((128, 128, 128)) == 1
    assert at.batch_cells((64, 64, 64)) == 8


def test_bench_refuses_more_than_one_gpu():
    """The bench runs on one GPU: --gpus above 1 exits with status 2 and a JSON line that says so, before any
    GPU work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_grid_limits_raise_value_errors(pkg):
    """Cells the FFT preconditioner or the operator cannot take are refused with a ValueError before any GPU
    work (the reference's jnp.fft takes any grid; this library takes sides that are products of 2, 3 and 5,
    nx, ny <= 512, nz <= 256 and nz split over at most 64 lanes of 1-4 values)."""
    import pytest
    from nlfh_amd.linear import check_grid
    for ok in ((64, 64, 64), (48, 48, 48), (30, 45, 75), (512, 512, 256), (8, 8, 1), (5, 9, 27)):
        check_grid(ok)
    for bad, what in (((14, 14, 14), "prime factor"), ((7, 8, 8), "prime factor"), ((1024, 8, 8), "outside"),
                      ((8, 8, 512), "outside"), ((8, 8, 250), "lanes"), ((8, 8, 0), "outside")):
        with pytest.raises(ValueError, match=what):
            check_grid(bad)
    # the classes check before touching the device (no HIP device here)
    with pytest.raises(ValueError, match="prime factor"):
        pkg.LinearHomogenization((7, 7, 7))
    with pytest.raises(ValueError, match="prime factor"):
        pkg.BatchedHomogenization([(8, 8, 8), (14, 14, 14)])


def test_dtype_other_than_float64_is_refused(pkg):
    """The reference's dtype argument (nlfiber/linear.py:102) sets its operators' precision; this class computes
    in float64 and refuses other dtypes instead of ignoring them."""
    import numpy as np
    import pytest
    import torch
    for bad in (torch.float32, np.float32, "float32", torch.bfloat16):
        with pytest.raises(ValueError, match="not supported"):
            pkg.LinearHomogenization((8, 8, 8), dtype=bad)
    from nlfh_amd.linear import _check_dtype
    for ok in (torch.float64, np.float64, "float64", np.dtype("float64")):
        _check_dtype(ok)

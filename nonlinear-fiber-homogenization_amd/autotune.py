"""Autotuner for the launch shapes of each grid shape of a plan: the matvec tiling (ty node rows per tile, xc
x-planes per chunk, zpt z values per lane) and, for the FP32 preconditioner of cubic 16-128 cells, the block shapes
of passes B and A ((ky, kz) positions per block of pass B: the default, half or double of it; x-planes per block of
the spectral pass A: 1 or 2; `tune_fft`).

Candidates are every valid combination from small sets; each is timed on the device with HIP events
over repeated launches of the operator K(s) u on m right-hand sides (the same kernel family the
solver launches every iteration) and the fastest is kept in the plan. Results are cached per
(device name, shape, precision, m) in a JSON file (NLFH_TUNE_CACHE, default
~/.cache/nlfh/tiling.json) so later plans skip the sweep. On gfx950 a table measured with
tune_table.py and shipped in data/tiling_gfx950.json is consulted before sweeping.

A launch over many cells of one shape fills the GPU with larger tiles than a single cell can (fewer
halo rows and planes recomputed, enough blocks either way), so batched plans are tuned on a plan of
`batch_cells(shape)` cells and keyed separately ("|batch").
"""
import ctypes
import json
import os

import numpy as np
import torch

from . import _lib

TY_CHOICES = (1, 2, 4, 6, 8, 14, 30, 62)
XC_CHOICES = (2, 4, 8, 16, 32, 64, 128)
ZPT_CHOICES = (1, 2, 3, 4)


def candidates(shape, elem_bytes=8, precision="f32"):
    """Tilings (ty, xc, zpt) for `shape` that fit a 512-thread block and the LDS of elem_bytes-wide kernels. FP64 plans
    get no three-z-value tilings: the FP64 PCG's direction-update kernels need more than 256 registers there and spill
    (the standalone operator the sweep times does not, so the sweep alone would pick them)."""
    nx, ny, nz = shape
    out = []
    for zpt in ZPT_CHOICES:
        if nz % zpt or nz // zpt > 64 or (precision == "f64" and zpt == 3):
            continue
        lz = nz // zpt
        lzs = 1
        while lzs < lz:
            lzs <<= 1
        for ty in TY_CHOICES:
            if ty > ny or lzs * (ty + 2) > 512:
                continue
            if 3 * (ty + 2) * 6 * nz * elem_bytes > 159 * 1024:
                continue
            for xc in XC_CHOICES:
                if xc > nx:
                    continue
                out.append((ty, xc, zpt))
    return out


_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tiling_gfx950.json")


def batch_cells(shape):
    """Cells of `shape` in the representative batched launch used for tuning: about 2^21 voxels."""
    nvox = int(np.prod(shape))
    return max(1, min(64, (1 << 21) // nvox))


def _key(shape, precision, m, batched, kind=""):
    return f"{tuple(shape)}|{precision}|{m}" + (f"|{kind}" if kind else "") + ("|batch" if batched else "")


def _packaged(shape, precision, m, dev, batched=False, kind=""):
    """Tiling measured on MI355X (gfx950) and shipped with the package (tune_table.py writes it)."""
    try:
        arch = torch.cuda.get_device_properties(dev).gcnArchName
    except Exception:
        return None
    if not arch.startswith("gfx950"):
        return None
    try:
        with open(_TABLE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    e = table.get(_key(shape, precision, m, batched, kind)) or table.get(_key(shape, precision, m, False, kind))
    return tuple(e["best"]) if e else None


def _cache_path():
    return os.environ.get("NLFH_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "nlfh", "tiling.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except Exception:
        return {}


def _save_cache(c):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(c, f, indent=1)
    except OSError:
        pass


def time_matvec(plan, s, u, out, m, dtype, reps):
    lib = _lib.load()
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    args = (plan.handle, ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            m, dtype, sp)
    for _ in range(2):
        _lib.check(lib.nlfh_apply_K(*args))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        _lib.check(lib.nlfh_apply_K(*args))
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e3  # microseconds


def tune(plan, shape_index=0, m=6, precision="f64", reps=10, use_cache=True, verbose=False, batched=False):
    """Pick the fastest tiling for distinct shape `shape_index` of `plan` (a linear.Plan) and set
    it. With `batched`, candidates are timed on a plan of batch_cells(shape) cells of that shape (the
    caller's plan is only tiled). Returns (best (ty, xc, zpt), [(tiling, microseconds), ...])."""
    # the distinct shapes in order of first appearance
    distinct = []
    for s in plan.shapes:
        if s not in distinct:
            distinct.append(s)
    shape = distinct[shape_index]
    batched = batched and batch_cells(shape) > 1
    dev = torch.device("cuda", torch.cuda.current_device())
    key = f"{torch.cuda.get_device_name(dev)}|" + _key(shape, precision, m, batched)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = tuple(cache[key]["best"])
        plan.set_tiling(shape_index, *best)
        return best, []
    if use_cache:
        best = _packaged(shape, precision, m, dev, batched)
        if best is not None:
            plan.set_tiling(shape_index, *best)
            return best, []
    dtype = 0 if precision == "f64" else 1
    tdt = torch.float64 if precision == "f64" else torch.float32
    if batched:
        from .linear import Plan
        nb = batch_cells(shape)
        tplan, ti = Plan([shape] * nb, 1.0, 0.3, m_layout=m), 0
    else:
        # a single-cell plan owns the voxel range of this shape at offset 0
        if len(plan.shapes) != 1:
            raise ValueError("tune() times single-cell plans; tune each shape with its own plan")
        nb, tplan, ti = 1, plan, shape_index
    nvox = int(np.prod(shape))
    s = torch.rand(nb * nvox, dtype=torch.float64, device=dev) * 0.99 + 0.01
    u = torch.randn(nb * m * 3 * nvox, dtype=tdt, device=dev)
    out = torch.empty_like(u)
    table = []
    for cand in candidates(shape, 8, precision):  # the plan's tiling must fit the float64 kernels too
        tplan.set_tiling(ti, *cand)
        t = time_matvec(tplan, s, u, out, m, dtype, reps)
        table.append((cand, t))
        if verbose:
            print(f"  tiling ty={cand[0]:3d} xc={cand[1]:3d} zpt={cand[2]}: {t:9.2f} us", flush=True)
    table.sort(key=lambda x: x[1])
    best = table[0][0]
    plan.set_tiling(shape_index, *best)
    if use_cache:
        cache[key] = {"best": list(best), "us": table[0][1]}
        _save_cache(cache)
    return best, table


def time_precond(plan, r, z, m, reps):
    """Microseconds per FP32 nlfh_precond over the plan's cells (HIP events, current stream)."""
    lib = _lib.load()
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    args = (plan.handle, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(z.data_ptr()), m, 1, sp)
    for _ in range(2):
        _lib.check(lib.nlfh_precond(*args))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        _lib.check(lib.nlfh_precond(*args))
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e3


def time_inner_iterations(plan, s, m, iters):
    """Microseconds per FP32 inner PCG iteration of a mixed solve over the plan's cells (HIP events around one solve
    with the iteration cap `iters`, which ends unconverged): the spectral pass A only runs there."""
    lib = _lib.load()
    dev = s.device
    nc = len(plan.shapes)
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    u = torch.empty(int(s.numel()) * 3 * m, dtype=torch.float64, device=dev)
    es = (ctypes.c_double * (6 * m))(*[1.0 if j == i else 0.0 for i in range(m) for j in range(6)])
    it = (ctypes.c_int * nc)()
    cv = (ctypes.c_int * nc)()
    rr = (ctypes.c_double * (nc * m))()

    def solve():
        _lib.check(lib.nlfh_solve(plan.handle, ctypes.c_void_p(s.data_ptr()), es, m, None, ctypes.c_void_p(u.data_ptr()),
                                  1e-14, iters, 1, it, cv, rr, sp))

    solve()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    solve()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / max(1, min(it))


def fft_candidates(default):
    """(pass B positions per block, pass A planes per block) to try: the default, half and double positions, each
    with one and two planes."""
    return [(b, q) for b in (default, default // 2, 2 * default) for q in (1, 2)]


def tune_fft(plan, shape_index=0, m=6, reps=10, use_cache=True, verbose=False, batched=False):
    """Pick the FP32 preconditioner block shapes of distinct shape `shape_index` of `plan` (cubic 16-128 shapes;
    others are left alone) and set them: pass B's positions per block by timing nlfh_precond in FP32, then pass A's
    planes per block by timing FP32 inner iterations of a capped mixed solve (the spectral pass A runs only there),
    on a plan of batch_cells(shape) cells with `batched`. Returns ((positions, planes) or None, [(candidate,
    microseconds), ...])."""
    distinct = []
    for s in plan.shapes:
        if s not in distinct:
            distinct.append(s)
    shape = distinct[shape_index]
    _, default, _ = plan.get_fft_tiling(shape_index)
    if default == 0:
        return None, []
    batched = batched and batch_cells(shape) > 1
    dev = torch.device("cuda", torch.cuda.current_device())
    key = f"{torch.cuda.get_device_name(dev)}|" + _key(shape, "f32", m, batched, "fftb")
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache and len(cache[key]["best"]) == 2:
        best = tuple(int(v) for v in cache[key]["best"])
        plan.set_fft_tiling(shape_index, *best)
        return best, []
    if use_cache:
        best = _packaged(shape, "f32", m, dev, batched, "fftb")
        if best is not None and len(best) == 2:
            best = tuple(int(v) for v in best)
            plan.set_fft_tiling(shape_index, *best)
            return best, []
    from .linear import Plan
    nb = batch_cells(shape) if batched else 1
    tplan = Plan([shape] * nb, 1.0, 0.3, m_layout=m)
    nvox = int(np.prod(shape))
    r = torch.randn(nb * m * 3 * nvox, dtype=torch.float32, device=dev)
    z = torch.empty_like(r)
    table = []
    for cand in sorted({c[0] for c in fft_candidates(default)}, key=lambda b: (b != default, b)):
        try:
            tplan.set_fft_tiling(0, cand, 1)
        except _lib.NlfhError:
            continue
        t = time_precond(tplan, r, z, m, reps)
        table.append(((cand, 1), t))
        if verbose:
            print(f"  pass B positions per block {cand:4d}: {t:9.2f} us per application", flush=True)
    table.sort(key=lambda x: x[1])
    b_best = table[0][0][0]
    # pass A planes per block: whole FP32 inner iterations with the chosen pass B
    s = torch.rand(nb * nvox, dtype=torch.float64, device=dev) * 0.99 + 0.01
    iters = 64
    t_it = []
    for q in (1, 2):
        try:
            tplan.set_fft_tiling(0, b_best, q)
        except _lib.NlfhError:  # one plane per block only (the in-place 128-point passes)
            continue
        t = time_inner_iterations(tplan, s, m, iters)
        t_it.append(((b_best, q), t))
        if verbose:
            print(f"  pass A planes per block {q}: {t:9.2f} us per inner iteration", flush=True)
    t_it.sort(key=lambda x: x[1])
    best = t_it[0][0]
    plan.set_fft_tiling(shape_index, *best)
    if use_cache:
        cache[key] = {"best": list(best), "us_per_iteration": t_it[0][1]}
        _save_cache(cache)
    return best, table + t_it

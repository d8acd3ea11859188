"""The learned warm-start network on the GPU against its float64 torch-CPU reference: forward output
and every weight and bias gradient (BF16 storage, FP32 accumulation on the GPU; the reference rounds
weights and activations to bfloat16 the same way, so the forward agrees to ~1e-2 and the gradients,
which the GPU also stores in bfloat16 between layers, to a few 1e-2 of their largest entry)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import convnet_ref as R
from oracle import linear_ref as O

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_reference(gpu_pkg):
    from nlfh_amd.learned import WarmStartNet
    net = WarmStartNet(seed=3)
    # non-zero biases and head weights so every path carries signal
    rng = np.random.default_rng(4)
    params = net.params()
    for k in params:
        if k.endswith(".b"):
            params[k] = (0.1 * rng.standard_normal(params[k].shape)).astype(np.float32)
    params["out.b"][18:] = 0.0
    params["out"][:, :18, :] = (rng.standard_normal(params["out"][:, :18, :].shape) * 0.05).astype(np.float32)
    net = WarmStartNet(params=params)
    B, n = 2, 16
    rho = torch.tensor(rng.uniform(0, 1, (B, n, n, n)))
    out = net.forward(net.input(rho.cuda()), keep=True)
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in params.items()}
    ref = R.forward(P, rho.double())
    o = out.cpu().double()
    assert torch.abs(o[..., :18] - ref[..., :18]).max() / torch.abs(ref[..., :18]).max() < 2e-2
    assert torch.abs(o[..., 18:]).max() == 0.0
    target = torch.tensor(rng.standard_normal((B, n, n, n, 18)), dtype=torch.float64)
    L, dout = net.loss_and_grad(out, target.float().cuda())
    grads = net.backward(dout)
    Lr = R.loss(ref, target)
    Lr.backward()
    assert abs(L - float(Lr)) < 2e-2 * float(Lr)
    for name, *_ in R.LAYERS:
        dw, db = grads[name]
        gw = P[name].grad
        gb = P[name + ".b"].grad
        ew = float(torch.abs(dw.cpu().double() - gw).max() / torch.abs(gw).max().clamp_min(1e-30))
        eb = float(torch.abs(db.cpu().double() - gb).max() / torch.abs(gb).max().clamp_min(1e-30))
        assert ew < 5e-2, (name, ew)
        assert eb < 5e-2, (name, eb)


def _bf16(x):
    return x.to(torch.bfloat16).to(torch.float64)


@pytest.mark.parametrize("n,ca,cb,a_half,cout,coutp,act,flip", [
    (16, 16, 0, 0, 16, 16, 1, 0),    # brick 4x4x16, 16-channel chunks
    (32, 32, 0, 0, 32, 32, 1, 1),    # 32-channel chunks, flipped taps
    (16, 64, 32, 1, 32, 32, 1, 0),   # half-resolution source + skip source (decoder)
    (8, 32, 16, 1, 16, 16, 0, 0),    # brick 4x8x8, mixed chunk width falls to 16
    (8, 32, 0, 0, 96, 96, 0, 1),     # six output blocks
    (16, 16, 0, 0, 18, 32, 0, 0),    # head: padded output channels
    (12, 16, 16, 0, 32, 32, 1, 0),   # a side the bricks do not tile: the generic kernel
    (64, 16, 0, 0, 16, 16, 1, 0),    # full-resolution layer of a 64^3 cell
])
def test_conv_layer_matches_torch(gpu_pkg, n, ca, cb, a_half, cout, coutp, act, flip):
    """nlfh_net_conv against torch's conv3d in float64 on the same bf16-rounded operands (circular padding;
    flip = 1 reads the taps mirrored, as the data gradient does): FP32 accumulation only, so the outputs
    agree to 1e-4 of their largest entry."""
    import ctypes
    import torch.nn.functional as F
    lib = gpu_pkg.load_library()
    rng = np.random.default_rng(n + ca + cb + cout)
    B = 2
    na = n // 2 if a_half else n
    A = _bf16(torch.tensor(rng.standard_normal((B, na, na, na, ca))))
    Bs = _bf16(torch.tensor(rng.standard_normal((B, n, n, n, cb)))) if cb else None
    W = _bf16(torch.tensor(rng.standard_normal((27, coutp, ca + cb)) * 0.1))
    bias = torch.tensor(rng.standard_normal(cout) * 0.1)
    # reference: nearest upsampling of A, concatenation, periodic 3x3x3 taps
    Af = A.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3) if a_half else A
    X = torch.cat([Af, Bs], dim=-1) if cb else Af
    Wk = W.reshape(3, 3, 3, coutp, ca + cb)
    if flip:
        Wk = Wk.flip(0, 1, 2)
    Wk = Wk.permute(3, 4, 0, 1, 2)
    xt = F.pad(X.permute(0, 4, 1, 2, 3), (1, 1, 1, 1, 1, 1), mode="circular")
    ref = (F.conv3d(xt, Wk) + F.pad(bias, (0, coutp - cout))[None, :, None, None, None]).permute(0, 2, 3, 4, 1)
    ref = ref[..., :cout]
    if act:
        ref = F.leaky_relu(ref, 0.1)

    def dev16(t):
        return t.float().to(torch.bfloat16).view(torch.int16).cuda().contiguous()

    a_d, w_d = dev16(A), dev16(W)
    b_d = dev16(Bs) if cb else None
    bias_d = bias.float().cuda()
    out = torch.full((B, n, n, n, cout), float("nan"), dtype=torch.float32, device="cuda")
    outh = torch.empty((B, n, n, n, cout), dtype=torch.int16, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gpu_pkg._lib.check(lib.nlfh_net_conv(P(a_d), ca, a_half, P(b_d), cb, P(w_d), P(bias_d), n, B, cout, coutp, act,
                                         flip, P(out), P(outh), st))
    o = out.cpu().double()
    assert torch.isfinite(o).all()
    assert float((o - ref).abs().max() / ref.abs().max()) < 1e-4
    oh = outh.cpu().view(torch.bfloat16).double()
    assert torch.equal(oh, out.cpu().to(torch.bfloat16).double())  # the bf16 copy is the rounded output


@pytest.mark.parametrize("n,ca,cb,a_half,cg,act", [
    (64, 32, 16, 1, 16, 1),   # runs of several bricks per block (three input blocks), half-resolution source
    (16, 64, 0, 0, 64, 1),    # four output-gradient blocks
    (12, 16, 16, 0, 32, 0),   # a side the bricks do not tile: the generic kernels
])
def test_weight_and_bias_gradients_match_torch(gpu_pkg, n, ca, cb, a_half, cg, act):
    """nlfh_net_wgrad, dW[t][co][ci] = sum_v g[v][co] in[v + d_t][ci] (periodic, the concatenated and upsampled
    input of nlfh_net_conv), and nlfh_net_bias_grad, db[c] = sum_v dy[v][c] act'(y[v][c]), against float64
    torch on the same bf16 operands (FP32 accumulation: 1e-4 of the largest entry)."""
    import ctypes
    lib = gpu_pkg.load_library()
    rng = np.random.default_rng(n + ca + cg)
    B = 2
    na = n // 2 if a_half else n
    A = _bf16(torch.tensor(rng.standard_normal((B, na, na, na, ca))))
    Bs = _bf16(torch.tensor(rng.standard_normal((B, n, n, n, cb)))) if cb else None
    G = _bf16(torch.tensor(rng.standard_normal((B, n, n, n, cg))))
    Af = A.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3) if a_half else A
    X = torch.cat([Af, Bs], dim=-1) if cb else Af
    cin = ca + cb
    ref = torch.empty((27, cg, cin), dtype=torch.float64)
    for t in range(27):
        d = (t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1)
        Xs = torch.roll(X, shifts=(-d[0], -d[1], -d[2]), dims=(1, 2, 3))  # Xs[v] = X[v + d]
        ref[t] = torch.einsum("bxyzo,bxyzi->oi", G, Xs)

    def dev16(t):
        return t.float().to(torch.bfloat16).view(torch.int16).cuda().contiguous()

    P = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    part = torch.empty(16 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dw = torch.full((27, cg, cin), float("nan"), dtype=torch.float32, device="cuda")
    a_d, g_d = dev16(A), dev16(G)
    b_d = dev16(Bs) if cb else None
    gpu_pkg._lib.check(lib.nlfh_net_wgrad(P(g_d), cg, P(a_d), ca, a_half, P(b_d), cb, n, B, P(part), part.numel(),
                                          P(dw), st))
    w = dw.cpu().double()
    assert torch.isfinite(w).all()
    assert float((w - ref).abs().max() / ref.abs().max()) < 1e-4
    # bias gradient with and without the partial-sum scratch
    dy = torch.tensor(rng.standard_normal((B * n ** 3, cg)), dtype=torch.float32)
    y = _bf16(torch.tensor(rng.standard_normal((B * n ** 3, cg))))
    gref = dy.double() * torch.where(y <= 0, 0.1, 1.0) if act else dy.double()
    dbr = gref.sum(0)
    for use_part in (True, False):
        db = torch.full((cg,), float("nan"), dtype=torch.float32, device="cuda")
        y_d = dev16(y)
        gpu_pkg._lib.check(lib.nlfh_net_bias_grad(P(dy.cuda()), P(y_d) if act else None, B * n ** 3, cg, act,
                                                  P(part) if use_part else None, part.numel() if use_part else 0,
                                                  P(db), st))
        torch.cuda.synchronize()
        assert float((db.cpu().double() - dbr).abs().max() / dbr.abs().max()) < 1e-4


def test_training_reduces_loss(gpu_pkg):
    """Adam steps on correctors produced by the solver lower the loss; the prediction used as u0
    gives the same C as a cold start."""
    from nlfh_amd.learned import WarmStartNet
    n = 16
    rhos, targets = [], []
    for seed in range(4):
        rho, _ = gpu_pkg.cells.random_cell(n, seed)
        lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-4, tol=1e-8, precision="mixed")
        _, u, _ = lh.C_full(rho)
        rhos.append(torch.tensor(rho))
        targets.append(u.permute(1, 2, 3, 0, 4).reshape(n, n, n, 18).float() / n)
    rho_b = torch.stack(rhos).cuda()
    tgt = torch.stack(targets).cuda()
    net = WarmStartNet(seed=0)
    x0 = net.input(rho_b)
    losses = []
    for _ in range(60):
        out = net.forward(x0, keep=True)
        L, dout = net.loss_and_grad(out, tgt)
        net.adam(net.backward(dout), lr=1e-3)
        losses.append(L)
    assert losses[-1] < 0.5 * losses[0]
    lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-4, tol=1e-8, precision="fp64")
    C0, _, it0 = lh.C_full(rhos[0])
    u0 = net.predict(rhos[0])[0]
    C1, _, it1 = lh.C_full(rhos[0], u0=u0)
    print("losses", losses[0], losses[-1], "iterations cold/warm", it0, it1)
    assert np.abs(np.asarray(C1) - np.asarray(C0)).max() < 1e-7 * np.abs(np.asarray(C0)).max()


def test_ritz_start_matches_oracle(gpu_pkg):
    """nlfh_ritz_start against the Galerkin system built from the oracle's K and load: for each
    right-hand side I, a_I solves (P_J . K P_K) a_I = -(f_I . P_J)."""
    from oracle import linear_ref as O
    n = 8
    rho, _ = gpu_pkg.cells.random_cell(n, 7)
    ref = O.LinearRef((n, n, n), emin=1e-3)
    s = ref.stiffness_scale(rho)
    rng = np.random.default_rng(11)
    P = rng.standard_normal((6, n, n, n, 3))
    P -= P.mean(axis=(1, 2, 3), keepdims=True)
    Es = np.eye(6)
    KP = np.stack([ref.apply_K(s, P[j]) for j in range(6)])
    f = ref.load(s, Es)
    G = P.reshape(6, -1) @ KP.reshape(6, -1).T
    H = f.reshape(6, -1) @ P.reshape(6, -1).T
    A = np.linalg.solve(G, -H.T).T
    lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3)
    out, coef = lh.ritz_start(torch.tensor(s), Es, torch.tensor(P))
    assert np.abs(coef - A).max() < 1e-9 * np.abs(A).max()
    want = np.einsum("ij,j...->i...", A, P)
    assert np.abs(out.cpu().numpy() - want).max() < 1e-9 * np.abs(want).max()


def test_ritz_start_reproduces_exact_correctors(gpu_pkg):
    """Started from the converged correctors, the Ritz weights are the identity (the fields already
    minimise the energy), and a zero prediction gives the zero start."""
    n = 16
    rho, _ = gpu_pkg.cells.random_cell(n, 2)
    lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3, tol=1e-12)
    s = lh.stiffness_scale(torch.tensor(rho))
    Es = np.eye(6)
    u, _ = lh.solve(s, Es)
    out, coef = lh.ritz_start(s, Es, u)
    assert np.abs(coef - np.eye(6)).max() < 1e-8
    assert float((out - u).abs().max()) < 1e-8 * float(u.abs().max())
    out0, coef0 = lh.ritz_start(s, Es, torch.zeros_like(u))
    assert np.all(coef0 == 0) and float(out0.abs().max()) == 0.0


def test_learned_start_falls_back_to_zero_start(gpu_pkg):
    """A warm-started C_full whose solve cannot converge (iteration cap) re-solves from zero and says so."""
    from nlfh_amd.learned import WarmStartNet
    n = 16
    rho, _ = gpu_pkg.cells.random_cell(n, 5)
    net = WarmStartNet(seed=1)
    lh = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3, tol=1e-8, maxit=3, warm_start=net)
    lh.C_full(torch.tensor(rho))
    assert lh.last_info["warm_start"] and lh.last_info["fallback"]
    lh2 = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3, tol=1e-8, warm_start=net)
    C, _, _ = lh2.C_full(torch.tensor(rho))
    C0, _, _ = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-3, tol=1e-8).C_full(torch.tensor(rho))
    assert lh2.last_info["warm_start"] and not lh2.last_info["fallback"]
    assert np.abs(np.asarray(C) - np.asarray(C0)).max() < 1e-6 * np.abs(np.asarray(C0)).max()


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 6), (16, 16, 16)])
def test_cheb_smooth_matches_oracle(gpu_pkg, shape):
    """nlfh_cheb_smooth (FP32 steps fused into the operator launch) against the oracle's FP64 three-term
    recurrence, 24 steps on [hi / 50, hi] with hi the element bound, six strains from a random start."""
    rng = np.random.default_rng(sum(shape))
    rho = rng.uniform(0.0, 1.0, shape)  # the operator test needs no smooth density
    lh = gpu_pkg.LinearHomogenization(shape, emin=1e-3)
    s = lh.stiffness_scale(torch.tensor(rho))
    Es = np.eye(6)
    u0 = rng.standard_normal((6,) + shape + (3,)) * 0.1
    hi = O.jacobi_bound()
    x = lh.cheb_smooth(s, Es, torch.tensor(u0), degree=24, ratio=50.0).cpu().numpy()
    ref = O.LinearRef(shape, emin=1e-3)
    xr = O.cheb_smooth(ref, s.cpu().numpy(), Es, u0, 24, hi / 50.0, hi)
    assert np.abs(x - xr).max() < 2e-4 * np.abs(xr).max()
    # degree 0 returns the start itself
    x0 = lh.cheb_smooth(s, Es, torch.tensor(u0), degree=0).cpu().numpy()
    assert np.array_equal(x0, u0)


def test_cheb_smooth_batched_cells_match_single(gpu_pkg):
    """On a batched plan of uneven cells the smoothing of every cell equals that cell smoothed alone (the launch
    covers all cells; the diagonal is per cell, periodic)."""
    from nlfh_amd.linear import _es_host
    shapes = [(16, 16, 16), (8, 8, 8), (12, 12, 12)]
    rhos = [gpu_pkg.cells.random_cell(n[0], 30 + i)[0] for i, n in enumerate(shapes)]
    bh = gpu_pkg.BatchedHomogenization(shapes, emin=1e-4, streams=1)
    rd = bh.pack(rhos)
    s = bh.stiffness_scale(rd)
    rng = np.random.default_rng(1)
    u = torch.tensor(rng.standard_normal(bh.tot * 18) * 0.1, device="cuda")
    out = torch.empty_like(u)
    lib = gpu_pkg.load_library()
    Es_np, Es_p = _es_host(np.eye(6))
    hi = float(lib.nlfh_jacobi_bound(1.0, 0.3))
    import ctypes
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gpu_pkg._lib.check(lib.nlfh_cheb_smooth(bh._plan.handle, ctypes.c_void_p(s.data_ptr()), Es_p, 6,
                                            ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(out.data_ptr()), 16,
                                            hi / 100.0, hi, None, st))
    views_in, views_out = bh.field_views(u), bh.field_views(out)
    for i, sh in enumerate(shapes):
        lh = gpu_pkg.LinearHomogenization(sh, emin=1e-4)
        si = lh.stiffness_scale(torch.tensor(rhos[i]))
        xi = lh.cheb_smooth(si, np.eye(6), views_in[i].contiguous(), degree=16, ratio=100.0)
        assert float((xi - views_out[i]).abs().max()) <= 1e-5 * float(xi.abs().max())


def test_cheb_first_step_beyond_2pow24_voxels(gpu_pkg):
    """Regression: on a plan of more voxels than one launch of 65536 blocks x 256 threads covers (65 cells of
    64^3, 17.0 M voxels) the Jacobi diagonal reaches every cell: one Chebyshev step from zero is
    x_1 = -D^-1 f / theta in every cell (checked on the first and the last three cells against D from rolled sums
    of s and the plan's load)."""
    import ctypes
    from nlfh_amd.linear import _es_host
    n, nc = 64, 65
    bh = gpu_pkg.BatchedHomogenization([(n, n, n)] * nc, emin=1e-3, streams=1)
    assert bh.tot > 65536 * 256
    g = torch.Generator(device="cuda").manual_seed(5)
    rd = torch.rand(bh.tot, dtype=torch.float64, device="cuda", generator=g)
    s = bh.stiffness_scale(rd)
    lib = gpu_pkg.load_library()
    Es_np, Es_p = _es_host(np.eye(6))
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    hi = float(lib.nlfh_jacobi_bound(1.0, 0.3))
    lo = hi / 900.0
    z = torch.zeros(bh.tot * 18, dtype=torch.float64, device="cuda")
    f = torch.empty_like(z)
    x1 = torch.empty_like(z)
    gpu_pkg._lib.check(lib.nlfh_load(bh._plan.handle, P(s), Es_p, 6, P(f), st))
    gpu_pkg._lib.check(lib.nlfh_cheb_smooth(bh._plan.handle, P(s), Es_p, 6, P(z), P(x1), 1, lo, hi, None, st))
    lam, mu = 0.3 / (1.3 * 0.4), 1.0 / 2.6
    fv, xv = bh.field_views(f), bh.field_views(x1)
    for i in (0, nc - 3, nc - 2, nc - 1):
        sc = s[bh.off[i]:bh.off[i + 1]].view(n, n, n)
        D = (lam + 4 * mu) / 9 * sum(torch.roll(sc, (a, b, c), (0, 1, 2)) for a in (0, 1) for b in (0, 1) for c in (0, 1))
        want = -fv[i] / D[None, ..., None] / (0.5 * (hi + lo))
        assert float((xv[i] - want).abs().max() / want.abs().max()) < 1e-5, i


def test_trained_network_on_heldout_cells(gpu_pkg):
    """The packaged network (trained on 5000 solver-generated 64^3 cells) on held-out cells: its prediction is
    a real approximation of the correctors (relative L2 error well below 1), the learned start (prediction,
    128 Chebyshev-Jacobi steps, Ritz weights) gives the same C as a zero start and needs fewer PCG iterations
    than the zero start on every cell; without the smoothing the raw Ritz start needs more (its error is
    rough in the soft phase, where the reference-medium preconditioner is slowest)."""
    import os
    from nlfh_amd.learned import WarmStartNet
    path = os.path.join(os.path.dirname(gpu_pkg.__file__), "data", "warmstart64.npz")
    net = WarmStartNet.load(path)
    n = 64
    kw = dict(emin=1e-4, tol=1e-8, precision="mixed")
    cold = gpu_pkg.LinearHomogenization((n, n, n), **kw)
    warm = gpu_pkg.LinearHomogenization((n, n, n), warm_start=net, **kw)
    rough = gpu_pkg.LinearHomogenization((n, n, n), warm_start=net, warm_smoothing=None, **kw)
    ratios, rough_ratios, errs = [], [], []
    for seed in (20000, 20001, 20002, 20003):
        rho, _ = gpu_pkg.cells.random_cell(n, seed)
        r = torch.tensor(rho)
        C0, u0, it0 = cold.C_full(r)
        C1, _, it1 = warm.C_full(r)
        assert warm.last_info["warm_start"] and not warm.last_info["fallback"]
        c0, c1 = np.asarray(C0), np.asarray(C1)
        assert np.abs(c1 - c0).max() < 1e-6 * np.abs(c0).max()
        _, _, it2 = rough.C_full(r)
        P = net.predict(r)[0]
        errs.append(float((P - u0).norm() / u0.norm()))
        ratios.append(it1 / it0)
        rough_ratios.append(it2 / it0)
    print("held-out relative L2 error", errs, "iteration ratio smoothed learned / zero", ratios,
          "unsmoothed / zero", rough_ratios)
    assert max(errs) < 0.4
    assert max(ratios) < 0.97
    assert min(rough_ratios) > 1.0


@pytest.mark.parametrize("B,n", [(3, 16), (2, 4)])
def test_readout_matches_torch(gpu_pkg, B, n):
    """nlfh_net_readout (head channels scaled by n, mean-free per cell and channel, float64 planar layout) against the
    same steps in torch; n = 4 leaves most of the 512 voxel chunks empty."""
    from nlfh_amd import learned as L
    lib = gpu_pkg.load_library()
    out = torch.randn((B, n, n, n, L.COUT_HEAD), dtype=torch.float32, device="cuda")
    ref = out[..., :L.NOUT].double() * n
    ref = ref.reshape(B, n, n, n, 6, 3).permute(0, 4, 1, 2, 3, 5)
    ref = ref - ref.mean(dim=(2, 3, 4), keepdim=True)
    dst = torch.empty((B, 6, 3, n, n, n), dtype=torch.float64, device="cuda")
    part = torch.empty(B * 513 * L.NOUT, dtype=torch.float64, device="cuda")
    rc = lib.nlfh_net_readout(ctypes.c_void_p(out.data_ptr()), L.COUT_HEAD, L.NOUT, n ** 3, B, float(n),
                              ctypes.c_void_p(part.data_ptr()), part.numel(), ctypes.c_void_p(dst.data_ptr()),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = dst.permute(0, 1, 3, 4, 5, 2)
    assert float((got - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    short = torch.empty(B * 513 * L.NOUT - 1, dtype=torch.float64, device="cuda")
    assert lib.nlfh_net_readout(ctypes.c_void_p(out.data_ptr()), L.COUT_HEAD, L.NOUT, n ** 3, B, float(n),
                                ctypes.c_void_p(short.data_ptr()), short.numel(), ctypes.c_void_p(dst.data_ptr()),
                                None) != 0


def test_input_matches_float32_path(gpu_pkg):
    """nlfh_net_input (bf16 channels-last input, channel 0 = 2 rho - 1 with rho rounded to float32) against the float32
    construction in torch converted by nlfh_net_f2bf: bitwise."""
    lib = gpu_pkg.load_library()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rho = torch.rand((2, 8, 8, 8), dtype=torch.float64, device="cuda")
    x = torch.zeros((2, 8, 8, 8, 16), dtype=torch.float32, device="cuda")
    x[..., 0] = 2.0 * rho.float() - 1.0
    ref = torch.empty(x.shape, dtype=torch.int16, device="cuda")
    assert lib.nlfh_net_f2bf(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(ref.data_ptr()), x.numel(), st) == 0
    got = torch.full(x.shape, 7, dtype=torch.int16, device="cuda")
    assert lib.nlfh_net_input(ctypes.c_void_p(rho.data_ptr()), rho.numel(), ctypes.c_void_p(got.data_ptr()), st) == 0
    assert torch.equal(got, ref)

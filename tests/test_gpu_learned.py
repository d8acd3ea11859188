"""The learned warm-start network on the GPU against its float64 torch-CPU reference: forward output
and every weight and bias gradient (BF16 storage, FP32 accumulation on the GPU; the reference rounds
weights and activations to bfloat16 the same way, so the forward agrees to ~1e-2 and the gradients,
which the GPU also stores in bfloat16 between layers, to a few 1e-2 of their largest entry)."""
import numpy as np
import pytest
import torch

from oracle import convnet_ref as R

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


def test_trained_network_on_heldout_cells(gpu_pkg):
    """The packaged network (trained on 5000 solver-generated 64^3 cells) on held-out cells: its
    prediction is a real approximation of the correctors (relative L2 error well below 1), the
    learned start gives the same C as a zero start, and its PCG iteration count relative to the zero
    start is recorded. The measured effect at tol 1e-8 is a small increase (about +10% over 512
    held-out cells): the slow PCG modes live in the soft phase, where the prediction's error is not
    smooth enough to help."""
    import os
    from nlfh_amd.learned import WarmStartNet
    path = os.path.join(os.path.dirname(gpu_pkg.__file__), "data", "warmstart64.npz")
    net = WarmStartNet.load(path)
    n = 64
    cold = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-4, tol=1e-8, precision="mixed")
    warm = gpu_pkg.LinearHomogenization((n, n, n), emin=1e-4, tol=1e-8, precision="mixed", warm_start=net)
    ratios, errs = [], []
    for seed in (20000, 20001):
        rho, _ = gpu_pkg.cells.random_cell(n, seed)
        r = torch.tensor(rho)
        C0, u0, it0 = cold.C_full(r)
        C1, _, it1 = warm.C_full(r)
        assert warm.last_info["warm_start"] and not warm.last_info["fallback"]
        c0, c1 = np.asarray(C0), np.asarray(C1)
        assert np.abs(c1 - c0).max() < 1e-6 * np.abs(c0).max()
        P = net.predict(r)[0]
        errs.append(float((P - u0).norm() / u0.norm()))
        ratios.append(it1 / it0)
    print("held-out relative L2 error", errs, "iteration ratio learned/zero", ratios)
    assert max(errs) < 0.4
    assert max(ratios) < 1.25

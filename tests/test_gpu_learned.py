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

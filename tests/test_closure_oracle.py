"""Pins the learned closure's FP64 oracle (oracle/closure.py) on the CPU: forward and gradients
against torch's FP64 autograd (conv2d + exact GELU, an independent library), gradients against
central finite differences, BF16 rounding against torch's own conversion, Adam against
torch.optim.Adam, and the gradient model / Π contraction against closed forms."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import closure as OC

CH = (4, 6, 5, 3)


def _params(seed=0, ch=CH):
    rng = np.random.Generator(np.random.PCG64(seed))
    ps = []
    for ci, co in zip(ch[:-1], ch[1:]):
        ps.append((rng.standard_normal((co, ci, 5, 5)) / np.sqrt(ci * 25) * 1.4, 0.1 * rng.standard_normal(co)))
    return ps


def _torch_net(params, x, target):
    ws = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w, _ in params]
    bs = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for _, b in params]
    a = torch.tensor(x, dtype=torch.float64)
    for i, (w, b) in enumerate(zip(ws, bs)):
        a = F.conv2d(a, w, b)
        if i < len(ws) - 1:
            a = F.gelu(a, approximate="none")
    loss = torch.mean((a - torch.tensor(target)) ** 2)
    loss.backward()
    return a.detach().numpy(), float(loss.detach()), [(w.grad.numpy(), b.grad.numpy()) for w, b in zip(ws, bs)]


def test_forward_and_gradients_match_torch_fp64_autograd():
    rng = np.random.Generator(np.random.PCG64(1))
    params = _params()
    x = rng.standard_normal((2, CH[0], 21, 19))
    target = rng.standard_normal((2, CH[-1], 21 - 12, 19 - 12))
    y, _, _ = OC.forward(params, x, bf16=False)
    yt, lt, gt = _torch_net(params, x, target)
    np.testing.assert_allclose(y, yt, rtol=1e-12, atol=1e-13)
    loss, grads = OC.backward(params, x, target, bf16=False)
    assert abs(loss - lt) <= 1e-12 * abs(lt)
    for (dw, db), (tw, tb) in zip(grads, gt):
        np.testing.assert_allclose(dw, tw, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(db, tb, rtol=1e-10, atol=1e-14)


def test_gradients_match_central_differences():
    rng = np.random.Generator(np.random.PCG64(2))
    params = _params(3)
    x = rng.standard_normal((1, CH[0], 17, 16))
    target = rng.standard_normal((1, CH[-1], 5, 4))
    _, grads = OC.backward(params, x, target, bf16=False)
    h = 1e-6
    picks = [(0, (1, 2, 0, 3)), (1, (4, 5, 2, 2)), (2, (2, 0, 4, 1))]
    for li, idx in picks:
        w = params[li][0]
        old = w[idx]
        w[idx] = old + h
        lp = OC.backward(params, x, target, bf16=False)[0]
        w[idx] = old - h
        lm = OC.backward(params, x, target, bf16=False)[0]
        w[idx] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[li][0][idx]) <= 1e-6 * max(1e-3, abs(fd)), (li, idx, fd, grads[li][0][idx])
    for li in range(len(params)):
        b = params[li][1]
        b[0] += h
        lp = OC.backward(params, x, target, bf16=False)[0]
        b[0] -= 2 * h
        lm = OC.backward(params, x, target, bf16=False)[0]
        b[0] += h
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[li][1][0]) <= 1e-6 * max(1e-3, abs(fd))


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.Generator(np.random.PCG64(4))
    a = rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))
    ref = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(OC.round_bf16(a), ref)
    ulp = 2.0 ** -7
    assert OC.round_bf16(1.0 + ulp / 2) == 1.0                   # tie → even mantissa 0
    assert OC.round_bf16(1.0 + 1.5 * ulp) == 1.0 + 2 * ulp       # tie → even mantissa 2
    assert OC.round_bf16(1.0 + 0.51 * ulp) == 1.0 + ulp


def test_adam_matches_torch_optim():
    rng = np.random.Generator(np.random.PCG64(5))
    p = rng.standard_normal(50)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 6):
        g = rng.standard_normal(50)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = OC.adam(p, m, v, g, step)
    np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_gradient_model_and_pi_closed_forms():
    # For a linear velocity (ū, v̄) = (a x + b y, c x + d y), τ_gm = σ²·G Gᵀ with G = [[a, b], [c, d]],
    # and Π = −S̄:τ with S̄ the symmetric part of G.
    a, b, c, d, s2 = 0.3, -1.2, 0.7, 0.5, 2.5
    grad4 = np.array([a, b, c, d]).reshape(4, 1, 1)
    G = np.array([[a, b], [c, d]])
    T = s2 * G @ G.T
    tau = OC.gradient_model(grad4, s2)
    np.testing.assert_allclose(tau[:, 0, 0], [T[0, 0], T[0, 1], T[1, 1]], rtol=1e-14)
    S = 0.5 * (G + G.T)
    pi = OC.closure_pi(grad4, tau)
    assert abs(pi[0, 0] + np.sum(S * T)) <= 1e-13


@pytest.mark.parametrize("bf16", [False, True])
def test_bf16_forward_stays_close_to_fp64(bf16):
    rng = np.random.Generator(np.random.PCG64(6))
    params = _params(7)
    x = rng.standard_normal((1, CH[0], 20, 20))
    y64, _, _ = OC.forward(params, x, bf16=False)
    y, _, _ = OC.forward(params, x, bf16=bf16)
    rel = np.linalg.norm(y - y64) / np.linalg.norm(y64)
    assert rel <= (2e-2 if bf16 else 0.0)


def test_tau_pi_loss_gradient_matches_central_differences():
    rng = np.random.Generator(np.random.PCG64(8))
    pred, target = rng.standard_normal((2, 3, 4, 5)), rng.standard_normal((2, 3, 4, 5))
    g = rng.standard_normal((2, 4, 4, 5))
    coef = rng.uniform(0.5, 2.0, 7)
    loss, grad = OC.tau_pi_loss(pred, target, g, coef, 3.0)
    # Π(τ) by its definition, independently: −S̄:τ with S̄ from the physical gradients, times the scales
    ux, uy, vx, vy = (g[:, k] * coef[k] for k in range(4))
    pi = lambda t: -(ux * t[:, 0] * coef[4] + (uy + vx) * t[:, 1] * coef[5] + vy * t[:, 2] * coef[6])
    ref = np.mean((pred - target) ** 2) + 3.0 * np.mean((pi(pred) - pi(target)) ** 2)
    assert abs(loss - ref) <= 1e-12 * ref
    h = 1e-6
    for idx in [(0, 0, 1, 2), (1, 1, 3, 4), (0, 2, 0, 0)]:
        p1, p2 = pred.copy(), pred.copy()
        p1[idx] += h
        p2[idx] -= h
        fd = (OC.tau_pi_loss(p1, target, g, coef, 3.0)[0] - OC.tau_pi_loss(p2, target, g, coef, 3.0)[0]) / (2 * h)
        assert abs(fd - grad[idx]) <= 1e-7 * max(1.0, abs(fd))

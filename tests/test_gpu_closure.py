"""GPU parity of the learned closure's kernels (libcgef_hip.so through cgef.closure) against the FP64
oracle (oracle/closure.py, pinned in test_closure_oracle.py), and its training / fallback behaviour.

Tolerances: the oracle rounds what the device stores in BF16 (input, weights, hidden
pre-activations), so the device differs from it by FP32 accumulation order, plus the rare BF16
rounding that lands on the other side of a tie-point: forward rel-L2 ≤ 5e-3, gradients ≤ 1e-2 (an
index or sign mistake gives O(1)). Exact-τ pairs and Π: FP32 storage of FP64 results, 1e-6."""
import numpy as np
import pytest
import torch

import cgef
from cgef import closure as CL
from cgef import synthetic as SY
from helpers import cart_axes, oracle_grid_cart, rel_l2
from oracle import closure as OC
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def N(t):
    torch.cuda.synchronize()
    return t.detach().float().cpu().numpy().astype(np.float64)


def _net_and_input(channels, n, h, w, seed, mfma=True):
    net = CL.ClosureNet(channels, seed=seed, device=0, mfma=mfma)
    # non-zero biases so the bias path is exercised
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    Ws, bs = net.state()
    bs = [0.1 * rng.standard_normal(b.shape) for b in bs]
    net.load_state(Ws, bs)
    x = rng.standard_normal((n, channels[0], h, w))
    xb = torch.tensor(x, dtype=torch.float32, device=DEV).to(torch.bfloat16)
    params = [(Ws[l], bs[l]) for l in range(len(Ws))]
    return net, xb, x, params


# 40 × 37 stages the matrix-core layers' inputs element by element (odd width); 44 × 48 keeps every
# layer's width a multiple of 4, so they stage 4-element groups, edge tiles partly outside the input
CASES = [((4, 64, 64, 64, 3), 2, 40, 37, True), ((4, 64, 64, 64, 3), 2, 44, 48, True),
         ((4, 64, 64, 64, 3), 2, 40, 37, False), ((4, 8, 20, 3), 3, 23, 50, True)]


@pytest.mark.parametrize("channels,n,h,w,mfma", CASES)
def test_conv_forward_matches_oracle(channels, n, h, w, mfma):
    print(f"closure forward {channels} n={n} {h}x{w} mfma={mfma}", flush=True)
    net, xb, x, params = _net_and_input(channels, n, h, w, 1, mfma)
    y = net.forward(xb)
    yo, zs, _ = OC.forward(params, x, bf16=True)
    assert rel_l2(N(y), yo) <= 5e-3
    for l, z in enumerate(zs):
        assert rel_l2(N(net.Z[l]), z) <= 5e-3, l


@pytest.mark.parametrize("channels,n,h,w,mfma", CASES)
def test_conv_backward_matches_oracle(channels, n, h, w, mfma):
    print(f"closure backward {channels} n={n} {h}x{w} mfma={mfma}", flush=True)
    net, xb, x, params = _net_and_input(channels, n, h, w, 2, mfma)
    L = len(channels) - 1
    rng = np.random.Generator(np.random.PCG64(9))
    target = rng.standard_normal((n, channels[-1], h - 4 * L, w - 4 * L))
    net.loss_sum.zero_()
    net.forward(xb)
    net.backward(torch.tensor(target, dtype=torch.float32, device=DEV))
    loss, grads = OC.backward(params, x, target, bf16=True)
    assert abs(float(net.loss_sum[0].item()) / target.size - loss) <= 5e-3 * loss
    for l in range(L):
        assert rel_l2(N(net.gW[l]), grads[l][0]) <= 1e-2, l
        assert rel_l2(N(net.gb[l]), grads[l][1]) <= 1e-2, l


def test_adam_and_bf16_layouts():
    print("closure adam", flush=True)
    channels = (4, 8, 20, 3)
    net, xb, x, params = _net_and_input(channels, 1, 20, 20, 3)
    rng = np.random.Generator(np.random.PCG64(4))
    Ws, bs = net.state()
    ms = [(np.zeros_like(w), np.zeros_like(w), np.zeros_like(b), np.zeros_like(b)) for w, b in zip(Ws, bs)]
    for step in range(1, 4):
        gs = [(rng.standard_normal(w.shape), rng.standard_normal(b.shape)) for w, b in zip(Ws, bs)]
        for l, (gw, gb) in enumerate(gs):
            net.gW[l].copy_(torch.tensor(gw, dtype=torch.float32))
            net.gb[l].copy_(torch.tensor(gb, dtype=torch.float32))
            gw32, gb32 = gw.astype(np.float32).astype(np.float64), gb.astype(np.float32).astype(np.float64)
            mw, vw, mb, vb = ms[l]
            Ws[l], mw, vw = OC.adam(Ws[l], mw, vw, gw32, step)
            bs[l], mb, vb = OC.adam(bs[l], mb, vb, gb32, step)
            ms[l] = (mw, vw, mb, vb)
        net.step()
    got_W, got_b = net.state()
    for l in range(len(Ws)):
        assert np.max(np.abs(got_W[l] - Ws[l])) <= 1e-6 * (1 + np.max(np.abs(Ws[l])))
        assert np.max(np.abs(got_b[l] - bs[l])) <= 1e-6
        ci, co = channels[l], channels[l + 1]
        wb16 = torch.tensor(got_W[l], dtype=torch.float32).to(torch.bfloat16).float().numpy().reshape(co, ci, 25)
        wf = N(net.wf[l]).reshape(ci, 25, net.ocs_f[l])
        wbk = N(net.wb[l]).reshape(co, 25, net.ocs_b[l])
        np.testing.assert_array_equal(wf[:, :, :co], wb16.transpose(1, 2, 0))
        np.testing.assert_array_equal(wbk[:, :, :ci], wb16[:, :, ::-1].transpose(0, 2, 1))
        assert not wf[:, :, co:].any() and not wbk[:, :, ci:].any()


def test_crop_is_a_periodic_gather():
    print("closure crop", flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    src = rng.standard_normal((3, 2, 17, 23)).astype(np.float32)
    org = np.array([[2, -5, 20], [0, 16, -30], [1, 3, 4]], dtype=np.int32)
    scale = np.array([2.0, -0.5], dtype=np.float32)
    S = torch.tensor(src, device=DEV)
    O_ = torch.tensor(org, device=DEV)
    sc = torch.tensor(scale, device=DEV)
    for dt in (torch.float32, torch.bfloat16):
        dst = torch.empty((3, 2, 9, 31), dtype=dt, device=DEV)
        CL.crop_(dst, S, O_, sc)
        ref = np.empty((3, 2, 9, 31), dtype=np.float32)
        for i, (img, y0, x0) in enumerate(org):
            ys = (y0 + np.arange(9)) % 17
            xs = (x0 + np.arange(31)) % 23
            ref[i] = src[img][:, ys][:, :, xs] * scale[:, None, None]
        if dt == torch.bfloat16:
            ref = torch.tensor(ref).to(torch.bfloat16).float().numpy()
        np.testing.assert_array_equal(N(dst), ref.astype(np.float64))


def _periodic_case(n=64, ell_cells=4.0, seed=0):
    u, v, xs = SY.cartesian_turbulence(N=n, L=1.0e6, seed=seed, kmax=n // 3)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    dx = xs[1] - xs[0]
    plan = cgef.plan_filter(grid, cgef.GaussianKernel(6.0), ell_cells * dx)
    og = oracle_grid_cart(xs, xs, periodic=(True, True))
    return u, v, xs, grid, plan, og, ell_cells * dx


def test_exact_pairs_match_oracle_filter():
    print("closure exact pairs", flush=True)
    u, v, xs, grid, plan, og, ell = _periodic_case()
    assert plan.engine == "separable"
    feat, tau = CL.closure_features(torch.tensor(u, device=DEV), torch.tensor(v, device=DEV), grid, plan)
    f = O.filter_fields(og, [u, v, u * u, u * v, v * v], kind=O.GAUSSIAN, alpha=6.0, ell=ell,
                        engine=O.ENGINE_SEPARABLE)
    ub, vb = f[0], f[1]
    tau_ref = np.stack([f[2] - ub * ub, f[3] - ub * vb, f[4] - vb * vb])
    grad_ref = np.stack([O.deriv(og, 0, ub), O.deriv(og, 1, ub), O.deriv(og, 0, vb), O.deriv(og, 1, vb)])
    assert rel_l2(N(tau[0]), tau_ref) <= 1e-6
    assert rel_l2(N(feat[0]), grad_ref) <= 1e-6
    feat2, none = CL.closure_features(torch.tensor(u, device=DEV), torch.tensor(v, device=DEV), grid, plan, tau=False)
    assert none is None
    np.testing.assert_array_equal(N(feat2), N(feat))
    # Π from the exact pairs is compute_Π!'s Π
    Pi = torch.empty((64, 64), dtype=torch.float64, device=DEV)
    cgef.compute_pi_(Pi, torch.tensor(u, device=DEV), torch.tensor(v, device=DEV), None, grid, plan.kernel, plan.scale,
                     filter_plan=plan)
    assert rel_l2(N(CL.pi_from_tau(feat, tau)[0]), N(Pi)) <= 1e-5
    # gradient model and Π kernels against the oracle's closed forms
    fn = N(feat[0])
    gm = N(CL.gradient_model_tau(feat, ell * ell / 12.0)[0])
    assert rel_l2(gm, OC.gradient_model(fn, ell * ell / 12.0)) <= 1e-6
    sc = torch.tensor([2.0, 3.0, 0.5], device=DEV)
    scaled = N(tau[0]) * np.array([2.0, 3.0, 0.5])[:, None, None]
    assert rel_l2(N(CL.pi_from_tau(feat, tau, sc)[0]), OC.closure_pi(fn, scaled)) <= 1e-6
    # the residual the network learns, and Π of residual + gradient model
    s2 = ell * ell / 12.0
    res = N(CL.gradient_model_tau(feat, -s2, base=tau)[0])
    assert rel_l2(res, N(tau[0]) - OC.gradient_model(fn, s2)) <= 1e-5
    assert rel_l2(N(CL.pi_from_tau(feat, tau, sc, s2)[0]),
                  OC.closure_pi(fn, scaled + OC.gradient_model(fn, s2))) <= 1e-5


def test_training_learns_tau_and_fallback_uses_exact_pi():
    """A short run at small size: the loss falls, the held-out Π of the learned closure (gradient
    model + learned residual) is closer to the exact Π than the gradient model's, and LearnedClosure
    returns the exact Π when the measured error is above its threshold."""
    print("closure training", flush=True)
    seeds = list(range(10))
    feat, tau, grid, plan = CL.closure_dataset(seeds, N=128, ell_cells=8.0, device=0)
    s2 = CL.gradient_model_sigma2(plan.kernel, plan.scale)
    assert s2 == pytest.approx(plan.scale ** 2 / 12.0)
    model = CL.fit_closure(feat[:8], tau[:8], s2, steps=300, batch=8, crop=32, log_every=50, seed=0)
    hist = model.history
    assert hist[-1][1] < 0.5 * hist[0][1], hist
    fte, tte = feat[8:].contiguous(), tau[8:].contiguous()
    exact = CL.pi_from_tau(fte, tte)
    corr, rel = CL.pi_errors(model.predict_pi(fte), exact)
    gcorr, grel = CL.pi_errors(CL.pi_from_tau(fte, CL.gradient_model_tau(fte, s2)), exact)
    print(f"held-out Π: learned corr {corr:.3f} rel-L2 {rel:.3f}; gradient model corr {gcorr:.3f} rel-L2 {grel:.3f}",
          flush=True)
    assert corr > 0.9 and rel < 0.5 and rel < grel

    u, v, _ = SY.cartesian_turbulence(N=128, L=1.0e6, seed=20)
    ut, vt = torch.tensor(u, device=DEV), torch.tensor(v, device=DEV)
    exact_pi = torch.empty((128, 128), dtype=torch.float32, device=DEV)
    cgef.compute_pi_(exact_pi, ut.float(), vt.float(), None, grid, plan.kernel, plan.scale, filter_plan=plan)
    strict = CL.LearnedClosure(model, grid, plan, threshold=1e-9)
    out = torch.empty_like(exact_pi)
    _, used_exact = strict.compute_pi_(out, ut, vt)
    assert used_exact and strict.disabled
    np.testing.assert_array_equal(N(out), N(exact_pi))
    _, used_exact = strict.compute_pi_(out, ut, vt)
    assert used_exact and strict.fallbacks == 2
    lax = CL.LearnedClosure(model, grid, plan, threshold=10.0)
    _, used_exact = lax.compute_pi_(out, ut, vt)
    assert not used_exact and lax.last_error is not None
    assert rel_l2(N(out), N(exact_pi)) == pytest.approx(lax.last_error, rel=1e-3)


def test_tau_pi_loss_gradient_matches_oracle():
    print("closure Π loss", flush=True)
    rng = np.random.Generator(np.random.PCG64(12))
    pred, target = rng.standard_normal((2, 3, 9, 13)), rng.standard_normal((2, 3, 9, 13))
    g = rng.standard_normal((2, 4, 9, 13))
    coef = rng.uniform(0.5, 2.0, 7)
    dev = lambda a: torch.tensor(a, dtype=torch.float32, device=DEV)
    pred32, target32, g32, coef32 = (a.astype(np.float32).astype(np.float64) for a in (pred, target, g, coef))
    net = CL.ClosureNet((4, 4, 3), seed=0, device=0)   # any network: only its loss kernel is used here
    dpred = torch.empty((2, 3, 9, 13), dtype=torch.float32, device=DEV)
    loss = torch.zeros(2, dtype=torch.float32, device=DEV)
    tp, tt, tg, tc = dev(pred), dev(target), dev(g), dev(coef)    # held for the call's duration
    cgef._lib.check(CL._L().cgef_closure_loss_grad(CL._p(tp), CL._p(tt), CL._p(tg), CL._p(tc), 2.5, 2, 9 * 13,
                                                    CL._p(dpred), CL._p(loss), CL._stream(0)))
    ref_loss, ref_grad = OC.tau_pi_loss(pred32, target32, g32, coef32, 2.5)
    assert rel_l2(N(dpred), ref_grad) <= 1e-6
    ls = N(loss)
    assert abs(ls[0] / pred.size + 2.5 * ls[1] / (pred.size / 3) - ref_loss) <= 1e-5 * ref_loss
    del net


@pytest.mark.parametrize("mfma", [True, False])
def test_gelu_on_load_matches_stored_activation(mfma):
    """The kernels' GELU-while-staging input mode (in_gelu) against the stored BF16 activation path:
    they differ only by the BF16 rounding of GELU(Z)."""
    print(f"closure in_gelu mfma={mfma}", flush=True)
    net, xb, x, params = _net_and_input((4, 64, 64, 64, 3), 1, 36, 34, 5, mfma)
    net.forward(xb)
    z2 = torch.empty_like(net.Z[1])
    if mfma:
        CL.conv5_mfma(z2, net.Z[0], net.wmf[1], net.b[1], in_gelu=True)
    else:
        CL.conv5(z2, net.Z[0], net.wf[1], net.ocs_f[1], net.b[1], in_gelu=True)
    assert rel_l2(N(z2), N(net.Z[1])) <= 5e-3
    gw, gb = torch.empty_like(net.gW[1]), torch.empty_like(net.gb[1])
    dz = torch.tensor(np.random.Generator(np.random.PCG64(3)).standard_normal(tuple(net.Z[1].shape)),
                      dtype=torch.float32, device=DEV)
    CL.conv5_wgrad(gw, gb, dz, net.Z[0], net.scratch, in_gelu=True, mfma=mfma)
    gw2, gb2 = torch.empty_like(gw), torch.empty_like(gb)
    CL.conv5_wgrad(gw2, gb2, dz, net.A[0], net.scratch, mfma=mfma)
    assert rel_l2(N(gw), N(gw2)) <= 5e-3
    np.testing.assert_allclose(N(gb), N(gb2), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,h,w", [(2, 40, 37), (1, 748, 748)])
def test_output_layer_both_tilings_match_fp64_conv(n, h, w):
    """The 64 → 3 layer runs on 8×8 tiles with the input channels split over the waves when the batch
    has few 16×16 tiles (n·tiles < 2048: the first case) and on k_conv5's 16×16 tiles otherwise (the
    second: 1 × 47² tiles); both against torch's FP64 convolution of the same BF16 values."""
    print(f"closure 64->3 layer n={n} {h}x{w}", flush=True)
    net = CL.ClosureNet((64, 3), seed=7, device=0, mfma=False)
    rng = np.random.Generator(np.random.PCG64(70))
    Ws, _ = net.state()
    b = [0.1 * rng.standard_normal(3)]
    net.load_state(Ws, b)
    x = torch.tensor(rng.standard_normal((n, 64, h, w)), dtype=torch.float32, device=DEV).to(torch.bfloat16)
    out = torch.empty((n, 3, h - 4, w - 4), dtype=torch.float32, device=DEV)
    CL.conv5(out, x, net.wf[0], net.ocs_f[0], net.b[0])
    xd = x.float().cpu().double().numpy()
    Wd = torch.tensor(np.asarray(Ws[0]), dtype=torch.float32).to(torch.bfloat16).double().numpy().reshape(3, 64, 5, 5)
    ho, wo = h - 4, w - 4
    ref = np.zeros((n, 3, ho, wo)) + np.asarray(b[0]).reshape(1, 3, 1, 1)
    for ky in range(5):           # valid cross-correlation, one tap at a time (a channel contraction each)
        for kx in range(5):
            ref += np.einsum("nihw,oi->nohw", xd[:, :, ky:ky + ho, kx:kx + wo], Wd[:, :, ky, kx], optimize=True)
    err = rel_l2(N(out), ref)
    print(f"  rel-L2 vs FP64 conv {err:.2e}", flush=True)
    assert err <= 1e-4

"""FP64 NumPy restatement of the learned closure's network (forward, backward and Adam).

TEST INFRASTRUCTURE ONLY: the checker, never the product. Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU legs may import this module.

The network has no counterpart in the reference (which obtains τ only from the exact filter); its
targets are the exact τ of the filter restated in ``cgef_oracle.c``. This module states what the
device kernels compute, in FP64:

    Z_1 = conv5(X, W_1) + b_1,   Z_l = conv5(GELU(Z_{l-1}), W_l) + b_l  (l = 2..L),   Y = Z_L

with valid 5×5 cross-correlation conv5(A, W)[n, o, y, x] = Σ_{i,ky,kx} W[o, i, ky, kx]·A[n, i, y+ky, x+kx]
and GELU(z) = z·Φ(z) (exact normal CDF). ``bf16=True`` rounds what the device stores in BF16 (the
input, the weights, every hidden pre-activation Z_l and activation GELU(Z_l)) to nearest even, so the
device differs from it by FP32 accumulation order (and, on the matrix-core path, by the BF16
rounding of the data gradient it multiplies). Gradients are those of the loss mean((Y − T)²).

Pinned in tests/test_closure_oracle.py against torch's CPU FP64 autograd (torch.nn.functional.conv2d and
gelu, an independent library) and against central finite differences of the FP64 forward.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def round_bf16(a):
    """Round FP64/FP32 values to the nearest BF16 (ties to even), returned as FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def conv5_valid(a, w):
    """a (n, ci, h, w), w (co, ci, 5, 5) → (n, co, h−4, w−4)."""
    n, ci, h, wd = a.shape
    out = np.zeros((n, w.shape[0], h - 4, wd - 4))
    for ky in range(5):
        for kx in range(5):
            out += np.einsum("nihw,oi->nohw", a[:, :, ky:ky + h - 4, kx:kx + wd - 4], w[:, :, ky, kx], optimize=True)
    return out


def conv5_full_transposed(d, w):
    """Data gradient of conv5_valid: d (n, co, h, w), w (co, ci, 5, 5) → (n, ci, h+4, w+4),
    out[n, i, y, x] = Σ_{o,ky,kx} d[n, o, y−ky, x−kx]·w[o, i, ky, kx]."""
    n, co, h, wd = d.shape
    out = np.zeros((n, w.shape[1], h + 4, wd + 4))
    for ky in range(5):
        for kx in range(5):
            out[:, :, ky:ky + h, kx:kx + wd] += np.einsum("nohw,oi->nihw", d, w[:, :, ky, kx], optimize=True)
    return out


def conv5_weight_grad(d, a):
    """dW[o, i, ky, kx] = Σ_{n,y,x} d[n, o, y, x]·a[n, i, y+ky, x+kx]."""
    n, co, h, wd = d.shape
    dw = np.zeros((co, a.shape[1], 5, 5))
    for ky in range(5):
        for kx in range(5):
            dw[:, :, ky, kx] = np.einsum("nohw,nihw->oi", d, a[:, :, ky:ky + h, kx:kx + wd], optimize=True)
    return dw


def forward(params, x, bf16=True):
    """params: list of (W (co, ci, 5, 5), b (co,)) FP64; x (n, c0, h, w). Returns (Y, Zs, As): the
    output, the hidden pre-activations (as stored) and each layer's input."""
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = r(x)
    zs, ins = [], []
    L = len(params)
    for li, (w, b) in enumerate(params):
        ins.append(a)
        z = conv5_valid(a, r(w)) + b[None, :, None, None]
        if li == L - 1:
            return z, zs, ins
        z = r(z)
        zs.append(z)
        a = r(gelu(z))
    raise ValueError("empty network")


def backward(params, x, target, bf16=True):
    """Loss mean((Y − T)²) and its gradients [(dW, db)] per layer, with the forward as in `forward`."""
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, zs, ins = forward(params, x, bf16)
    diff = y - target
    loss = float(np.mean(diff ** 2))
    d = 2.0 * diff / diff.size
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        w = r(params[li][0])
        grads[li] = (conv5_weight_grad(d, ins[li]), d.sum(axis=(0, 2, 3)))
        if li > 0:
            d = conv5_full_transposed(d, w) * gelu_grad(zs[li - 1])
    return loss, grads


def adam(p, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015, Algorithm 1) with bias correction; returns the new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def gradient_model(grad4, sigma2):
    """τ_ab = σ² Σ_k ∂_k ū_a ∂_k ū_b from (∂xū, ∂yū, ∂xv̄, ∂yv̄) (…, 4, …) on axis −3."""
    ux, uy, vx, vy = (grad4[..., c, :, :] for c in range(4))
    return sigma2 * np.stack([ux * ux + uy * uy, ux * vx + uy * vy, vx * vx + vy * vy], axis=-3)


def closure_pi(grad4, tau3):
    """Π = −(S̄xx τxx + 2 S̄xy τxy + S̄yy τyy) with S̄xy = ½(∂yū + ∂xv̄) (Diagnostics.jl:460-466)."""
    ux, uy, vx, vy = (grad4[..., c, :, :] for c in range(4))
    return -(ux * tau3[..., 0, :, :] + (uy + vx) * tau3[..., 1, :, :] + vy * tau3[..., 2, :, :])


def tau_pi_loss(pred, target, g, coef, lam):
    """Loss mean((pred − target)²) + lam·mean((Π(pred) − Π(target))²) and its gradient in pred, with
    Π(τ) = −(a0 τxx + a1 τxy + a2 τyy), a0 = g0·c[0]·c[4], a1 = (g1·c[1] + g2·c[2])·c[5],
    a2 = g3·c[3]·c[6]; pred, target (n, 3, h, w), g (n, 4, h, w)."""
    a = np.stack([g[:, 0] * coef[0] * coef[4], (g[:, 1] * coef[1] + g[:, 2] * coef[2]) * coef[5],
                  g[:, 3] * coef[3] * coef[6]], axis=1)
    e = pred - target
    dpi = -np.sum(a * e, axis=1)
    loss = float(np.mean(e ** 2) + lam * np.mean(dpi ** 2))
    grad = 2.0 * e / e.size - 2.0 * lam * (dpi / dpi.size)[:, None] * a
    return loss, grad

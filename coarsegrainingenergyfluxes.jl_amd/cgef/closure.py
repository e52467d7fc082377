"""Learned closure for the subfilter stress τ_ℓ (no counterpart in the reference, which computes τ
only by the exact filter, Diagnostics.jl:496-527).

A 4-layer 5×5 convolutional network maps the filtered-velocity gradients (∂xū, ∂yū, ∂xv̄, ∂yv̄)
to what the gradient model τ ≈ σ² G Gᵀ (G = ∇ū, σ² the footprint's second moment) misses of
τ = (τxx, τxy, τyy): in 2-D incompressible flow the gradient model's Π vanishes identically
(tr S̄ = 0 makes −S̄:G Gᵀ = 0), so Π is carried by that residual. It is trained on the device by this library's own kernels (libcgef_hip.so:
cgef_conv5 forward and data gradient, cgef_conv5_wgrad, cgef_adam_update, cgef_mse_grad) on pairs
that the exact filter itself produces (closure_dataset), with BF16 weights and activations and FP32
accumulation. The exact Π remains the check and the fallback: LearnedClosure measures the network's
Π against the exact Π and falls back to the exact pipeline when the relative L2 error exceeds a
threshold.

Tensors are torch device tensors; torch provides memory only. Every compute step is a HIP kernel.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _lib
from .diagnostics import compute_pi_
from .derivatives import ddx_, ddy_
from .filtering import _stream, filter_apply_, plan_filter
from .grids import StructuredGrid
from .kernels import GaussianKernel

BF16 = 2
HALO = 2          # each 5×5 valid layer trims 2 cells per side


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _ocb(c):
    return 4 if c <= 4 else (16 if c <= 16 else 64)


def _round_up(a, b):
    return (a + b - 1) // b * b


_sig_done = False


def _L():
    global _sig_done
    L = _lib.load()
    if not _sig_done:
        vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sigs = {
            "cgef_conv5": [vp, i32, i32, i64, i64, i64, i64, vp, i64, vp, i64, i32, vp, i32, vp, vp, vp],
            "cgef_conv5_mfma": [vp, i32, i32, i64, i64, i64, vp, vp, i32, vp, i32, vp, vp, vp],
            "cgef_conv5_pack_mfma": [vp, vp, vp, vp],
            "cgef_conv5_mfma_ex": [vp, i32, i32, i64, i64, i64, i64, vp, vp, i64, i32, vp, i32, vp, vp, vp],
            "cgef_conv5_pack_mfma_ex": [vp, i64, i64, vp, vp, vp],
            "cgef_conv5_wgrad_scratch": [i64, i64, i64, i64, i64, i32, C.POINTER(i64)],
            "cgef_conv5_wgrad": [vp, vp, i32, i32, i64, i64, i64, i64, i64, i32, vp, vp, vp, i64, vp],
            "cgef_conv5_pack_weights": [vp, i64, i64, vp, i64, vp, i64, vp],
            "cgef_adam_update": [vp, vp, vp, vp, i64, dbl, dbl, dbl, dbl, i64, i64, i64, vp, i64, vp, i64, vp],
            "cgef_mse_grad": [vp, vp, i64, vp, vp, vp],
            "cgef_closure_loss_grad": [vp, vp, vp, vp, dbl, i64, i64, vp, vp, vp],
            "cgef_closure_crop": [vp, i32, i64, i64, i64, i64, vp, i64, i64, i64, vp, vp, i32, vp],
            "cgef_closure_products": [vp, vp, i32, i64, i64, vp, vp],
            "cgef_closure_targets": [vp, vp, i64, i64, vp, vp, vp],
            "cgef_closure_pi": [vp, vp, vp, dbl, i64, i64, vp, vp],
            "cgef_closure_gradient_model": [vp, dbl, vp, i64, i64, vp, vp],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _sig_done = True
    return L


def _check_dev(t, name, dtype=None):
    if not t.is_cuda:
        raise _lib.ArgumentError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise _lib.ArgumentError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise _lib.ArgumentError(f"{name} must be {dtype}")


# ------------------------------------------------------------------------------------------------
# Layer-level entry points (thin wrappers over the C ABI)
# ------------------------------------------------------------------------------------------------

def conv5(out, x, wf, ocs, bias, *, in_gelu=False, pad=0, zgrad=None, act=None):
    """One 5×5 layer into `out` (see cgef_conv5). x BF16 or FP32 (n, cin, h, w); wf BF16 forward
    layout [cin][25][ocs] (or the backward-data layout for pad=4)."""
    n, cin, h, w = x.shape
    cout = out.shape[1]
    ho, wo = h + 2 * pad - 4, w + 2 * pad - 4
    if tuple(out.shape) != (n, cout, ho, wo):
        raise _lib.DimensionMismatch(f"out is {tuple(out.shape)}, expected {(n, cout, ho, wo)}")
    if wf.numel() != cin * 25 * ocs:
        raise _lib.DimensionMismatch("weight layout does not match cin / ocs")
    if zgrad is not None and tuple(zgrad.shape) != tuple(out.shape):
        raise _lib.DimensionMismatch("zgrad must be shaped as out")
    for t, nm in ((out, "out"), (x, "x"), (wf, "weights")):
        _check_dev(t, nm)
    in_dt = BF16 if x.dtype == torch.bfloat16 else _lib.F32
    out_dt = BF16 if out.dtype == torch.bfloat16 else _lib.F32
    if act is not None and (tuple(act.shape) != tuple(out.shape) or act.dtype != torch.bfloat16):
        raise _lib.DimensionMismatch("act must be a BF16 tensor shaped as out")
    _lib.check(_L().cgef_conv5(_p(x), in_dt, int(in_gelu), n, cin, h, w, _p(wf), ocs, _p(bias), cout, pad, _p(out),
                               out_dt, _p(zgrad), _p(act), _stream(x.device.index)))
    return out


def mfma_pads(cin, cout):
    """(input-channel pad, output-channel pad) of a matrix-core layer: 32 or 64, 16 or 64."""
    return (32 if cin <= 32 else 64), (16 if cout <= 16 else 64)


def mfma_layout_size(cin, cout):
    """Entries of the forward matrix-core layout of a cin → cout layer (cgef_conv5_pack_mfma_ex)."""
    ci, co = mfma_pads(cin, cout)
    return 25 * ci * co


def conv5_mfma(out, x, wm, bias, *, in_gelu=False, pad=0, zgrad=None, act=None):
    """A 5×5 layer of 1 to 64 input and output channels on the matrix cores (cgef_conv5_mfma_ex); wm
    is the forward layout (pad 0) or the backward-data layout (pad 4) of cgef_conv5_pack_mfma_ex."""
    n, cin, h, w = x.shape
    cout = out.shape[1]
    ho, wo = h + 2 * pad - 4, w + 2 * pad - 4
    if not (1 <= cin <= 64 and 1 <= cout <= 64) or tuple(out.shape) != (n, cout, ho, wo):
        raise _lib.DimensionMismatch(f"matrix-core layer takes 1 to 64 channels in and out; got x {tuple(x.shape)}, "
                                     f"out {tuple(out.shape)}")
    if wm.numel() != mfma_layout_size(cin, cout):
        raise _lib.DimensionMismatch(f"matrix-core weight layout has {mfma_layout_size(cin, cout)} entries")
    if zgrad is not None and tuple(zgrad.shape) != tuple(out.shape):
        raise _lib.DimensionMismatch("zgrad must be shaped as out")
    for t, nm in ((out, "out"), (x, "x"), (wm, "weights")):
        _check_dev(t, nm)
    in_dt = BF16 if x.dtype == torch.bfloat16 else _lib.F32
    out_dt = BF16 if out.dtype == torch.bfloat16 else _lib.F32
    if act is not None and (tuple(act.shape) != tuple(out.shape) or act.dtype != torch.bfloat16):
        raise _lib.DimensionMismatch("act must be a BF16 tensor shaped as out")
    _lib.check(_L().cgef_conv5_mfma_ex(_p(x), in_dt, int(in_gelu), n, cin, h, w, _p(wm), _p(bias), cout, pad, _p(out),
                                       out_dt, _p(zgrad), _p(act), _stream(x.device.index)))
    return out


def conv5_wgrad(dw, db, dz, a, scratch, *, in_gelu=False, mfma=False):
    n, cout, ho, wo = dz.shape
    cin = a.shape[1]
    if tuple(a.shape) != (n, cin, ho + 4, wo + 4):
        raise _lib.DimensionMismatch("layer input must be (n, cin, hout+4, wout+4)")
    in_dt = BF16 if a.dtype == torch.bfloat16 else _lib.F32
    _lib.check(_L().cgef_conv5_wgrad(_p(dz), _p(a), in_dt, int(in_gelu), n, cin, cout, ho, wo, int(mfma), _p(dw),
                                     _p(db), _p(scratch), scratch.numel(), _stream(dz.device.index)))
    return dw, db


def wgrad_scratch_floats(n, cin, cout, ho, wo, mfma=False):
    out = C.c_int64()
    _lib.check(_L().cgef_conv5_wgrad_scratch(n, cin, cout, ho, wo, int(mfma), C.byref(out)))
    return int(out.value)


# ------------------------------------------------------------------------------------------------
# The network
# ------------------------------------------------------------------------------------------------

class ClosureNet:
    """conv5 → GELU → conv5 → GELU → conv5 → GELU → conv5 with `channels` = (4, 64, 64, 64, 3):
    FP32 master weights and Adam moments, BF16 weight layouts and activations, FP32 accumulation.
    The input must carry HALO·L cells of context on each side of the predicted window. With
    mfma=True every layer of at most 64 channels in and out runs its forward and data gradient on the
    matrix cores (channels padded to the MFMA tiles), and the 64 → 64 layers their weight gradients."""

    def __init__(self, channels=(4, 64, 64, 64, 3), *, seed=0, device=None, lr=1e-3, betas=(0.9, 0.999),
                 eps=1e-8, mfma=True):
        self.device = torch.cuda.current_device() if device is None else device
        self.mfma = bool(mfma)
        self.channels = tuple(int(c) for c in channels)
        self.L = len(self.channels) - 1
        self.lr, self.betas, self.eps = float(lr), (float(betas[0]), float(betas[1])), float(eps)
        self.t = 0
        rng = np.random.Generator(np.random.PCG64(seed))
        dev = f"cuda:{self.device}"
        self.W, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], []
        self.wf, self.wb, self.ocs_f, self.ocs_b, self.wmf, self.wmb, self.wgmm = [], [], [], [], [], [], []
        for l in range(self.L):
            ci, co = self.channels[l], self.channels[l + 1]
            gain = math.sqrt(2.0) if l < self.L - 1 else 1.0          # He init before GELU
            w = rng.standard_normal((co, ci, 5, 5)) * (gain / math.sqrt(ci * 25))
            self.W.append(torch.tensor(w, dtype=torch.float32, device=dev))
            self.b.append(torch.zeros(co, dtype=torch.float32, device=dev))
            for lst in (self.mW, self.vW):
                lst.append(torch.zeros_like(self.W[-1]))
            for lst in (self.mb, self.vb):
                lst.append(torch.zeros_like(self.b[-1]))
            of, ob = _round_up(co, _ocb(co)), _round_up(ci, _ocb(ci))
            self.ocs_f.append(of)
            self.ocs_b.append(ob)
            self.wf.append(torch.zeros(ci * 25 * of, dtype=torch.bfloat16, device=dev))
            self.wb.append(torch.zeros(co * 25 * ob, dtype=torch.bfloat16, device=dev))
            mm = self.mfma and ci <= 64 and co <= 64
            self.wmf.append(torch.zeros(mfma_layout_size(ci, co), dtype=torch.bfloat16, device=dev) if mm else None)
            # the data gradient through layer l (l ≥ 1) is a co → ci layer; from at most 4 channels (the
            # 64 → 3 output layer) the VALU kernel is faster: one K slice per tap would be ≥ 7/8 padding
            self.wmb.append(torch.zeros(mfma_layout_size(co, ci), dtype=torch.bfloat16, device=dev)
                            if mm and l > 0 and co > 4 else None)
            self.wgmm.append(self.mfma and ci == 64 and co == 64)   # matrix-core weight gradient
        self.pack()
        self._shape = None
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.loss_sum = torch.zeros(2, dtype=torch.float32, device=dev)    # (Σ τ part, Σ Π part)

    @property
    def halo(self):
        return HALO * self.L

    def pack(self):
        """Write the BF16 layouts from the FP32 master weights."""
        for l in range(self.L):
            co, ci = self.channels[l + 1], self.channels[l]
            _lib.check(_L().cgef_conv5_pack_weights(_p(self.W[l]), co, ci, _p(self.wf[l]), self.ocs_f[l],
                                                    _p(self.wb[l]), self.ocs_b[l], _stream(self.device)))
            self._pack_mfma(l)

    def _pack_mfma(self, l):
        if self.wmf[l] is not None:
            co, ci = self.channels[l + 1], self.channels[l]
            _lib.check(_L().cgef_conv5_pack_mfma_ex(_p(self.W[l]), co, ci, _p(self.wmf[l]), _p(self.wmb[l]),
                                                    _stream(self.device)))

    def load_state(self, weights, biases):
        for l in range(self.L):
            self.W[l].copy_(torch.as_tensor(np.asarray(weights[l]), dtype=torch.float32))
            self.b[l].copy_(torch.as_tensor(np.asarray(biases[l]), dtype=torch.float32))
        self.pack()

    def state(self):
        return ([w.detach().cpu().double().numpy() for w in self.W], [b.detach().cpu().double().numpy() for b in self.b])

    def _buffers(self, n, h, w):
        if self._shape == (n, h, w):
            return
        dev = f"cuda:{self.device}"
        self.Z, self.A, self.dZ = [], [], []
        for l in range(self.L):
            co = self.channels[l + 1]
            ho, wo = h - 4 * (l + 1), w - 4 * (l + 1)
            if ho < 1 or wo < 1:
                raise _lib.DimensionMismatch(f"input {h}x{w} is too small for {self.L} valid 5x5 layers")
            last = l == self.L - 1
            self.Z.append(torch.empty((n, co, ho, wo), dtype=torch.float32 if last else torch.bfloat16, device=dev))
            self.A.append(None if last else torch.empty((n, co, ho, wo), dtype=torch.bfloat16, device=dev))
            self.dZ.append(torch.empty((n, co, ho, wo), dtype=torch.float32, device=dev))
        need = 0
        for l in range(self.L):
            ho, wo = h - 4 * (l + 1), w - 4 * (l + 1)
            need = max(need, wgrad_scratch_floats(n, self.channels[l], self.channels[l + 1], ho, wo,
                                                  self.wgmm[l]))
        self.scratch = torch.empty(need, dtype=torch.float32, device=dev)
        self._shape = (n, h, w)

    def forward(self, x):
        """x BF16 (n, c0, h, w) → FP32 (n, cL, h − 4L, w − 4L) (a buffer reused by the next call)."""
        _check_dev(x, "x", torch.bfloat16)
        n, c0, h, w = x.shape
        if c0 != self.channels[0]:
            raise _lib.DimensionMismatch(f"x has {c0} channels, the network takes {self.channels[0]}")
        self._buffers(n, h, w)
        self.x = x
        a = x
        for l in range(self.L):
            if self.wmf[l] is not None:
                conv5_mfma(self.Z[l], a, self.wmf[l], self.b[l], pad=0, act=self.A[l])
            else:
                conv5(self.Z[l], a, self.wf[l], self.ocs_f[l], self.b[l], pad=0, act=self.A[l])
            a = self.A[l]
        return self.Z[-1]

    def backward(self, target, *, accumulate_loss=True, feat=None, coef=None, lambda_pi=0.0):
        """Gradients of mean((Y − target)²) after forward(); fills gW, gb. With `feat` (the normalized
        gradients at the output pixels, (n, 4, h, w)) and `coef` (cgef_closure_loss_grad) the loss
        adds lambda_pi · mean((Π(Y) − Π(target))²)."""
        y = self.Z[-1]
        if tuple(target.shape) != tuple(y.shape):
            raise _lib.DimensionMismatch(f"target is {tuple(target.shape)}, output is {tuple(y.shape)}")
        _check_dev(target, "target", torch.float32)
        ls = _p(self.loss_sum) if accumulate_loss else None
        if feat is not None:
            n, c, h, w = y.shape
            if c != 3 or tuple(feat.shape) != (n, 4, h, w) or coef is None or coef.numel() != 7:
                raise _lib.DimensionMismatch("the Π loss needs 3 output channels, feat (n, 4, h, w) and 7 coefficients")
            _check_dev(feat, "feat", torch.float32)
            _lib.check(_L().cgef_closure_loss_grad(_p(y), _p(target), _p(feat), _p(coef), float(lambda_pi), n, h * w,
                                                   _p(self.dZ[-1]), ls, _stream(self.device)))
        else:
            _lib.check(_L().cgef_mse_grad(_p(y), _p(target), y.numel(), _p(self.dZ[-1]), ls, _stream(self.device)))
        self.backward_from(self.dZ[-1])

    def backward_from(self, dy):
        """Backpropagate dL/dY (FP32, shaped as the output) through the layers."""
        if dy.data_ptr() != self.dZ[-1].data_ptr():
            self.dZ[-1].copy_(dy)
        for l in range(self.L - 1, -1, -1):
            a = self.x if l == 0 else self.A[l - 1]
            conv5_wgrad(self.gW[l], self.gb[l], self.dZ[l], a, self.scratch, mfma=self.wgmm[l])
            if l > 0 and self.wmb[l] is not None:
                conv5_mfma(self.dZ[l - 1], self.dZ[l], self.wmb[l], None, pad=4, zgrad=self.Z[l - 1])
            elif l > 0:
                conv5(self.dZ[l - 1], self.dZ[l], self.wb[l], self.ocs_b[l], None, pad=4, zgrad=self.Z[l - 1])

    def step(self):
        self.t += 1
        b1, b2 = self.betas
        for l in range(self.L):
            co, ci = self.channels[l + 1], self.channels[l]
            _lib.check(_L().cgef_adam_update(_p(self.W[l]), _p(self.mW[l]), _p(self.vW[l]), _p(self.gW[l]),
                                             self.W[l].numel(), self.lr, b1, b2, self.eps, self.t, co, ci,
                                             _p(self.wf[l]), self.ocs_f[l], _p(self.wb[l]), self.ocs_b[l],
                                             _stream(self.device)))
            self._pack_mfma(l)
            _lib.check(_L().cgef_adam_update(_p(self.b[l]), _p(self.mb[l]), _p(self.vb[l]), _p(self.gb[l]),
                                             co, self.lr, b1, b2, self.eps, self.t, co, 1, None, 0, None, 0,
                                             _stream(self.device)))

    def train_step(self, x, target, **loss):
        self.forward(x)
        self.backward(target, **loss)
        self.step()

    def flops_per_sample(self, h, w):
        """Multiply-adds × 2 of one forward over an (h, w) input (backward: twice this)."""
        f = 0
        for l in range(self.L):
            ho, wo = h - 4 * (l + 1), w - 4 * (l + 1)
            f += 2 * ho * wo * self.channels[l] * self.channels[l + 1] * 25
        return f


# ------------------------------------------------------------------------------------------------
# Data: exact (features, τ) pairs from the filter
# ------------------------------------------------------------------------------------------------

def crop_(dst, src, origins, chan_scale=None):
    """Periodic crops of src (S, c, H, W) FP32 at origins (n, 3) int32 (img, y0, x0) into dst
    (n, c, h, w) FP32 or BF16, each channel times chan_scale[c]."""
    S, c, H, W = src.shape
    n, c2, h, w = dst.shape
    if c2 != c or tuple(origins.shape) != (n, 3):
        raise _lib.DimensionMismatch("crop: channel count or origins shape mismatch")
    _check_dev(src, "src", torch.float32)
    _check_dev(origins, "origins", torch.int32)
    dt = BF16 if dst.dtype == torch.bfloat16 else _lib.F32
    _lib.check(_L().cgef_closure_crop(_p(src), _lib.F32, S, c, H, W, _p(origins), n, h, w, _p(chan_scale), _p(dst), dt,
                                      _stream(src.device.index)))
    return dst


def closure_features(u, v, grid: StructuredGrid, plan, *, tau=True):
    """Exact pairs for one batch of velocity slices u, v (nb, Ny, Nx) (FP32 or FP64): the filter
    (`plan`) of u, v and, with tau=True, of u², uv, v²; the gradients of ū, v̄ by the reference's
    stencils (ddx!/ddy!). Returns feat (nb, 4, Ny, Nx) FP32 and tau (nb, 3, Ny, Nx) FP32 or None."""
    if u.dim() == 2:
        u, v = u[None], v[None]
    nb = u.shape[0]
    npts = u.shape[-1] * u.shape[-2]
    dev = u.device
    if u.dtype != v.dtype or u.shape != v.shape:
        raise _lib.ArgumentError("u and v must share shape and dtype")
    prod = torch.empty((nb, 5, *u.shape[-2:]), dtype=torch.float64, device=dev)
    _lib.check(_L().cgef_closure_products(_p(u.contiguous()), _p(v.contiguous()),
                                          _lib.F32 if u.dtype == torch.float32 else _lib.F64, npts, nb, _p(prod),
                                          _stream(dev.index)))
    if not tau:
        prod = prod[:, :2].contiguous()          # ū, v̄ only
    filt = torch.empty_like(prod)
    filter_apply_(filt, prod, plan)
    grad = torch.empty((nb, 4, *u.shape[-2:]), dtype=torch.float64, device=dev)
    for s in range(nb):
        ddx_(grad[s, 0], filt[s, 0], grid)
        ddy_(grad[s, 1], filt[s, 0], grid)
        ddx_(grad[s, 2], filt[s, 1], grid)
        ddy_(grad[s, 3], filt[s, 1], grid)
    feat = torch.empty((nb, 4, *u.shape[-2:]), dtype=torch.float32, device=dev)
    tau_out = torch.empty((nb, 3, *u.shape[-2:]), dtype=torch.float32, device=dev) if tau else None
    _lib.check(_L().cgef_closure_targets(_p(filt) if tau else None, _p(grad), npts, nb, _p(tau_out), _p(feat),
                                         _stream(dev.index)))
    return feat, tau_out


def closure_dataset(seeds, *, N=1024, L=1.0e6, alpha=6.0, ell_cells=8.0, device=None, generator=None):
    """(feat, tau, grid, plan) for the doubly periodic N² Cartesian box: one slice per seed of the
    random-phase k⁻³ streamfunction generator (cgef.synthetic.cartesian_turbulence), Gaussian α,
    ℓ = ell_cells·Δx, filtered by this library's default engine for that grid."""
    from . import synthetic as S
    from .grids import CartesianGeometry
    dev = torch.cuda.current_device() if device is None else device
    gen = generator or S.cartesian_turbulence
    _, _, xs = gen(N=N, L=L, seed=int(seeds[0]))
    grid = StructuredGrid(CartesianGeometry(), xs, xs, periodic=(True, True))
    dx = xs[1] - xs[0]
    plan = plan_filter(grid, GaussianKernel(alpha), ell_cells * dx)
    feat = torch.empty((len(seeds), 4, N, N), dtype=torch.float32, device=f"cuda:{dev}")
    tau = torch.empty((len(seeds), 3, N, N), dtype=torch.float32, device=f"cuda:{dev}")
    for i, s in enumerate(seeds):
        u, v, _ = gen(N=N, L=L, seed=int(s))
        ut = torch.tensor(u, dtype=torch.float64, device=f"cuda:{dev}")
        vt = torch.tensor(v, dtype=torch.float64, device=f"cuda:{dev}")
        f, t = closure_features(ut, vt, grid, plan)
        feat[i].copy_(f[0])
        tau[i].copy_(t[0])
    return feat, tau, grid, plan


def rms_per_channel(x):
    """√mean(x²) per channel of (S, c, H, W) (a data statistic for normalization)."""
    return torch.sqrt(torch.mean(x.double() ** 2, dim=(0, 2, 3))).float()


# ------------------------------------------------------------------------------------------------
# Training and evaluation
# ------------------------------------------------------------------------------------------------

def train_closure(feat, tau, *, steps=2000, batch=16, crop=64, lr=1e-3, seed=0, net=None, channels=(4, 64, 64, 64, 3),
                  lambda_pi=0.0, log_every=0, log=None):
    """Adam on the mean-squared error of the RMS-normalized τ over random periodic crops (batch ×
    crop² predicted cells, each with its 8-cell context), plus lambda_pi × the mean-squared error of
    the Π those τ give with the exact gradients, in units of the training set's RMS Π. Returns
    (net, feat_scale, tau_scale, history): feat_scale = 1/RMS of each input channel, tau_scale =
    RMS of each τ channel, history = (step, τ mse, Π mse, seconds)."""
    dev = feat.device
    S, c0, N, _ = feat.shape
    fs = (1.0 / rms_per_channel(feat)).to(dev)
    ts = rms_per_channel(tau).to(dev)
    inv_ts = (1.0 / ts).to(dev)
    coef = None
    if lambda_pi > 0:
        pi = pi_from_tau(feat, tau)
        ps = 1.0 / float(torch.sqrt(torch.mean(pi.double() ** 2)))
        coef = torch.cat([1.0 / fs, ts * ps]).float().contiguous()
        del pi
    if net is None:
        net = ClosureNet(channels, seed=seed, device=dev.index, lr=lr)
    halo = net.halo
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    img = rng.integers(0, S, size=(steps, batch))
    y0 = rng.integers(0, N, size=(steps, batch))
    x0 = rng.integers(0, N, size=(steps, batch))
    org_in = torch.tensor(np.stack([img, y0 - halo, x0 - halo], -1).astype(np.int32), device=dev)
    org_t = torch.tensor(np.stack([img, y0, x0], -1).astype(np.int32), device=dev)
    X = torch.empty((batch, c0, crop + 2 * halo, crop + 2 * halo), dtype=torch.bfloat16, device=dev)
    T = torch.empty((batch, tau.shape[1], crop, crop), dtype=torch.float32, device=dev)
    G = torch.empty((batch, c0, crop, crop), dtype=torch.float32, device=dev) if coef is not None else None
    hist = []
    net.loss_sum.zero_()
    t0 = time.perf_counter()
    since = 0
    for s in range(steps):
        crop_(X, feat, org_in[s], fs)
        crop_(T, tau, org_t[s], inv_ts)
        if G is not None:
            crop_(G, feat, org_t[s], fs)
            net.train_step(X, T, feat=G, coef=coef, lambda_pi=lambda_pi)
        else:
            net.train_step(X, T)
        since += 1
        if log_every and ((s + 1) % log_every == 0 or s == steps - 1):
            tot = net.loss_sum.tolist()
            hist.append((s + 1, tot[0] / (since * T.numel()), tot[1] / (since * T.numel() / 3),
                         time.perf_counter() - t0))
            net.loss_sum.zero_()
            since = 0
            if log:
                log(f"closure step {s + 1}: τ mse {hist[-1][1]:.4e}, Π mse {hist[-1][2]:.4e} ({hist[-1][3]:.1f} s)")
    return net, fs, ts, hist


def predict_tau(net, feat, feat_scale, *, chunk=2):
    """The network's τ over whole periodic fields feat (S, 4, N, M), in the normalized units it was
    trained in (pi_from_tau applies tau_scale); `chunk` slices per forward."""
    S, c0, N, M = feat.shape
    h = net.halo
    dev = feat.device
    out = torch.empty((S, net.channels[-1], N, M), dtype=torch.float32, device=dev)
    for s0 in range(0, S, chunk):
        n = min(chunk, S - s0)
        org = torch.tensor([[s0 + i, -h, -h] for i in range(n)], dtype=torch.int32, device=dev)
        X = torch.empty((n, c0, N + 2 * h, M + 2 * h), dtype=torch.bfloat16, device=dev)
        crop_(X, feat, org, feat_scale)
        out[s0:s0 + n].copy_(net.forward(X))
    return out


def pi_from_tau(feat, tau, tau_scale=None, sigma2=0.0):
    """Π = −S̄:τ per slice (FP32) from gradients (S, 4, …) and τ = tau·tau_scale + sigma2·G Gᵀ."""
    S = feat.shape[0]
    npts = feat.shape[-1] * feat.shape[-2]
    if tau.shape[0] != S or tau.shape[1] != 3 or not tau.is_contiguous() or not feat.is_contiguous():
        raise _lib.DimensionMismatch("pi_from_tau: feat (S, 4, …) and tau (S, 3, …), both contiguous")
    Pi = torch.empty((S, *feat.shape[-2:]), dtype=torch.float32, device=feat.device)
    _lib.check(_L().cgef_closure_pi(_p(feat), _p(tau), _p(tau_scale), float(sigma2), npts, S, _p(Pi),
                                    _stream(feat.device.index)))
    return Pi


def gradient_model_tau(feat, sigma2, base=None):
    """base + σ² G Gᵀ (the gradient model when base is None) per slice, FP32 (S, 3, …)."""
    S = feat.shape[0]
    npts = feat.shape[-1] * feat.shape[-2]
    if base is not None and (tuple(base.shape) != (S, 3, *feat.shape[-2:]) or not base.is_contiguous()):
        raise _lib.DimensionMismatch("base must be a contiguous (S, 3, …) tensor")
    tau = torch.empty((S, 3, *feat.shape[-2:]), dtype=torch.float32, device=feat.device)
    _lib.check(_L().cgef_closure_gradient_model(_p(feat), float(sigma2), _p(base), npts, S, _p(tau),
                                                _stream(feat.device.index)))
    return tau


def gradient_model_sigma2(kernel, scale):
    """Per-axis second moment σ² of the kernel's footprint, the gradient model's coefficient:
    Gaussian exp(−α d²/ℓ²) → ℓ²/(2α); top-hat disk of diameter ℓ → ℓ²/16."""
    if isinstance(kernel, GaussianKernel):
        return scale * scale / (2.0 * kernel.alpha)
    return scale * scale / 16.0


def pi_errors(pred, exact):
    """(correlation, relative L2) of a predicted Π against the exact Π, over all slices (FP64 host)."""
    p = pred.double().cpu().numpy().ravel()
    e = exact.double().cpu().numpy().ravel()
    corr = float(np.corrcoef(p, e)[0, 1])
    rel = float(np.linalg.norm(p - e) / np.linalg.norm(e))
    return corr, rel


class ClosureModel:
    """A trained closure: τ ≈ σ² G Gᵀ (gradient model) + tau_scale · net(feat_scale · ∇ū)."""

    def __init__(self, net, feat_scale, tau_scale, sigma2, history=()):
        self.net, self.feat_scale, self.tau_scale, self.sigma2 = net, feat_scale, tau_scale, float(sigma2)
        self.history = list(history)

    def predict_pi(self, feat, *, chunk=2):
        """Π of whole periodic fields feat (S, 4, N, M) from the network's τ."""
        r = predict_tau(self.net, feat, self.feat_scale, chunk=chunk)
        return pi_from_tau(feat, r, self.tau_scale, self.sigma2)


def fit_closure(feat, tau, sigma2, **train):
    """Train the network on the residual τ − σ² G Gᵀ (what the gradient model misses) of the exact pairs
    (feat, tau); keyword arguments go to train_closure. Returns a ClosureModel."""
    resid = gradient_model_tau(feat, -float(sigma2), base=tau.contiguous())
    net, fs, ts, hist = train_closure(feat, resid, **train)
    del resid
    return ClosureModel(net, fs, ts, sigma2, hist)


class LearnedClosure:
    """Π from a ClosureModel, checked against the exact Π: the first call and every `check_every`-th
    call also compute the exact Π (compute_Π! through this library's filter) and the relative L2
    error; above `threshold` (default 20 %) that call returns the exact Π and the closure stays on
    the exact path until `revalidate()`."""

    def __init__(self, model, grid, plan, *, threshold=0.2, check_every=16):
        self.model = model
        self.grid, self.plan = grid, plan
        self.threshold, self.check_every = float(threshold), int(check_every)
        self.calls = 0
        self.disabled = False
        self.last_error = None
        self.fallbacks = 0

    def revalidate(self):
        self.disabled = False

    def compute_pi_(self, Pi, u, v):
        """Π (Ny, Nx) FP32 for one slice of u, v. Returns (Pi, used_exact)."""
        self.calls += 1
        if self.disabled:
            self._exact(Pi, u, v)
            self.fallbacks += 1
            return Pi, True
        feat, _ = closure_features(u, v, self.grid, self.plan, tau=False)
        pred = self.model.predict_pi(feat)[0]
        if self.check_every and (self.calls - 1) % self.check_every == 0:
            exact = torch.empty_like(pred)
            self._exact(exact, u, v)
            _, rel = pi_errors(pred, exact)
            self.last_error = rel
            if rel > self.threshold:
                self.disabled = True
                self.fallbacks += 1
                Pi.copy_(exact)
                return Pi, True
        Pi.copy_(pred)
        return Pi, False

    def _exact(self, Pi, u, v):
        compute_pi_(Pi, u.float(), v.float(), None, self.grid, self.plan.kernel, self.plan.scale, filter_plan=self.plan)

"""The ring engine's Fourier-mode cut, which FP32 calls run under.

Target row r keeps the modes m < M(r): the smallest M at which the weight tail 2·Σ_{m≥M} Σ_offsets
|Ŵ(m)| is at most 3e-8 of the row's in-range weight sum and the same tail weighted by
max(1, 4ℓ·sin(πm/N)/Δx) at most 3e-6 of it. Every |DFT coefficient| of a row is at most N·max|q| and
every x-difference stencil multiplies mode m by at most 4·sin(πm/N)/Δx, so under ZeroFill (whose
denominator is that weight sum) the dropped part of a filtered value is at most 3e-8·max|q| and of an
x-difference at most 3e-6·max|q|/ℓ. Each test runs one FP32 call under the cut and again with CGEF_MODE_CUT=0 (every mode:
the exact operator the FP64 calls run), through every kernel that reads the contracted spectra — the
sphere's Fourier-space-rotation inverse + epilogue, the per-point-rotation fused kernel, the 256-point
fused kernel (2-D and 3-D), the staged inverse with the separate epilogue, the LDS ping-pong inverse —
and both contraction kernels, and checks the difference against that bound plus the FP32 outputs'
rounding, and the cut result against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

import cgef
from helpers import cart_axes, oracle_grid_cart, oracle_grid_sphere, rel_l2, smooth_random_field, sphere_axes
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
R_EARTH = 6.371e6
TOL = 3e-8          # the cut's value-tail bound, relative to the row's weight sum
F32_ULP = 2.0 ** -23


def N(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def T32(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=DEV)


def say(msg):
    print(msg, flush=True)


def within_cut(cut, exact, qmax, factor=1.0):
    """|cut − exact| ≤ factor·TOL·qmax + the two FP32 roundings of values that differ by that much."""
    lim = factor * TOL * qmax + F32_ULP * (np.abs(exact) + np.abs(cut)) + 1e-30
    return float(np.max(np.abs(cut - exact) / lim))


def blob_mask(shape, seed, frac=0.3):
    f = smooth_random_field(shape, seed, 1, smooth=8)[0]
    return f > np.quantile(f, frac)


def test_kept_fraction_falls_with_scale():
    lam, phi = sphere_axes(1440, 720)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(R_EARTH), lam, phi)
    ker = cgef.GaussianKernel(4.0)
    kept = []
    for ell in (50e3, 200e3, 1000e3, 2000e3):
        plan = cgef.plan_filter(g, ker, ell, engine="ring")
        info = plan.query()
        kept.append(info.ring_modes_kept)
        cut = plan.mode_cut()
        say(f"  {ell / 1e3:.0f} km: {info.ring_modes_kept:.3f} of the (row, mode) pairs kept; "
            f"rows keep {cut.min()}..{cut.max()} of {info.ring_modes} modes")
        assert cut.shape == (720,) and cut.min() >= 1 and cut.max() <= info.ring_modes
        assert np.array_equal(cut, cut[::-1])                 # the grid is symmetric about the equator
        assert abs(cut.sum() / (720 * info.ring_modes) - info.ring_modes_kept) < 1e-12
        assert cut[0] < cut[360]                               # the polar rows keep the fewest
    assert kept[0] > 0.9 and kept[-1] < 0.1
    assert all(a >= b for a, b in zip(kept, kept[1:]))
    # a translation-invariant grid keeps the same modes on every row
    xs = cart_axes(256, 1.0)
    gc = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    assert cgef.plan_filter(gc, cgef.GaussianKernel(6.0), 48.0, engine="ring").query().ring_modes_kept < 0.5
    # engines other than the ring have no cut
    banded = cgef.plan_filter(gc, cgef.GaussianKernel(6.0), 48.0, engine="banded")
    assert banded.query().ring_modes_kept == 0.0
    with pytest.raises(cgef._lib.ArgumentError, match="unsupported"):
        banded.mode_cut()


def _filter_case(kind):
    if kind == "sphere360":      # staged inverse (N = 360), per-row bands, masked
        lam, phi = sphere_axes(360, 180)
        mask = blob_mask((180, 360), 21)
        g = cgef.StructuredGrid(cgef.SphericalGeometry(R_EARTH), lam, phi, mask=mask)
        og = oracle_grid_sphere(lam, phi, mask=mask)
        return g, og, cgef.GaussianKernel(4.0), O.GAUSSIAN, 4.0, 3000e3
    if kind == "cart182":        # LDS ping-pong inverse (182 = 2·7·13), one band
        xs, ys = cart_axes(182, 1.0), cart_axes(96, 1.0)
        mask = blob_mask((96, 182), 22)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, True))
        og = oracle_grid_cart(xs, ys, periodic=(True, True), mask=mask)
        return g, og, cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0, 30.0
    xs, ys = cart_axes(256, 1.0), cart_axes(64, 1.0)   # 256-point rows, non-periodic y
    mask = blob_mask((64, 256), 23)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, False))
    og = oracle_grid_cart(xs, ys, periodic=(True, False), mask=mask)
    return g, og, cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0, 40.0


@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("kind", ["sphere360", "cart182", "cart256"])
def test_filter_apply_cut_within_bound(kind, strategy, monkeypatch):
    g, og, ker, okind, alpha, ell = _filter_case(kind)
    s_cg, s_o = (cgef.ZeroFill(), O.ZEROFILL) if strategy == "zerofill" else (cgef.Deformable(), O.DEFORMABLE)
    plan = cgef.plan_filter(g, ker, ell, mask_strategy=s_cg, engine="ring")
    kept = plan.query().ring_modes_kept
    assert kept < 0.9, kept   # the cut drops modes here, so the comparison below tests it
    f = smooth_random_field((3,) + g.field_shape, 31, 1, smooth=1)[0]
    inp = T32(f)
    fin = N(inp)
    out = torch.empty_like(inp)
    cgef.filter_apply_(out, inp, plan)
    cut = N(out)
    monkeypatch.setenv("CGEF_MODE_CUT", "0")
    cgef.filter_apply_(out, inp, plan)
    exact = N(out)
    monkeypatch.delenv("CGEF_MODE_CUT")
    ref = O.filter_fields(og, list(fin), kind=okind, alpha=alpha, ell=ell, strategy=s_o)
    for b in range(3):
        qmax = np.max(np.abs(fin[b]))
        # Deformable divides by the filtered mask instead of the weight sum: the bound grows by their ratio
        r = within_cut(cut[b], exact[b], qmax, factor=1.0 if strategy == "zerofill" else 50.0)
        say(f"  {kind} {strategy} slice {b}: kept {kept:.3f}, |cut − exact| / bound {r:.2f}, "
            f"rel-L2 vs oracle {rel_l2(cut[b], ref[b]):.2e}")
        assert r <= 1.0
        assert np.max(np.abs(cut[b] - ref[b])) <= 1e-6 * np.max(np.abs(ref[b]))


@pytest.fixture(scope="module")
def sphere1440():
    lam, phi = sphere_axes(1440, 720)
    mask = blob_mask((720, 1440), 41)
    u, v = smooth_random_field((720, 1440), 42, 2, smooth=6)
    return lam, phi, mask, u, v


@pytest.mark.parametrize("strategy,fuse", [("zerofill", None), ("zerofill", "1"), ("zerofill", "0"),
                                           ("deformable", None), ("deformable", "0")])
def test_sphere_pi_cut_every_inverse_path(sphere1440, strategy, fuse, monkeypatch):
    lam, phi, mask, u, v = sphere1440
    say(f"config: 1440x720 sphere, 1000 km, {strategy}, CGEF_FUSE={fuse}")
    if fuse is None:
        monkeypatch.delenv("CGEF_FUSE", raising=False)
    else:
        monkeypatch.setenv("CGEF_FUSE", fuse)
    s_cg, s_o = (cgef.ZeroFill(), O.ZEROFILL) if strategy == "zerofill" else (cgef.Deformable(), O.DEFORMABLE)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(R_EARTH), lam, phi, mask=mask)
    ker, ell = cgef.GaussianKernel(4.0), 1000e3
    plan = cgef.plan_filter(g, ker, ell, mask_strategy=s_cg, engine="ring")
    assert plan.query().ring_modes_kept < 0.3
    U, V = T32(u), T32(v)
    outs = {}
    for cut in (True, False):
        if not cut:
            monkeypatch.setenv("CGEF_MODE_CUT", "0")
        Pi = torch.empty_like(U)
        ws = cgef.PiWorkspace(g, (), torch.float32)
        cgef.compute_pi_(Pi, U, V, None, g, ker, ell, workspace=ws, filter_plan=plan, mask_strategy=s_cg)
        outs[cut] = (N(Pi), N(ws.u_filt), N(ws.v_filt))
    monkeypatch.delenv("CGEF_MODE_CUT")
    (pc, uc, vc), (pe, ue, ve) = outs[True], outs[False]
    smax = float(np.max(np.hypot(N(U), N(V))))
    fac = 2.0 if strategy == "zerofill" else 100.0   # the frame rotation's norm ≤ √3 < 2
    ru, rv = within_cut(uc, ue, smax, fac), within_cut(vc, ve, smax, fac)
    dpi = rel_l2(pc, pe)
    say(f"  ū, v̄ |cut − exact| / bound {ru:.2f}, {rv:.2f}; Π rel-L2 cut vs exact {dpi:.2e}")
    assert ru <= 1.0 and rv <= 1.0
    assert dpi <= 1e-6
    rows = [0, 119, 360, 600, 719]
    og = oracle_grid_sphere(lam, phi, mask=mask)
    ref, ub, vb = O.compute_pi_rows(og, N(U), N(V), rows, kind=O.GAUSSIAN, alpha=4.0, ell=ell, strategy=s_o)
    err = rel_l2(pc[rows], ref[rows])
    per_row = [rel_l2(pc[r], ref[r]) for r in rows]
    say(f"  Π rows rel-L2 vs the FP64 oracle {err:.2e}, per row {['%.1e' % e for e in per_row]}")
    assert err <= 1e-5
    assert max(per_row) <= 1e-5   # the pole rows too, where Δx ≪ ℓ and the slope bound sets the cut
    assert rel_l2(uc[rows], ub[rows]) <= 1e-6


@pytest.mark.parametrize("dims", ["2d256", "3d256"])
def test_cartesian_pi_cut_256_point_rows(dims, monkeypatch):
    nx = 256
    xs = cart_axes(nx, 1.0)
    if dims == "2d256":
        ys = cart_axes(128, 1.0)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, periodic=(True, True))
        shape, ell, w = (128, nx), 40.0, None
    else:
        ys, zs = cart_axes(32, 1.0), cart_axes(6, 1.0)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, periodic=(True, True, False))
        shape, ell = (6, 32, nx), 24.0
    ker = cgef.GaussianKernel(6.0)
    plan = cgef.plan_filter(g, ker, ell, engine="ring")
    kept = plan.query().ring_modes_kept
    assert kept < 0.7, kept
    f = smooth_random_field(shape, 51, 3, smooth=2)
    U, V = T32(f[0]), T32(f[1])
    W = T32(f[2]) if dims == "3d256" else None
    outs = {}
    for cut in (True, False):
        if not cut:
            monkeypatch.setenv("CGEF_MODE_CUT", "0")
        Pi = torch.empty_like(U)
        ws = cgef.PiWorkspace(g, (), torch.float32)
        cgef.compute_pi_(Pi, U, V, W, g, ker, ell, workspace=ws, filter_plan=plan)
        outs[cut] = (N(Pi), N(ws.u_filt))
    monkeypatch.delenv("CGEF_MODE_CUT")
    (pc, uc), (pe, ue) = outs[True], outs[False]
    r = within_cut(uc, ue, float(np.max(np.abs(N(U)))))
    dpi = rel_l2(pc, pe)
    say(f"  {dims}: kept {kept:.3f}, ū |cut − exact| / bound {r:.2f}, Π rel-L2 cut vs exact {dpi:.2e}")
    assert r <= 1.0 and dpi <= 1e-6

"""GPU parity: libcgef_hip.so (through the C ABI, via cgef) against the FP64 oracle on the same seeded
inputs. Tolerances: FP64 kernels 1e-10 relative (summation order and FMA contraction are the only
differences); the FP32 fused Π path rel-L2 ≤ 1e-5 against the FP64 oracle (the stated target)."""
import numpy as np
import pytest
import torch

import cgef
from helpers import (cart_axes, oracle_grid_cart, oracle_grid_sphere, rel_l2, smooth_random_field, sphere_axes)
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
STRATS = {"zerofill": (cgef.ZeroFill(), O.ZEROFILL), "deformable": (cgef.Deformable(), O.DEFORMABLE)}


def T(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def N(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def blob_mask(shape, seed, frac=0.25):
    f = smooth_random_field(shape, seed, 1, smooth=6)[0]
    return f > np.quantile(f, frac)


# ------------------------------------------------------------------------------------------ 2-D --

def cart_case(periodic=True, masked=False, nx=48, ny=40, dx=1000.0):
    xs, ys = cart_axes(nx, dx), cart_axes(ny, dx * 1.25)
    mask = blob_mask((ny, nx), 11) if masked else None
    og = oracle_grid_cart(xs, ys, periodic=(periodic, periodic), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(periodic, periodic))
    return og, g


@pytest.mark.parametrize("engine", ["banded", "baseline"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("masked,periodic", [(False, True), (True, True), (True, False)])
def test_filter_cartesian_fp64(engine, strategy, masked, periodic):
    og, g = cart_case(periodic, masked)
    f = smooth_random_field(g.field_shape, 3, 2)
    s_cg, s_o = STRATS[strategy]
    ker = cgef.GaussianKernel(6.0)
    plan = cgef.plan_filter(g, ker, 6000.0, mask_strategy=s_cg, engine=engine)
    assert plan.engine == engine
    inp = T(np.stack(f))
    out = torch.empty_like(inp)
    cgef.filter_apply_(out, inp, plan)
    ref = O.filter_fields(og, f, kind=O.GAUSSIAN, alpha=6.0, ell=6000.0, strategy=s_o)
    got = N(out)
    for m in range(2):
        assert np.max(np.abs(got[m] - ref[m])) <= 1e-12 * np.max(np.abs(ref[m]))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_filter_banded_fp32_fp64_batch(dtype, tol):
    og, g = cart_case(True, True)
    f = smooth_random_field((5,) + g.field_shape, 4, 1)[0]
    plan = cgef.plan_filter(g, cgef.TopHatKernel(), 9000.0, engine="banded")
    inp = T(f, dtype)
    out = torch.empty_like(inp)
    cgef.filter_apply_(out, inp, plan)
    fin = N(inp)
    ref = O.filter_fields(og, list(fin), kind=O.TOPHAT, alpha=0.0, ell=9000.0)
    got = N(out)
    for b in range(5):
        assert np.max(np.abs(got[b] - ref[b])) <= tol * np.max(np.abs(ref[b]))


@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("masked,periodic", [(False, True), (True, True), (True, False)])
def test_filter_separable_fp64(strategy, masked, periodic):
    og, g = cart_case(periodic, masked)
    f = smooth_random_field(g.field_shape, 5, 1)[0]
    s_cg, s_o = STRATS[strategy]
    plan = cgef.plan_filter(g, cgef.GaussianKernel(6.0), 7000.0, mask_strategy=s_cg)
    assert plan.engine == "separable"  # the reference's default for a Cartesian Gaussian
    out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    cgef.filter_apply_(out, T(f), plan)
    ref = O.apply_separable(og, 6.0, 7000.0, s_o, [f])[0]
    assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))


def sphere_case(masked=True, nlon=96, nlat=48):
    lam, phi = sphere_axes(nlon, nlat)
    mask = blob_mask((nlat, nlon), 21, 0.3) if masked else None
    og = oracle_grid_sphere(lam, phi, mask=mask)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
    return og, g, lam, phi


@pytest.mark.parametrize("engine", ["banded", "baseline"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
def test_filter_sphere_fp64(engine, strategy):
    og, g, lam, phi = sphere_case(True)
    f = smooth_random_field(g.field_shape, 6, 1)[0]
    s_cg, s_o = STRATS[strategy]
    plan = cgef.plan_filter(g, cgef.GaussianKernel(4.0), 1.6e6, mask_strategy=s_cg, engine=engine)
    fp = O.Footprint(og, O.GAUSSIAN, 4.0, 1.6e6)
    assert plan.info.nentries == len(fp.w)
    out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    cgef.filter_apply_(out, T(f), plan)
    ref = O.apply_banded(og, fp, s_o, [f])[0]
    assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_derivatives_sphere_and_cartesian():
    og, g, lam, phi = sphere_case(True)
    f = smooth_random_field(g.field_shape, 7, 1)[0]
    for ax, fn in ((0, cgef.ddx_), (1, cgef.ddy_)):
        out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
        fn(out, T(f), g)
        ref = O.deriv(og, ax, f)
        assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))
    og2, g2 = cart_case(False, True)
    f2 = smooth_random_field(g2.field_shape, 8, 1)[0]
    for ax, fn in ((0, cgef.ddx_), (1, cgef.ddy_)):
        out = torch.empty(g2.field_shape, dtype=torch.float64, device=DEV)
        fn(out, T(f2), g2)
        ref = O.deriv(og2, ax, f2)
        assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))


def _run_lengths(mask, axis):
    """Lengths of the runs of active cells along `axis` (domain ends cut runs)."""
    m = np.moveaxis(np.asarray(mask, bool), axis, -1).reshape(-1, np.shape(mask)[axis])
    out = []
    for row in m:
        n = 0
        for a in row:
            if a:
                n += 1
            elif n:
                out.append(n)
                n = 0
        if n:
            out.append(n)
    return np.array(out)


@pytest.mark.parametrize("periodic", [True, False])
def test_derivatives_and_pi_speckled_mask(periodic):
    # a speckled mask (35 % of cells masked, independently) gives every derivative stencil along both
    # axes: centred, 3-node windows shifted at run ends, 2-node runs and isolated cells (ReduceInRun),
    # and on periodic axes runs that cross the seam
    rng = np.random.default_rng(5)
    nx, ny = 36, 29
    mask = rng.random((ny, nx)) > 0.35
    for ax in (0, 1):
        rl = _run_lengths(mask, 1 - ax)
        assert (rl == 1).any() and (rl == 2).any() and (rl >= 4).any()
    xs, ys = cart_axes(nx, 1000.0), cart_axes(ny, 1250.0)
    og = oracle_grid_cart(xs, ys, periodic=(periodic, periodic), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(periodic, periodic))
    f = smooth_random_field(g.field_shape, 9, 1)[0]
    for ax, fn in ((0, cgef.ddx_), (1, cgef.ddy_)):
        out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
        fn(out, T(f), g)
        ref = O.deriv(og, ax, f)
        assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))
    u, v = smooth_random_field(g.field_shape, 10, 2)
    ref = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=4000.0)[0]
    for engine in ("banded", "ring") if periodic else ("banded",):
        plan = cgef.plan_filter(g, cgef.GaussianKernel(6.0), 4000.0, engine=engine)
        Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
        cgef.compute_pi_(Pi, T(u), T(v), None, g, cgef.GaussianKernel(6.0), 4000.0, filter_plan=plan)
        assert rel_l2(N(Pi), ref) <= 1e-10, engine


def test_derivative_z_speckled_mask():
    rng = np.random.default_rng(6)
    xs, ys, zs = cart_axes(12, 1.0), cart_axes(10, 1.0), np.concatenate([[0.0], np.cumsum(0.8 * 1.15 ** np.arange(13))])
    shape = (len(zs), len(ys), len(xs))
    mask = rng.random(shape) > 0.35
    rl = _run_lengths(mask, 0)
    assert (rl == 1).any() and (rl == 2).any() and (rl >= 4).any()
    og = oracle_grid_cart(xs, ys, zs, periodic=(True, True, False), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, mask=mask, periodic=(True, True, False))
    f = smooth_random_field(shape, 11, 1)[0]
    out = torch.empty(shape, dtype=torch.float64, device=DEV)
    cgef.ddz_(out, T(f), g)
    ref = O.deriv(og, 2, f)
    assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))
    u, v, w = smooth_random_field(shape, 12, 3)
    Pi = torch.empty(shape, dtype=torch.float64, device=DEV)
    cgef.compute_pi_(Pi, T(u), T(v), T(w), g, cgef.TopHatKernel(), 2.5)
    ref = O.compute_pi(og, u, v, w, kind=O.TOPHAT, alpha=0.0, ell=2.5, engine=O.ENGINE_ND3)[0]
    assert rel_l2(N(Pi), ref) <= 1e-10


# ------------------------------------------------------------------------------------ Π, 2-D --

@pytest.mark.parametrize("engine", ["banded", "baseline", "separable"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("masked", [False, True])
def test_pi_cartesian_fp64(engine, strategy, masked):
    og, g = cart_case(True, masked)
    u, v = smooth_random_field(g.field_shape, 9, 2)
    s_cg, s_o = STRATS[strategy]
    ker = cgef.GaussianKernel(6.0)
    plan = cgef.plan_filter(g, ker, 6000.0, mask_strategy=s_cg, engine=engine)
    Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    ws = cgef.PiWorkspace(g, (), torch.float64)
    _, E = cgef.compute_pi_(Pi, T(u), T(v), None, g, ker, 6000.0, workspace=ws, filter_plan=plan, return_energy=True)
    eng = {"banded": O.ENGINE_BANDED, "baseline": O.ENGINE_BANDED, "separable": O.ENGINE_SEPARABLE}[engine]
    ref, ub, vb, _ = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=6000.0, strategy=s_o, engine=eng)
    assert rel_l2(N(Pi), ref) <= 1e-10
    assert rel_l2(N(ws.u_filt), ub) <= 1e-12 and rel_l2(N(ws.v_filt), vb) <= 1e-12
    assert abs(E - O.energy(og, ub, vb)) <= 1e-12 * abs(O.energy(og, ub, vb))


@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("masked", [False, True])
def test_pi_sphere_fp64(strategy, masked):
    og, g, lam, phi = sphere_case(masked)
    u, v = smooth_random_field(g.field_shape, 10, 2)
    s_cg, s_o = STRATS[strategy]
    ker = cgef.GaussianKernel(4.0)
    Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    ws = cgef.PiWorkspace(g, (), torch.float64)
    _, E = cgef.compute_pi_(Pi, T(u), T(v), None, g, ker, 1.2e6, workspace=ws, mask_strategy=s_cg, return_energy=True)
    ref, ub, vb, _ = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=4.0, ell=1.2e6, strategy=s_o)
    assert rel_l2(N(Pi), ref) <= 1e-10
    assert rel_l2(N(ws.u_filt), ub) <= 1e-12
    assert abs(E - O.energy(og, ub, vb)) <= 1e-12 * abs(O.energy(og, ub, vb))


@pytest.mark.parametrize("geom", ["cartesian", "sphere"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
def test_pi_fp32_meets_tolerance(geom, strategy):
    """FP32 inputs, FP32 fused kernel vs the FP64 oracle on the same (FP32-rounded) inputs."""
    if geom == "cartesian":
        og, g = cart_case(True, True, nx=96, ny=80)
        ell, alpha = 8000.0, 6.0
    else:
        og, g, _, _ = sphere_case(True, 144, 72)
        ell, alpha = 9e5, 4.0
    u, v = smooth_random_field(g.field_shape, 12, 2, smooth=2)
    u32, v32 = T(u + 0.3, torch.float32), T(v - 0.2, torch.float32)   # a mean flow exercises the shift
    s_cg, s_o = STRATS[strategy]
    Pi = torch.empty(g.field_shape, dtype=torch.float32, device=DEV)
    cgef.compute_pi_(Pi, u32, v32, None, g, cgef.GaussianKernel(alpha), ell, mask_strategy=s_cg,
                     filter_plan=cgef.plan_filter(g, cgef.GaussianKernel(alpha), ell, mask_strategy=s_cg,
                                                  engine="banded"))
    ref, _, _, _ = O.compute_pi(og, N(u32), N(v32), kind=O.GAUSSIAN, alpha=alpha, ell=ell, strategy=s_o)
    assert rel_l2(N(Pi), ref) <= 1e-5


def test_pi_batch_equals_slices():
    og, g, _, _ = sphere_case(True)
    fs = smooth_random_field((3,) + g.field_shape, 13, 2)
    u, v = T(fs[0], torch.float32), T(fs[1], torch.float32)
    ker = cgef.GaussianKernel(4.0)
    plan = cgef.plan_filter(g, ker, 1.0e6)
    Pb = torch.empty_like(u)
    _, Eb = cgef.compute_pi_(Pb, u, v, None, g, ker, 1.0e6, filter_plan=plan, return_energy=True)
    for b in range(3):
        Ps = torch.empty_like(u[b])
        _, Es = cgef.compute_pi_(Ps, u[b].contiguous(), v[b].contiguous(), None, g, ker, 1.0e6, filter_plan=plan,
                                 return_energy=True)
        assert torch.equal(Ps, Pb[b])
        assert Es == Eb[b]


def test_cumulative_energy_and_spectrum():
    og, g, _, _ = sphere_case(True)
    u, v = smooth_random_field(g.field_shape, 14, 2)
    scales = [6e5, 9e5, 1.3e6, 2e6]
    E = cgef.cumulative_energy(T(u), T(v), None, g, cgef.GaussianKernel(4.0), scales)
    Eo = O.cumulative_energy(og, u, v, scales, kind=O.GAUSSIAN, alpha=4.0)
    assert np.max(np.abs(E - Eo) / np.abs(Eo)) <= 1e-12
    k, Et = cgef.filtering_spectrum(T(u), T(v), None, g, cgef.GaussianKernel(4.0), scales, L=1.0)
    assert np.allclose(Et, O.spectral_density(Eo, 1.0 / np.asarray(scales)), rtol=1e-9)


# ------------------------------------------------------------------------------------------ 3-D --

def volume_case(stretched=True, masked=False):
    xs, ys = cart_axes(24, 1.0), cart_axes(20, 1.0)
    if stretched:
        zs = np.concatenate([[0.0], np.cumsum(0.8 * 1.15 ** np.arange(11))])
    else:
        zs = cart_axes(12, 1.0)
    shape = (len(zs), len(ys), len(xs))
    mask = blob_mask(shape, 31, 0.2) if masked else None
    og = oracle_grid_cart(xs, ys, zs, periodic=(True, True, False), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, mask=mask, periodic=(True, True, False))
    return og, g


@pytest.mark.parametrize("stretched", [True, False])
@pytest.mark.parametrize("kernel", ["tophat", "gaussian"])
@pytest.mark.parametrize("masked", [False, True])
def test_pi_3d_fp64(stretched, kernel, masked):
    og, g = volume_case(stretched, masked)
    u, v, w = smooth_random_field(g.field_shape, 15, 3)
    ker = cgef.TopHatKernel() if kernel == "tophat" else cgef.GaussianKernel(6.0)
    kind, alpha, ell = (O.TOPHAT, 0.0, 6.0) if kernel == "tophat" else (O.GAUSSIAN, 6.0, 3.0)
    Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    ws = cgef.PiWorkspace(g, (), torch.float64)
    _, E = cgef.compute_pi_(Pi, T(u), T(v), T(w), g, ker, ell, workspace=ws, return_energy=True)
    ref, ub, vb, wb = O.compute_pi(og, u, v, w, kind=kind, alpha=alpha, ell=ell, engine=O.ENGINE_ND3)
    assert rel_l2(N(Pi), ref) <= 1e-10
    assert rel_l2(N(ws.w_filt), wb) <= 1e-12
    assert abs(E - O.energy(og, ub, vb, wb)) <= 1e-12 * O.energy(og, ub, vb, wb)


def test_pi_3d_fp32():
    og, g = volume_case(True, False)
    u, v, w = smooth_random_field(g.field_shape, 16, 3)
    f32 = [T(a, torch.float32) for a in (u, v, w)]
    Pi = torch.empty(g.field_shape, dtype=torch.float32, device=DEV)
    cgef.compute_pi_(Pi, *f32, g, cgef.TopHatKernel(), 6.0)
    ref, _, _, _ = O.compute_pi(og, *[N(a) for a in f32], kind=O.TOPHAT, alpha=0.0, ell=6.0, engine=O.ENGINE_ND3)
    assert rel_l2(N(Pi), ref) <= 1e-5


@pytest.mark.parametrize("masked", [False, True])
def test_pi_3d_contraction_variants_same_bits(masked, monkeypatch):
    """The three VALU contraction variants of a multi-plane (dz) footprint (8 or 4 columns per
    block at one wave per SIMD, 6 columns at six) give the same bits, and the oracle's Π."""
    og, g = volume_case(True, masked)
    u, v, w = smooth_random_field(g.field_shape, 18, 3)
    ker = cgef.TopHatKernel()
    plan = cgef.plan_filter(g, ker, 6.0, engine="ring")
    outs = []
    for variant in ("0", "1", "2"):
        print(f"3-D contraction variant {variant} masked={masked}", flush=True)
        monkeypatch.setenv("CGEF_CONTRACT_VARIANT", variant)
        Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
        cgef.compute_pi_(Pi, T(u), T(v), T(w), g, ker, 6.0, filter_plan=plan)
        outs.append(N(Pi))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref, _, _, _ = O.compute_pi(og, u, v, w, kind=O.TOPHAT, alpha=0.0, ell=6.0, engine=O.ENGINE_ND3)
    assert rel_l2(outs[-1], ref) <= 1e-10


def test_derivative_z_stretched():
    og, g = volume_case(True, True)
    f = smooth_random_field(g.field_shape, 17, 1)[0]
    out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    cgef.ddz_(out, T(f), g)
    ref = O.deriv(og, 2, f)
    assert np.max(np.abs(N(out) - ref)) <= 1e-12 * np.max(np.abs(ref))


# --------------------------------------------------------------------------- errors on the GPU --

def test_error_paths():
    og, g = cart_case(True, False)
    with pytest.raises(cgef.ArgumentError):
        cgef.plan_filter(g, cgef.GaussianKernel(), -1.0)
    with pytest.raises(cgef.ArgumentError):
        cgef.plan_filter(g, cgef.GaussianKernel(), float("nan"))
    plan = cgef.plan_filter(g, cgef.GaussianKernel(), 5000.0, engine="banded")
    bad = torch.zeros((3, 3), dtype=torch.float64, device=DEV)
    with pytest.raises(cgef.DimensionMismatch):
        cgef.filter_apply_(bad, bad, plan)
    u = torch.zeros(g.field_shape, dtype=torch.float64, device=DEV)
    v = torch.zeros((2,) + g.field_shape, dtype=torch.float64, device=DEV)
    with pytest.raises(cgef.DimensionMismatch):
        cgef.compute_pi_(torch.zeros_like(u), u, v, None, g, cgef.GaussianKernel(), 5000.0)
    with pytest.raises(cgef.ArgumentError):
        cgef.filter_apply_(torch.empty_like(u), u.to(torch.float32), plan)


# ------------------------------------------------------------------------------ ring engine --
# The ring engine evaluates the same banded footprint in Fourier space along the periodic x axis
# (FP64 throughout): parity with the oracle's direct sum to FFT rounding.

# ℓ = 6000 km spans ±35 latitude rows (71 offsets): the register-blocked wide-footprint contraction
@pytest.mark.parametrize("ell", [1.6e6, 6.0e6])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-6)])
def test_ring_filter_sphere(ell, strategy, dtype, tol):
    og, g, _, _ = sphere_case(True)
    f = smooth_random_field((3,) + g.field_shape, 18, 1)[0]
    s_cg, s_o = STRATS[strategy]
    plan = cgef.plan_filter(g, cgef.GaussianKernel(4.0), ell, mask_strategy=s_cg)
    assert plan.engine == "ring"
    if ell > 5e6:
        assert plan.info.ring_offsets >= 64
    inp = T(f, dtype)
    out = torch.empty_like(inp)
    cgef.filter_apply_(out, inp, plan)
    fin = N(inp)
    fp = O.Footprint(og, O.GAUSSIAN, 4.0, ell)
    got = N(out)
    for b in range(3):
        ref = O.apply_banded(og, fp, s_o, [fin[b]])[0]
        assert np.max(np.abs(got[b] - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("nx,ny,ell", [(45, 30, 9000.0), (56, 24, 8000.0), (32, 32, 40000.0), (44, 20, 7000.0),
                                       (26, 18, 5000.0), (256, 6, 9000.0), (360, 6, 9000.0), (1440, 4, 30000.0),
                                       (2048, 4, 9000.0), (3000, 4, 9000.0), (4096, 3, 12000.0)])
def test_ring_filter_cartesian_odd_sizes_and_multiwrap(nx, ny, ell):
    # odd length (radices 3·3·5), a radix-7 length, a radius larger than the box (runs that wrap the
    # ring more than once), radix-11 / radix-13 lengths (served by the LDS ping-pong FFT), and the
    # register-staged factorizations 16·16, 3·12·10, 9·16·10, 8·16·16, 3·5·…, 16·16·16 (padded LDS)
    xs, ys = cart_axes(nx, 1000.0), cart_axes(ny, 1000.0)
    mask = blob_mask((ny, nx), 41)
    og = oracle_grid_cart(xs, ys, periodic=(True, True), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, True))
    f = smooth_random_field(g.field_shape, 19, 1)[0]
    for ker, kind, alpha in ((cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0), (cgef.TopHatKernel(), O.TOPHAT, 0.0)):
        plan = cgef.plan_filter(g, ker, ell, engine="ring")
        out = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
        cgef.filter_apply_(out, T(f), plan)
        ref = O.filter_fields(og, [f], kind=kind, alpha=alpha, ell=ell)[0]
        assert np.max(np.abs(N(out) - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
def test_ring_latitude_grids(symmetric, strategy):
    # a latitude grid symmetric about the equator and one from 80°S to 85°N (bands of no symmetry)
    lam, phi = sphere_axes(96, 48)
    if not symmetric:
        phi = np.radians(np.linspace(-80.0, 85.0, 48))
    mask = blob_mask((48, 96), 47, 0.3)
    og = oracle_grid_sphere(lam, phi, mask=mask)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
    s_cg, s_o = STRATS[strategy]
    plan = cgef.plan_filter(g, cgef.GaussianKernel(4.0), 1.6e6, mask_strategy=s_cg, engine="ring")
    assert plan.engine == "ring"
    f = smooth_random_field((2,) + g.field_shape, 48, 1)[0]
    out = torch.empty(f.shape, dtype=torch.float64, device=DEV)
    cgef.filter_apply_(out, T(f), plan)
    fp = O.Footprint(og, O.GAUSSIAN, 4.0, 1.6e6)
    got = N(out)
    for b in range(2):
        ref = O.apply_banded(og, fp, s_o, [f[b]])[0]
        assert np.max(np.abs(got[b] - ref)) <= 1e-11 * np.max(np.abs(ref))
    u, v = smooth_random_field(g.field_shape, 49, 2)
    Pi = torch.empty(g.field_shape, dtype=torch.float64, device=DEV)
    cgef.compute_pi_(Pi, T(u), T(v), None, g, cgef.GaussianKernel(4.0), 1.6e6, filter_plan=plan)
    ref = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=4.0, ell=1.6e6, strategy=s_o, fp=fp)[0]
    assert rel_l2(N(Pi), ref) <= 1e-10


@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ring_pi_wide_footprint(strategy, dtype):
    # 71 offsets per mode on the sphere: the register-blocked contraction, bands differing per row
    og, g, _, _ = sphere_case(True)
    s_cg, s_o = STRATS[strategy]
    u, v = smooth_random_field(g.field_shape, 46, 2)
    U, V = T(u, dtype), T(v, dtype)
    ker = cgef.GaussianKernel(4.0)
    plan = cgef.plan_filter(g, ker, 6.0e6, mask_strategy=s_cg)
    assert plan.engine == "ring" and plan.info.ring_offsets >= 64
    Pi = torch.empty_like(U)
    cgef.compute_pi_(Pi, U, V, None, g, ker, 6.0e6, filter_plan=plan)
    ref = O.compute_pi(og, N(U), N(V), kind=O.GAUSSIAN, alpha=4.0, ell=6.0e6, strategy=s_o)[0]
    assert rel_l2(N(Pi), ref) <= (1e-10 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("geom", ["sphere1440", "cartesian256", "volume256"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("form", ["default", "per_point_rotation"])
def test_fused_inverse_epilogue(geom, strategy, dtype, form, monkeypatch):
    # Row lengths with a compiled-in factorization (1440: the 0.25° longitude; 256) take the inverse
    # transforms fused with the epilogue: against the oracle, and against the two separate kernels
    # (CGEF_FUSE=0), for the sphere's local-frame τ, the 2-D and 3-D Cartesian sets, the mask plane of
    # Deformable and two snapshots. On the sphere the default rotates to the local frame in Fourier
    # space (longitudes ascending from an arbitrary λ₀, so the mode shifts carry a phase); CGEF_FUSE=1
    # rotates per point after the transforms.
    print(f"fused inverse+epilogue {geom} {strategy} {dtype} {form}", flush=True)
    s_cg, s_o = STRATS[strategy]
    if geom == "sphere1440":
        lam, phi = sphere_axes(1440, 36)
        lam = lam - 2.0
        mask = blob_mask((36, 1440), 44, 0.3)
        og = oracle_grid_sphere(lam, phi, mask=mask)
        g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
        ker, kind, alpha, ell, eng3 = cgef.GaussianKernel(4.0), O.GAUSSIAN, 4.0, 8.0e5, None
    elif geom == "cartesian256":
        xs, ys = cart_axes(256, 1000.0), cart_axes(40, 1000.0)
        mask = blob_mask((40, 256), 45)
        og = oracle_grid_cart(xs, ys, periodic=(True, True), mask=mask)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, True))
        ker, kind, alpha, ell, eng3 = cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0, 9000.0, None
    else:
        xs, ys = cart_axes(256, 1.0), cart_axes(12, 1.0)
        zs = np.concatenate([[0.0], np.cumsum(0.8 * 1.1 ** np.arange(9))])
        mask = blob_mask((10, 12, 256), 46)
        og = oracle_grid_cart(xs, ys, zs, periodic=(True, True, False), mask=mask)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, mask=mask, periodic=(True, True, False))
        ker, kind, alpha, ell, eng3 = cgef.TopHatKernel(), O.TOPHAT, 0.0, 5.0, O.ENGINE_ND3
    plan = cgef.plan_filter(g, ker, ell, mask_strategy=s_cg, engine="ring")
    comps = 3 if geom == "volume256" else 2
    fs = smooth_random_field((2,) + g.field_shape, 49, comps)
    F = [T(f, dtype) for f in fs]
    W = F[2] if comps == 3 else None
    ws = cgef.PiWorkspace(g, (2,), dtype)
    Pi = torch.empty_like(F[0])
    if form == "default":
        monkeypatch.delenv("CGEF_FUSE", raising=False)
    else:
        monkeypatch.setenv("CGEF_FUSE", "1")
    cgef.compute_pi_(Pi, F[0], F[1], W, g, ker, ell, workspace=ws, filter_plan=plan)
    got, ub = N(Pi), N(ws.u_filt)
    monkeypatch.setenv("CGEF_FUSE", "0")
    P2 = torch.empty_like(Pi)
    ws2 = cgef.PiWorkspace(g, (2,), dtype)
    cgef.compute_pi_(P2, F[0], F[1], W, g, ker, ell, workspace=ws2, filter_plan=plan)
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    assert rel_l2(got, N(P2)) <= tol
    assert rel_l2(ub, N(ws2.u_filt)) <= tol
    if geom != "cartesian256":   # the radial component on the sphere, w on the volume
        assert rel_l2(N(ws.w_filt), N(ws2.w_filt)) <= tol
    for b in range(2):
        args = [N(f[b]) for f in F]
        kw = dict(kind=kind, alpha=alpha, ell=ell, strategy=s_o)
        if eng3 is not None:
            kw["engine"] = eng3
        ref, rub, _, _ = O.compute_pi(og, *args, **kw)
        assert rel_l2(got[b], ref) <= (1e-10 if dtype == torch.float64 else 1e-5)
        assert rel_l2(ub[b], rub) <= (1e-11 if dtype == torch.float64 else 1e-6)


@pytest.mark.parametrize("case", ["sphere", "sphere_ragged", "cartesian_periodic", "cartesian_bounded"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("nbatch", [1, 3])
def test_matrix_core_contraction(case, strategy, nbatch, monkeypatch):
    # The contraction on the matrix cores (kernel variant 3, forced) against the oracle and against the
    # VALU contraction: row counts that are not multiples of the 16-row tiles or 64-row workgroups, one
    # band per latitude and a translation-invariant band, wrapped and clipped source rows, a mask
    # column (Deformable), and 9-30 plane columns (the different column-tile instances).
    print(f"matrix-core contraction {case} {strategy} nbatch={nbatch}", flush=True)
    s_cg, s_o = STRATS[strategy]
    if case == "sphere":
        og, g, _, _ = sphere_case(True)
        ker, kind, alpha, ell = cgef.GaussianKernel(4.0), O.GAUSSIAN, 4.0, 6.0e6
    elif case == "sphere_ragged":
        og, g, _, _ = sphere_case(True, nlon=72, nlat=70)
        ker, kind, alpha, ell = cgef.GaussianKernel(4.0), O.GAUSSIAN, 4.0, 4.0e6
    else:
        per = case == "cartesian_periodic"
        ny = 80 if per else 50
        xs, ys = cart_axes(64, 1000.0), cart_axes(ny, 1000.0)
        mask = blob_mask((ny, 64), 43)
        og = oracle_grid_cart(xs, ys, periodic=(True, per), mask=mask)
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, per))
        ker, kind, alpha, ell = cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0, 15000.0
    plan = cgef.plan_filter(g, ker, ell, mask_strategy=s_cg, engine="ring")
    assert plan.info.ring_offsets >= 48
    fs = smooth_random_field((nbatch,) + g.field_shape, 48, 2)
    U, V = T(fs[0]), T(fs[1])
    Pi = torch.empty_like(U)
    monkeypatch.setenv("CGEF_CONTRACT_VARIANT", "3")
    cgef.compute_pi_(Pi, U, V, None, g, ker, ell, filter_plan=plan)
    got = N(Pi)
    monkeypatch.setenv("CGEF_CONTRACT_VARIANT", "0")
    P0 = torch.empty_like(U)
    cgef.compute_pi_(P0, U, V, None, g, ker, ell, filter_plan=plan)
    assert rel_l2(got, N(P0)) <= 1e-12
    for b in range(nbatch):
        ref = O.compute_pi(og, fs[0][b], fs[1][b], kind=kind, alpha=alpha, ell=ell, strategy=s_o)[0]
        assert rel_l2(got[b], ref) <= 1e-10


def test_ring_falls_back_when_the_contraction_slab_exceeds_lds():
    # A footprint spanning more source rows than the contraction can stage in 160 KiB of LDS (5 columns
    # × (256 + dy range) × 16 B): engine = auto takes the banded direct sum of the same footprint,
    # engine = ring is refused at plan time rather than failing at launch.
    print("ring LDS-slab fallback", flush=True)
    lam, phi = sphere_axes(4, 2400)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi)
    ker, ell = cgef.GaussianKernel(4.0), 3.2e6        # radius 7,680 km: about ±921 rows of 0.075°
    plan = cgef.plan_filter(g, ker, ell)
    assert plan.engine == "banded"
    with pytest.raises(cgef.ArgumentError, match="LDS"):
        cgef.plan_filter(g, ker, ell, engine="ring")
    assert cgef.plan_filter(g, ker, 2.0e5).engine == "ring"
    f = smooth_random_field(g.field_shape, 47, 1)[0]
    out = torch.empty((2400, 4), dtype=torch.float64, device=DEV)
    cgef.filter_apply_(out, T(f), plan)
    og = oracle_grid_sphere(lam, phi)
    ref = O.filter_fields(og, [f], kind=O.GAUSSIAN, alpha=4.0, ell=ell)[0]
    assert np.max(np.abs(N(out) - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("geom", ["cartesian", "sphere", "volume"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ring_pi(geom, strategy, dtype):
    s_cg, s_o = STRATS[strategy]
    if geom == "cartesian":
        og, g = cart_case(True, True)
        ker, kw = cgef.GaussianKernel(6.0), dict(kind=O.GAUSSIAN, alpha=6.0, ell=6000.0)
        fields = smooth_random_field(g.field_shape, 20, 2)
    elif geom == "sphere":
        og, g, _, _ = sphere_case(True)
        ker, kw = cgef.GaussianKernel(4.0), dict(kind=O.GAUSSIAN, alpha=4.0, ell=1.2e6)
        fields = smooth_random_field(g.field_shape, 21, 2)
    else:
        og, g = volume_case(True, True)
        ker, kw = cgef.TopHatKernel(), dict(kind=O.TOPHAT, alpha=0.0, ell=6.0, engine=O.ENGINE_ND3)
        fields = smooth_random_field(g.field_shape, 22, 3)
    fields = [f + 0.5 for f in fields]        # a mean flow: τ is a small difference of large terms
    ins = [T(f, dtype) for f in fields]
    plan = cgef.plan_filter(g, ker, kw["ell"], mask_strategy=s_cg, engine="ring")
    Pi = torch.empty(g.field_shape, dtype=dtype, device=DEV)
    ws = cgef.PiWorkspace(g, (), dtype)
    w = ins[2] if geom == "volume" else None
    _, E = cgef.compute_pi_(Pi, ins[0], ins[1], w, g, ker, kw["ell"], workspace=ws, filter_plan=plan,
                            return_energy=True)
    ref, ub, vb, wb = O.compute_pi(og, *[N(a) for a in ins], strategy=s_o, **kw)
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    assert rel_l2(N(Pi), ref) <= tol
    assert rel_l2(N(ws.u_filt), ub) <= (1e-12 if dtype == torch.float64 else 1e-7)
    Eo = O.energy(og, ub, vb, wb if geom == "volume" else None)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)


# --------------------------------------------------------------------------------- pipeline --

def test_coarse_grain_matches_oracle_and_fallback():
    og, g, _, _ = sphere_case(True)
    u, v = smooth_random_field(g.field_shape, 23, 2)
    scales = [8e5, 1.2e6, 1.8e6]
    ker = cgef.GaussianKernel(4.0)
    res = cgef.coarse_grain(T(u), T(v), None, g, scales=scales, kernel=ker)
    Eo = []
    for s_idx, ell in enumerate(scales):
        ref, ub, vb, _ = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=4.0, ell=ell)
        assert rel_l2(N(res.Pi[s_idx]), ref) <= 1e-10
        Eo.append(O.energy(og, ub, vb))
        assert abs(res.cumulative_energy[s_idx] - Eo[-1]) <= 1e-12 * abs(Eo[-1])
    k = 1.0 / np.asarray(scales)
    assert np.allclose(res.wavenumber, k)
    assert np.allclose(res.filtering_spectrum, O.spectral_density(np.asarray(Eo), k), rtol=1e-9)
    # the reference's default kernel (top-hat) cannot give a spectral density under StrictSpectrum
    with pytest.raises(cgef.ArgumentError):
        cgef.coarse_grain(T(u), T(v), None, g, scales=scales)
    nos = cgef.coarse_grain(T(u), T(v), None, g, scales=scales, spectrum=cgef.NoSpectrum())
    assert np.all(np.isnan(nos.filtering_spectrum))
    # non-ring plans take the per-scale path and agree
    plans = [cgef.plan_filter(g, ker, s, engine="banded") for s in scales]
    res2 = cgef.pipeline._allocate_result(g, len(scales), torch.float64, DEV)
    cgef.coarse_grain_(res2, T(u), T(v), None, g, scales=scales, kernel=ker, filter_plans=plans)
    assert rel_l2(N(res2.Pi), N(res.Pi)) <= 1e-10


def test_coarse_grain_batch_equals_slices():
    og, g, _, _ = sphere_case(True)
    fs = smooth_random_field((2, 3) + g.field_shape, 24, 2)
    U, V = T(fs[0], torch.float32), T(fs[1], torch.float32)
    scales = [9e5, 1.5e6]
    ker = cgef.GaussianKernel(4.0)
    batch = cgef.CoarseGrainBatchResult.allocate(g, len(scales), (2, 3), torch.float32, DEV)
    cgef.coarse_grain_batch_(batch, U, V, None, g, scales=scales, kernel=ker)
    for b0 in range(2):
        for b1 in range(3):
            r = cgef.coarse_grain(U[b0, b1].contiguous(), V[b0, b1].contiguous(), None, g, scales=scales, kernel=ker)
            assert torch.equal(r.Pi, batch.Pi[b0, b1])
            assert np.array_equal(r.cumulative_energy, batch.cumulative_energy[b0, b1])
            assert np.allclose(r.filtering_spectrum, batch.filtering_spectrum[b0, b1])


def test_coarse_grain_profile_is_the_level_batch():
    # Pipeline.jl:991-1010: per-level sweeps on a shared horizontal grid = coarse_grain_batch! over levels
    og, g, _, _ = sphere_case(True)
    fs = smooth_random_field((3,) + g.field_shape, 25, 2)
    U, V = T(fs[0], torch.float32), T(fs[1], torch.float32)
    scales = [9e5, 1.5e6]
    ker = cgef.GaussianKernel(4.0)
    prof = cgef.coarse_grain_profile(U, V, None, g, scales=scales, kernel=ker)
    assert tuple(prof.Pi.shape) == (3, 2) + g.field_shape
    for lev in range(3):
        r = cgef.coarse_grain(U[lev].contiguous(), V[lev].contiguous(), None, g, scales=scales, kernel=ker)
        assert torch.equal(r.Pi, prof.Pi[lev])
        assert np.array_equal(r.cumulative_energy, prof.cumulative_energy[lev])


def test_coarse_grain_slices_and_filter_slices_ragged_grids():
    # Pipeline.jl:866-919 / Filtering.jl:3433-3470: independent slices, each with its own grid
    og1, g1 = cart_case(True, True, nx=40, ny=24)
    og2, g2 = cart_case(False, True, nx=28, ny=36)
    ker = cgef.GaussianKernel(6.0)
    scales = [4000.0, 7000.0]
    us, vs, refs = [], [], []
    for og, g, seed in ((og1, g1, 26), (og2, g2, 27)):
        u, v = smooth_random_field(g.field_shape, seed, 2)
        us.append(T(u))
        vs.append(T(v))
        # the default engine for a Gaussian on a uniform Cartesian grid is the reference's separable one
        refs.append([O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=s, engine=O.ENGINE_SEPARABLE)[0]
                     for s in scales])
    results = [cgef.pipeline._allocate_result(g, len(scales), torch.float64, DEV) for g in (g1, g2)]
    cgef.coarse_grain_slices_(results, us, vs, None, [g1, g2], scales=scales, kernel=ker)
    for r, ref in zip(results, refs):
        for q in range(len(scales)):
            assert rel_l2(N(r.Pi[q]), ref[q]) <= 1e-10
    plans = [cgef.plan_filter(g, ker, 5000.0) for g in (g1, g2)]
    outs = [torch.empty_like(u) for u in us]
    cgef.filter_slices_(outs, us, plans)
    for o, u, p in zip(outs, us, plans):
        single = torch.empty_like(u)
        cgef.filter_apply_(single, u, p)
        assert torch.equal(o, single)
    with pytest.raises(cgef.DimensionMismatch):
        cgef.filter_slices_(outs, us, plans[:1])


# ------------------------------------------------------------------- prefix-sum top-hat engine --
# Filtering.jl:751-1141: the reference's engine for a top-hat on a 2-D grid, against the oracle's
# restatement of the same algorithm (itself pinned to the disk footprint and brute force).

def prefix_case(case, masked=True):
    rng = np.random.Generator(np.random.PCG64(31))
    if case == "sphere":
        lam, phi = sphere_axes(96, 48)
        mask = blob_mask((48, 96), 32) if masked else None
        og = oracle_grid_sphere(lam, phi, mask=mask)
        g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
        return og, g, 1.1e6
    ny, nx = 30, 44
    mask = blob_mask((ny, nx), 33) if masked else None
    ys = cart_axes(ny, 1250.0)
    if case.startswith("nonuniform"):
        xs = np.cumsum(rng.uniform(600.0, 1400.0, nx))
        per = case.endswith("periodic")
        L = float(xs[-1] - xs[0] + 1000.0)
        og = O.Grid([xs, ys], spherical=False, periodic=[per, per], period=[L, ny * 1250.0], mask=mask,
                    uniform=[False, True])
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(per, per),
                                period=(L, ny * 1250.0))
        return og, g, 9000.0
    per = case in ("periodic", "full_row")
    xs = cart_axes(nx, 1000.0, 500.0)
    og = oracle_grid_cart(xs, ys, periodic=(per, per), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(per, per))
    return og, g, (60000.0 if case == "full_row" else 9000.0)


@pytest.mark.parametrize("case", ["periodic", "bounded", "nonuniform", "nonuniform_periodic", "sphere", "full_row"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
def test_prefix_filter(case, strategy):
    og, g, ell = prefix_case(case)
    s_cg, s_o = STRATS[strategy]
    plan = cgef.plan_filter(g, cgef.TopHatKernel(), ell, mask_strategy=s_cg)
    assert plan.engine == "prefixsum"
    f = smooth_random_field((2,) + g.field_shape, 34, 1)[0]
    ref = O.apply_prefix(og, ell, s_o, [f[0], f[1]])
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-6)):
        inp = T(f, dtype)
        out = torch.empty_like(inp)
        cgef.filter_apply_(out, inp, plan)
        got = N(out)
        fin = N(inp)
        ref = O.apply_prefix(og, ell, s_o, [fin[0], fin[1]])
        for b in range(2):
            assert np.max(np.abs(got[b] - ref[b])) <= tol * np.max(np.abs(ref[b]))


@pytest.mark.parametrize("case", ["periodic", "sphere", "nonuniform"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_prefix_pi_and_default_pipeline(case, dtype):
    og, g, ell = prefix_case(case)
    u, v = smooth_random_field(g.field_shape, 35, 2)
    u, v = u + 0.3, v - 0.2
    U, V = T(u, dtype), T(v, dtype)
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(g, (), dtype)
    _, E = cgef.compute_pi_(Pi, U, V, None, g, cgef.TopHatKernel(), ell, workspace=ws, return_energy=True)
    ref, ub, vb, _ = O.compute_pi(og, N(U), N(V), kind=O.TOPHAT, alpha=0.0, ell=ell, engine=O.ENGINE_PREFIX)
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    assert rel_l2(N(Pi), ref) <= tol
    assert rel_l2(N(ws.u_filt), ub) <= (1e-12 if dtype == torch.float64 else 1e-7)
    assert rel_l2(N(ws.v_filt), vb) <= (1e-12 if dtype == torch.float64 else 1e-7)
    assert abs(E - O.energy(og, ub, vb)) <= 1e-10 * abs(O.energy(og, ub, vb))
    # the pipeline's default kernel is the top-hat (Pipeline.jl:344): coarse_grain takes this engine
    res = cgef.coarse_grain(U, V, None, g, scales=[ell], spectrum=cgef.NoSpectrum())
    assert rel_l2(N(res.Pi[0]), ref) <= tol


# ------------------------------------------------------------------ spectral (FFTWExt.jl:121-208) --

def spectral_case(ny, nx, masked, seed=41):
    xs, ys = cart_axes(nx, 1000.0, 500.0), cart_axes(ny, 1250.0)
    mask = blob_mask((ny, nx), seed) if masked else None
    og = oracle_grid_cart(xs, ys, periodic=(True, True), mask=mask)
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, mask=mask, periodic=(True, True))
    return og, g


KERNELS = {"gaussian": (cgef.GaussianKernel(6.0), O.GAUSSIAN, 6.0), "tophat": (cgef.TopHatKernel(), O.TOPHAT, 0.0)}


# (26, 44): radix 13 along y takes the LDS ping-pong transform; (90, 120): radices 9·10 and 15·8;
# (256, 192): 16·16 and 12·16 in the padded layout; (4096, 6): a 4096-point column transform
@pytest.mark.parametrize("shape", [(24, 32), (21, 30), (26, 44), (90, 120), (256, 192), (4096, 6)])
@pytest.mark.parametrize("kernel", ["gaussian", "tophat"])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
def test_spectral_filter(shape, kernel, strategy):
    og, g = spectral_case(*shape, masked=True)
    ker, kind, alpha = KERNELS[kernel]
    s_cg, s_o = STRATS[strategy]
    ell = 6000.0
    plan = cgef.plan_filter(g, ker, ell, mask_strategy=s_cg, method=cgef.Spectral())
    assert plan.engine == "spectral"
    f = smooth_random_field((2,) + g.field_shape, 42, 1)[0] + 0.5
    # FP64 in: transform round-off only; FP32 in: the FP32 input and output roundings
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-6)):
        inp = T(f, dtype)
        out = torch.empty_like(inp)
        cgef.filter_apply_(out, inp, plan)
        fin = N(inp)
        ref = O.apply_spectral(og, kind, alpha, ell, s_o, [fin[0], fin[1]])
        got = N(out)
        for b in range(2):
            assert np.max(np.abs(got[b] - ref[b])) <= tol * np.max(np.abs(ref[b]))


def test_spectral_unmasked_constant_and_method_resolution():
    og, g = spectral_case(24, 30, masked=False)
    ker = cgef.GaussianKernel(6.0)
    plan = cgef.plan_filter(g, ker, 5000.0, method=cgef.Spectral())
    # Ĝ(0) = 1: a constant passes unchanged, with no renormalization on an unmasked grid
    c = torch.full(g.field_shape, 2.5, dtype=torch.float64, device=DEV)
    out = torch.empty_like(c)
    cgef.filter_apply_(out, c, plan)
    assert np.max(np.abs(N(out) - 2.5)) <= 1e-13
    # AutoMethod takes the transform on a periodic uniform grid, RealSpace elsewhere (Filtering.jl:3012-3019)
    assert cgef.plan_filter(g, ker, 5000.0, method=cgef.AutoMethod()).engine == "spectral"
    assert cgef.plan_filter(g, ker, 5000.0).engine == "separable"
    _, gb = cart_case(periodic=False)
    assert cgef.plan_filter(gb, ker, 5000.0, method=cgef.AutoMethod()).engine != "spectral"
    # a length with a prime factor above 13 has no device transform
    _, g17 = spectral_case(24, 34, masked=False)
    with pytest.raises(Exception, match="spectral engine"):
        cgef.plan_filter(g17, ker, 5000.0, method=cgef.Spectral())


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("strategy", ["zerofill", "deformable"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spectral_pi(masked, strategy, dtype):
    og, g = spectral_case(40, 60, masked, seed=44)
    s_cg, s_o = STRATS[strategy]
    u, v = smooth_random_field(g.field_shape, 45, 2)
    u, v = u + 0.3, v - 0.2
    U, V = T(u, dtype), T(v, dtype)
    ell = 7000.0
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(g, (), dtype)
    _, E = cgef.compute_pi_(Pi, U, V, None, g, cgef.GaussianKernel(6.0), ell, workspace=ws, mask_strategy=s_cg,
                            method=cgef.Spectral(), return_energy=True)
    ref, ub, vb, _ = O.compute_pi(og, N(U), N(V), kind=O.GAUSSIAN, alpha=6.0, ell=ell, strategy=s_o,
                                  engine=O.ENGINE_SPECTRAL)
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    assert rel_l2(N(Pi), ref) <= tol
    assert rel_l2(N(ws.u_filt), ub) <= (1e-12 if dtype == torch.float64 else 1e-7)
    assert rel_l2(N(ws.v_filt), vb) <= (1e-12 if dtype == torch.float64 else 1e-7)
    assert abs(E - O.energy(og, ub, vb)) <= 1e-10 * abs(O.energy(og, ub, vb))
    # several scales share one forward transform of the inputs
    scales = [4000.0, 9000.0]
    res = cgef.coarse_grain(U, V, None, g, scales=scales, kernel=cgef.TopHatKernel(), mask_strategy=s_cg,
                            method=cgef.Spectral(), spectrum=cgef.NoSpectrum())
    for q, s in enumerate(scales):
        rq, ubq, vbq, _ = O.compute_pi(og, N(U), N(V), kind=O.TOPHAT, alpha=0.0, ell=s, strategy=s_o,
                                       engine=O.ENGINE_SPECTRAL)
        assert rel_l2(N(res.Pi[q]), rq) <= tol
        Eq = O.energy(og, ubq, vbq)
        assert abs(res.cumulative_energy[q] - Eq) <= 1e-10 * abs(Eq)
    # filtering_spectrum's cumulative energy from the filtered raw velocities
    cum = cgef.cumulative_energy(U, V, None, g, cgef.TopHatKernel(), scales, mask_strategy=s_cg,
                                 method=cgef.Spectral())
    for q, s in enumerate(scales):
        fu, fv = O.apply_spectral(og, O.TOPHAT, 0.0, s, s_o, [N(U), N(V)])
        Eq = O.energy(og, fu, fv)
        assert abs(cum[q] - Eq) <= (1e-12 if dtype == torch.float64 else 1e-6) * abs(Eq)

"""Parity at the benchmark configurations' own workloads, full size, through the C ABI on the GPU
against the FP64 oracle (the reference's algorithm restated; Pipeline.jl:519-563 is the batch driver
the reference gates the same way in test/test_backends.jl:694-793).

* 256² doubly periodic Cartesian k⁻³ box, FP64, Gaussian α = 6, ℓ = 8Δx, ρ₀ = 1025: every
  engine that serves it, rel-L2(Π) ≤ 1e-10.
* 0.25° sphere (1440 × 720), FP32 in and out, 30 % land ZeroFill, Gaussian α = 4, ℓ = 200 km:
  rel-L2(Π) ≤ 1e-5 overall, on the coast band and in the open-ocean interior; E(ℓ) and ū, v̄.
* 8 snapshots × 24 scales (50–2000 km) on the same grid, one batched sweep per contraction kernel
  (three VALU variants and the matrix-core one, forced, and the measured choice): Π at the poles, ±60° and the equator for the two widest-footprint scales (≈1000 and
  2000 km: up to 347 row offsets, source slabs crossing the pole) and the full field at 50 km,
  rel-L2 ≤ 1e-5; the oracle evaluates only the checked rows (and their stencil neighbours).
* 256 × 256 × 64 volume, stretched z, FP32, top-hat ℓ = 16Δx: rel-L2(Π) ≤ 1e-5.
* Learned closure on the 1024² box: the exact (∇ū, τ) pairs against the oracle's filter on a
  sampled field, then 2000 Adam steps at batch 16 and the held-out gain over the gradient model.
"""
import os

import numpy as np
import pytest
import torch

import cgef
from cgef import closure as CL
from cgef import synthetic as S
from helpers import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
R_EARTH = 6.371e6
NTHREADS = min(len(os.sched_getaffinity(0)), 16)


def N(t):
    return t.detach().cpu().double().numpy()


def say(msg):
    print(msg, flush=True)


def _sphere_grids(mask):
    lam, phi = S.sphere_axes()
    g = cgef.StructuredGrid(cgef.SphericalGeometry(R_EARTH), lam, phi, mask=mask)
    og = O.Grid([lam, phi], spherical=True, R=R_EARTH, periodic=[True, False], period=[2 * np.pi, 0.0],
                mask=mask, uniform=[True, True])
    return g, og


# ------------------------------------------------------------------------------- 256² box, FP64 --

def test_cartesian_256_fp64_every_engine():
    say("config: cartesian 256^2 FP64")
    n, L = 256, 1.0e6
    ell = 8 * L / n
    u, v, xs = S.cartesian_turbulence(N=n, L=L, seed=0)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    og = O.Grid([xs, xs], periodic=[True, True], period=[L, L], uniform=[True, True])
    ker = cgef.GaussianKernel(6.0)
    U, V = torch.tensor(u, device=DEV), torch.tensor(v, device=DEV)
    refs = {}
    for eng, oeng in (("separable", O.ENGINE_SEPARABLE), ("banded", O.ENGINE_BANDED)):
        P, ub, vb, _ = O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=ell, engine=oeng, nthreads=NTHREADS)
        refs[eng] = (1025.0 * P, O.energy(og, ub, vb), ub)
    for eng, ref in (("auto", "separable"), ("banded", "banded"), ("ring", "banded")):
        plan = cgef.plan_filter(grid, ker, ell, engine=eng)
        assert eng != "auto" or plan.engine == "separable"   # the reference's default (Filtering.jl:1533-1540)
        Pi = torch.empty_like(U)
        ws = cgef.PiWorkspace(grid, (), torch.float64)
        _, E = cgef.compute_pi_(Pi, U, V, None, grid, ker, ell, workspace=ws, filter_plan=plan, rho0=1025.0,
                                return_energy=True)
        ubar = ws.u_filt
        rp, re, rub = refs[ref]
        err = rel_l2(N(Pi), rp)
        say(f"  {plan.engine}: rel-L2 {err:.2e}")
        assert err <= 1e-10
        assert abs(E - re) <= 1e-12 * re
        assert rel_l2(N(ubar), rub) <= 1e-12


# -------------------------------------------------------------------- 0.25° sphere, FP32 ---------

@pytest.fixture(scope="module")
def sphere_cfg():
    say("config: generating the 0.25-degree workload")
    u, v, _, _ = S.sphere_turbulence()
    mask = S.sphere_land_mask()
    g, og = _sphere_grids(mask)
    return u, v, mask, g, og


def test_sphere_quarter_degree_fp32(sphere_cfg):
    u, v, mask, g, og = sphere_cfg
    ker, ell = cgef.GaussianKernel(4.0), 200e3
    U = torch.tensor(u, dtype=torch.float32, device=DEV)
    V = torch.tensor(v, dtype=torch.float32, device=DEV)
    uin, vin = N(U), N(V)   # the oracle filters the same FP32 values, in FP64
    fp = O.Footprint(og, O.GAUSSIAN, 4.0, ell, copy=False)
    ref, ub, vb, _ = O.compute_pi(og, uin, vin, kind=O.GAUSSIAN, alpha=4.0, ell=ell, fp=fp, nthreads=NTHREADS)
    Eo = O.energy(og, ub, vb)
    frac = O.apply_banded(og, fp, O.ZEROFILL, [mask.astype(np.float64)], nthreads=NTHREADS)[0]
    coast = mask & (frac < 1.0 - 1e-9)
    interior = mask & ~coast
    assert coast.sum() > 10000 and interior.sum() > 100000
    for eng in ("auto", "banded"):
        plan = cgef.plan_filter(g, ker, ell, engine=eng)
        assert eng != "auto" or plan.engine == "ring"
        Pi = torch.empty_like(U)
        ws = cgef.PiWorkspace(g, (), torch.float32)
        _, E = cgef.compute_pi_(Pi, U, V, None, g, ker, ell, workspace=ws, filter_plan=plan, return_energy=True)
        ubar, vbar = ws.u_filt, ws.v_filt
        got = N(Pi)
        errs = {"all": rel_l2(got, ref), "coast": rel_l2(got[coast], ref[coast]),
                "interior": rel_l2(got[interior], ref[interior])}
        say(f"  {plan.engine}: rel-L2 {errs}")
        assert all(e <= 1e-5 for e in errs.values()), errs
        assert np.all(got[~mask] == 0.0)
        assert abs(E - Eo) <= 1e-6 * Eo
        assert rel_l2(N(ubar), ub) <= 1e-6 and rel_l2(N(vbar), vb) <= 1e-6


# -------------------------------------------------- 8 snapshots × 24 scales, widest footprints ----

SWEEP_SCALES = np.geomspace(50e3, 2000e3, 24)
CHECK_ROWS = [0, 119, 360, 600, 719]   # south pole, −60°, equator, +60°, north pole


@pytest.fixture(scope="module")
def sweep_cfg(sphere_cfg):
    _, _, mask, g, og = sphere_cfg
    say("config: generating 8 snapshots (seeds 100-107)")
    us, vs, _, _ = S.sphere_turbulence_batch(range(100, 108))
    U = torch.tensor(np.stack(us), dtype=torch.float32, device=DEV)
    V = torch.tensor(np.stack(vs), dtype=torch.float32, device=DEV)
    ker = cgef.GaussianKernel(4.0)
    plans = [cgef.plan_filter(g, ker, float(s)) for s in SWEEP_SCALES]
    assert all(p.engine == "ring" for p in plans)
    info = plans[-1].info
    say(f"  2000 km plan: {info.ring_offsets} row offsets, {info.nentries} footprint entries")
    assert info.ring_offsets > 300
    # the oracle on the checked rows of snapshots 0 and 5 at the two widest-footprint scales
    refs = {}
    for q in (19, 23):
        for s in (0, 5):
            say(f"  oracle rows: scale {SWEEP_SCALES[q] / 1e3:.0f} km, snapshot {s}")
            refs[(q, s)] = O.compute_pi_rows(og, N(U[s]), N(V[s]), CHECK_ROWS, kind=O.GAUSSIAN, alpha=4.0,
                                             ell=float(SWEEP_SCALES[q]), nthreads=NTHREADS)
    P50, ub50, vb50, _ = O.compute_pi(og, N(U[3]), N(V[3]), kind=O.GAUSSIAN, alpha=4.0, ell=float(SWEEP_SCALES[0]),
                                      nthreads=NTHREADS)
    refs["full50"] = (P50, O.energy(og, ub50, vb50))
    return U, V, g, ker, plans, refs


@pytest.mark.parametrize("variant", ["0", "1", "2", "3", None])
def test_batched_sweep_widest_scales(sweep_cfg, variant, monkeypatch):
    U, V, g, ker, plans, refs = sweep_cfg
    say(f"config: 8x24 sweep, contraction variant {variant or 'measured'}")
    if variant is None:
        monkeypatch.delenv("CGEF_CONTRACT_VARIANT", raising=False)
    else:
        monkeypatch.setenv("CGEF_CONTRACT_VARIANT", variant)
    batch = cgef.CoarseGrainBatchResult.allocate(g, len(SWEEP_SCALES), (8,), torch.float32, DEV)
    ws = cgef.PiWorkspace(g, (8,), torch.float32)
    cgef.coarse_grain_batch_(batch, U, V, None, g, scales=SWEEP_SCALES, kernel=ker, filter_plans=plans,
                             workspaces=[ws])
    torch.cuda.synchronize()
    for (q, s), (ref, _, _) in ((k, r) for k, r in refs.items() if k != "full50"):
        got = N(batch.Pi[s, q])[CHECK_ROWS]
        err = rel_l2(got, ref[CHECK_ROWS])
        say(f"  {SWEEP_SCALES[q] / 1e3:.0f} km snapshot {s}: rows rel-L2 {err:.2e}")
        assert err <= 1e-5
        for r in CHECK_ROWS:   # each row on its own (pole rows carry the widest footprints)
            assert rel_l2(got[CHECK_ROWS.index(r)], ref[r]) <= 1e-5, r
    P50, E50 = refs["full50"]
    assert rel_l2(N(batch.Pi[3, 0]), P50) <= 1e-5
    assert abs(batch.cumulative_energy[3, 0] - E50) <= 1e-6 * E50


def test_widest_scale_filtered_velocity_single_snapshot(sweep_cfg):
    U, V, g, ker, plans, refs = sweep_cfg
    say("config: 2000 km single-snapshot Π with filtered-velocity export")
    Pi = torch.empty_like(U[0])
    ws = cgef.PiWorkspace(g, (), torch.float32)
    cgef.compute_pi_(Pi, U[5].contiguous(), V[5].contiguous(), None, g, ker, float(SWEEP_SCALES[23]), workspace=ws,
                     filter_plan=plans[23])
    ubar, vbar = ws.u_filt, ws.v_filt
    ref, ub, vb = refs[(23, 5)]
    assert rel_l2(N(Pi)[CHECK_ROWS], ref[CHECK_ROWS]) <= 1e-5
    assert rel_l2(N(ubar)[CHECK_ROWS], ub[CHECK_ROWS]) <= 1e-6
    assert rel_l2(N(vbar)[CHECK_ROWS], vb[CHECK_ROWS]) <= 1e-6


# ------------------------------------------------------------- 256 × 256 × 64 volume, FP32 -------

def test_volume_256x256x64_fp32():
    say("config: volume 256x256x64 FP32 top-hat")
    u, v, w = S.volume_turbulence()
    nz, ny, nx = u.shape
    xs, ys = np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64)
    zs = S.stretched_levels(nz)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, periodic=(True, True, False))
    ker, ell = cgef.TopHatKernel(), 16.0
    U, V, W = (torch.tensor(a, dtype=torch.float32, device=DEV) for a in (u, v, w))
    plan = cgef.plan_filter(grid, ker, ell)
    Pi = torch.empty_like(U)
    _, E = cgef.compute_pi_(Pi, U, V, W, grid, ker, ell, filter_plan=plan, return_energy=True)
    og = O.Grid([xs, ys, zs], periodic=[True, True, False], period=[float(nx), float(ny), 0.0],
                uniform=[True, True, False])
    ref, ub, vb, wb = O.compute_pi(og, N(U), N(V), N(W), kind=O.TOPHAT, alpha=0.0, ell=ell, engine=O.ENGINE_ND3,
                                   nthreads=NTHREADS)
    err = rel_l2(N(Pi), ref)
    say(f"  {plan.engine}: rel-L2 {err:.2e}")
    assert err <= 1e-5
    Eo = O.energy(og, ub, vb, wb)
    assert abs(E - Eo) <= 1e-6 * Eo


# ------------------------------------------------------------- learned closure, 1024² box ---------

def test_closure_box_1024_pairs_and_gain():
    say("config: learned closure 1024^2")
    n, L = 1024, 1.0e6
    ntrain, ntest = 64, 16
    feat, tau, grid, plan = CL.closure_dataset(list(range(1000, 1000 + ntrain + ntest)), N=n, ell_cells=8.0, device=0)
    # a sampled field's exact pairs against the oracle's filter and derivatives (the grid's default
    # engine, the separable Gaussian)
    k = ntrain + 3
    u, v, xs = S.cartesian_turbulence(N=n, L=L, seed=1000 + k)
    og = O.Grid([xs, xs], periodic=[True, True], period=[L, L], uniform=[True, True])
    f = O.filter_fields(og, [u, v, u * u, u * v, v * v], kind=O.GAUSSIAN, alpha=6.0, ell=plan.scale,
                        engine=O.ENGINE_SEPARABLE)
    ub, vb = f[0], f[1]
    tau_ref = np.stack([f[2] - ub * ub, f[3] - ub * vb, f[4] - vb * vb])
    grad_ref = np.stack([O.deriv(og, 0, ub), O.deriv(og, 1, ub), O.deriv(og, 0, vb), O.deriv(og, 1, vb)])
    assert rel_l2(N(tau[k]), tau_ref) <= 1e-6
    assert rel_l2(N(feat[k]), grad_ref) <= 1e-6
    # the configured training run and its gain over the gradient model on held-out seeds
    s2 = CL.gradient_model_sigma2(plan.kernel, plan.scale)
    model = CL.fit_closure(feat[:ntrain], tau[:ntrain], s2, steps=2000, batch=16, crop=64, seed=0)
    fte, tte = feat[ntrain:].contiguous(), tau[ntrain:].contiguous()
    exact = CL.pi_from_tau(fte, tte)
    corr, rel = CL.pi_errors(model.predict_pi(fte), exact)
    gcorr, grel = CL.pi_errors(CL.pi_from_tau(fte, CL.gradient_model_tau(fte, s2)), exact)
    say(f"  held-out Π: learned corr {corr:.4f} rel-L2 {rel:.4f}; gradient model corr {gcorr:.4f} rel-L2 {grel:.4f}")
    assert corr > 0.99 and rel < 0.2
    assert grel > 0.5 and rel < 0.25 * grel
    # the exact-Π check keeps the learned closure on a held-out field (error below the 20 % threshold)
    u2, v2, _ = S.cartesian_turbulence(N=n, L=L, seed=1000 + ntrain + 1)
    ut = torch.tensor(u2, dtype=torch.float32, device=DEV)
    vt = torch.tensor(v2, dtype=torch.float32, device=DEV)
    checked = CL.LearnedClosure(model, grid, plan, threshold=0.2)
    out = torch.empty_like(ut)
    _, used_exact = checked.compute_pi_(out, ut, vt)
    assert not used_exact and checked.last_error < 0.2

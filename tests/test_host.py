"""Host-side checks that need no GPU: the C ABI library loads and exports every symbol include/cgef.h
declares, the host-only entry points behave as the reference's, and the Python mirror's argument
handling (kernels, spectrum policies, grids) matches the reference."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import cgef
from cgef import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "cgef.h")).read()
    return sorted(set(re.findall(r"\b(cgef_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTED)
    assert cgef.version().startswith("cgef-mi355x")


def test_spectral_density_host_entry_point():
    # Diagnostics.jl:1319-1333: exact for a quadratic on a uniform grid, one-sided at the ends
    k = np.arange(1.0, 6.0)
    g = cgef.spectral_density(k ** 2, k)
    assert g[2] == pytest.approx(2 * k[2])
    assert g[0] == pytest.approx((4 - 1) / 1.0) and g[-1] == pytest.approx((25 - 16) / 1.0)
    assert np.all(cgef.spectral_density(np.array([3.0]), np.array([1.0])) == 0.0)
    with pytest.raises(cgef.DimensionMismatch):
        cgef.spectral_density_(np.zeros(2), np.zeros(3), np.zeros(3))
    st = _lib.load().cgef_spectral_density(-1, None, None, None)
    assert st == _lib.CGEF_ERR_DIMENSION


def test_kernels_and_spectrum_policy():
    g = cgef.GaussianKernel(6.0)
    assert cgef.kernel_weight(g, 1.0, 1.0) == pytest.approx(math.exp(-6.0))
    assert cgef.kernel_radius(cgef.TopHatKernel(), 8.0) == 4.0
    assert cgef.spectral_transfer(g, 0.0, 5.0) == 1.0
    assert cgef.transfer_monotone(g) and not cgef.transfer_monotone(cgef.TopHatKernel())
    assert cgef.gate_spectrum(g, cgef.StrictSpectrum())
    with pytest.raises(cgef.ArgumentError):
        cgef.gate_spectrum(cgef.TopHatKernel(), cgef.StrictSpectrum())
    assert not cgef.gate_spectrum(cgef.TopHatKernel(), cgef.NoSpectrum())
    with pytest.warns(UserWarning):
        assert cgef.gate_spectrum(cgef.TopHatKernel(), cgef.ForceSpectrum())


def test_plan_filter_rejects_bad_scale_before_touching_a_device():
    x = np.arange(8.0)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), x, x)
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(cgef.ArgumentError):
            cgef.plan_filter(grid, cgef.GaussianKernel(), bad)


def test_spectral_method_checks_the_grid_before_touching_a_device():
    x = np.arange(8.0)
    bounded = cgef.StructuredGrid(cgef.CartesianGeometry(), x, x)
    periodic = cgef.StructuredGrid(cgef.CartesianGeometry(), x, x, periodic=(True, True))
    lam = (np.arange(16) + 0.5) * (2 * np.pi / 16)
    phi = -np.pi / 2 + (np.arange(8) + 0.5) * (np.pi / 8)
    sphere = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi)
    k = cgef.GaussianKernel()
    # FFTWExt.jl:133-136: a transform filter needs both axes periodic
    with pytest.raises(cgef.ArgumentError, match="doubly-periodic"):
        cgef.plan_filter(bounded, k, 2.0, method=cgef.Spectral())
    with pytest.raises(cgef.ArgumentError, match="no transform"):
        cgef.plan_filter(sphere, k, 2.0, method=cgef.Spectral())
    with pytest.raises(cgef.ArgumentError, match="real-space engine"):
        cgef.plan_filter(periodic, k, 2.0, method=cgef.Spectral(), engine="ring")
    with pytest.raises(cgef.ArgumentError, match="method must be"):
        cgef.plan_filter(periodic, k, 2.0, method="spectral")
    # AutoMethod is Spectral only where the transform is exact (Filtering.jl:3000-3019)
    from cgef.filtering import _resolve_method
    assert isinstance(_resolve_method(periodic, cgef.AutoMethod()), cgef.Spectral)
    assert isinstance(_resolve_method(bounded, cgef.AutoMethod()), cgef.RealSpace)
    assert isinstance(_resolve_method(sphere, cgef.AutoMethod()), cgef.RealSpace)
    xn = np.cumsum(np.linspace(1.0, 2.0, 8))
    stretched = cgef.StructuredGrid(cgef.CartesianGeometry(), xn, x, periodic=(True, True), period=(20.0, 8.0))
    assert isinstance(_resolve_method(stretched, cgef.AutoMethod()), cgef.RealSpace)


def test_grid_description():
    lam = (np.arange(16) + 0.5) * (2 * np.pi / 16)
    phi = -np.pi / 2 + (np.arange(8) + 0.5) * (np.pi / 8)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(), lam, phi)
    assert g.periodic == [True, False] and g.uniform == [True, True]
    assert g.field_shape == (8, 16) and g.shape == (16, 8)
    with pytest.raises(cgef.DimensionMismatch):
        cgef.StructuredGrid(cgef.CartesianGeometry(), np.arange(4.0), np.arange(5.0), mask=np.ones((4, 5)))
    z = np.cumsum(1.03 ** np.arange(6))
    g3 = cgef.StructuredGrid(cgef.CartesianGeometry(), np.arange(4.0), np.arange(5.0), z)
    assert g3.uniform == [True, True, False] and g3.field_shape == (6, 5, 4)
    r = cgef.Range(0.0, 0.25, 4)
    assert np.allclose(r.values, [0, 0.25, 0.5, 0.75])


def test_compute_pi_validates_layout_on_host():
    x = np.arange(8.0)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), x, x)
    with pytest.raises(cgef.ArgumentError):
        cgef.filtering.filter_apply_(np.zeros((8, 8)), np.zeros((8, 8)), _FakePlan(grid))


class _FakePlan:
    def __init__(self, grid):
        self.grid, self.device = grid, 0


def test_slice_entry_points_check_lengths_before_touching_a_device():
    # Pipeline.jl:892-906, Filtering.jl:3452-3454: ragged collections are DimensionMismatch
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), np.arange(8.0), np.arange(6.0))
    with pytest.raises(cgef.DimensionMismatch):
        cgef.coarse_grain_slices_([None, None], [None], [None, None], None, [g, g], scales=[1.0],
                                  kernel=cgef.GaussianKernel(6.0))
    with pytest.raises(cgef.DimensionMismatch):
        cgef.filter_slices_([None], [None, None], [None])
    assert cgef.slice_pipeline_costs([g, g], 3) == [144, 144]


def test_coarse_grain_checks_plan_and_result_sizes_before_touching_a_device():
    # Pipeline.jl:443-444 (one plan per scale) and the result shape (Pipeline.jl:1286-1298): a list of
    # plans or a result sized for other scales is a DimensionMismatch before any library call
    import torch
    g = cgef.StructuredGrid(cgef.CartesianGeometry(), np.arange(8.0), np.arange(6.0))
    u = torch.zeros(g.field_shape, dtype=torch.float64)
    res = cgef.pipeline._allocate_result(g, 2, torch.float64, "cpu")
    fake = [_FakePlan(g)] * 3
    with pytest.raises(cgef.DimensionMismatch, match="3 filter plans for 2 scales"):
        cgef.coarse_grain_(res, u, u, None, g, scales=[1.0, 2.0], kernel=cgef.GaussianKernel(6.0), filter_plans=fake)
    with pytest.raises(cgef.DimensionMismatch, match="result.Π has shape"):
        cgef.coarse_grain_(res, u, u, None, g, scales=[1.0, 2.0, 3.0], kernel=cgef.GaussianKernel(6.0))
    batch = cgef.CoarseGrainBatchResult.allocate(g, 2, (4,), torch.float64, "cpu")
    ub = torch.zeros((4,) + g.field_shape, dtype=torch.float64)
    with pytest.raises(cgef.DimensionMismatch, match="batch.Π has shape"):
        cgef.coarse_grain_batch_(batch, ub, ub, None, g, scales=[1.0, 2.0, 3.0], kernel=cgef.GaussianKernel(6.0))
    with pytest.raises(cgef.DimensionMismatch, match="3 filter plans for 2 scales"):
        cgef.coarse_grain_batch_(batch, ub, ub, None, g, scales=[1.0, 2.0], kernel=cgef.GaussianKernel(6.0),
                                 filter_plans=fake)
    other = cgef.CoarseGrainBatchResult.allocate(g, 2, (3,), torch.float64, "cpu")
    other.batch_size = (4,)   # Π allocated for 3 slices, the call claims 4
    with pytest.raises(cgef.DimensionMismatch, match="batch.Π has shape"):
        cgef.coarse_grain_batch_(other, ub, ub, None, g, scales=[1.0, 2.0], kernel=cgef.GaussianKernel(6.0))


def test_closure_entry_points_validate_on_the_host():
    """The learned closure's C-ABI entry points reject bad arguments before any device call."""
    import ctypes as C

    from cgef import closure as CL
    L = CL._L()
    vp = C.c_void_p(1)   # never dereferenced: every call below fails its host checks first
    with pytest.raises(_lib.ArgumentError):
        _lib.check(L.cgef_conv5(None, 2, 0, 1, 4, 20, 20, vp, 64, vp, 64, 0, vp, 2, None, None, None))
    with pytest.raises(_lib.ArgumentError):     # pad must be 0 or 4
        _lib.check(L.cgef_conv5(vp, 2, 0, 1, 4, 20, 20, vp, 64, vp, 64, 2, vp, 2, None, None, None))
    with pytest.raises(_lib.DimensionMismatch):  # smaller than the stencil
        _lib.check(L.cgef_conv5(vp, 2, 0, 1, 4, 3, 20, vp, 64, vp, 64, 0, vp, 2, None, None, None))
    with pytest.raises(_lib.DimensionMismatch):  # layout stride not a whole number of channel blocks
        _lib.check(L.cgef_conv5(vp, 2, 0, 1, 4, 20, 20, vp, 60, vp, 60, 0, vp, 2, None, None, None))
    with pytest.raises(_lib.ArgumentError):     # FP32 input cannot take GELU on load
        _lib.check(L.cgef_conv5_mfma(vp, 0, 1, 1, 20, 20, vp, vp, 0, vp, 2, None, None, None))
    with pytest.raises(_lib.ArgumentError):     # the activation comes with a BF16 output only
        _lib.check(L.cgef_conv5_mfma(vp, 2, 0, 1, 20, 20, vp, vp, 0, vp, 0, None, vp, None))
    with pytest.raises(_lib.ArgumentError):     # Adam steps count from 1
        _lib.check(L.cgef_adam_update(vp, vp, vp, vp, 10, 1e-3, 0.9, 0.999, 1e-8, 0, 1, 1, None, 0, None, 0, None))
    with pytest.raises(_lib.DimensionMismatch):  # the matrix-core weight gradient takes 64 -> 64 layers
        CL.wgrad_scratch_floats(16, 4, 64, 76, 76, mfma=True)


def test_closure_wgrad_scratch_sizes():
    from cgef import closure as CL
    # VALU: 16x16 tiles, channel groups of 16, about 2048 blocks; matrix cores: 8x16 tiles, ~1024 blocks
    n, ci, co, ho, wo = 16, 64, 64, 72, 72
    tiles = 5 * 5
    want = max(1, 2048 // (n * 4 * 4))
    tpb = -(-tiles // min(want, tiles))
    assert CL.wgrad_scratch_floats(n, ci, co, ho, wo) == n * (-(-tiles // tpb)) * (co * ci * 25 + co)
    tiles_m = 5 * 9
    want_m = max(1, 1024 // (5 * n))
    tpb_m = -(-tiles_m // min(want_m, tiles_m))
    assert CL.wgrad_scratch_floats(n, ci, co, ho, wo, mfma=True) == n * (-(-tiles_m // tpb_m)) * (co * ci * 25 + co)
    assert CL.gradient_model_sigma2(cgef.GaussianKernel(6.0), 8.0) == pytest.approx(64.0 / 12.0)
    assert CL.gradient_model_sigma2(cgef.TopHatKernel(), 8.0) == pytest.approx(4.0)

"""Pipeline (Pipeline.jl): multi-scale Π + E(ℓ) + filtering-spectrum sweeps over one grid, single
snapshot (coarse_grain!) or a batch of snapshots computed batch-native on the GPU
(coarse_grain_batch!, the GPU driver of Pipeline.jl:519-563)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .diagnostics import (PiWorkspace, StrictSpectrum, NoSpectrum, compute_pi_, gate_spectrum, spectral_density_)
from .filtering import ZeroFill, _check_field, batch_dims, plan_filter
from .grids import StructuredGrid
from .kernels import TopHatKernel


@dataclass
class CoarseGrainResult:
    """Pipeline.jl:270-276. Π is one tensor (scale, …grid) — the reference's (grid…, Nscales)."""
    scales: np.ndarray
    Pi: torch.Tensor
    cumulative_energy: np.ndarray
    wavenumber: np.ndarray
    filtering_spectrum: np.ndarray

    @property
    def Π(self):
        return self.Pi


def _allocate_result(grid, nscales, dtype, device):
    return CoarseGrainResult(np.zeros(nscales), torch.zeros((nscales,) + grid.field_shape, dtype=dtype, device=device),
                             np.zeros(nscales), np.zeros(nscales), np.zeros(nscales))


def _fill_spectrum(g, Cv, k, policy):
    if isinstance(policy, NoSpectrum):
        g[:] = np.nan
    else:
        spectral_density_(g, Cv, k)


def coarse_grain_(result: CoarseGrainResult, u, v, w, grid: StructuredGrid, *, scales, kernel=TopHatKernel(),
                  workspace=None, filter_plans=None, deriv_plan=None, backend=None, mask_strategy=ZeroFill(),
                  L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0):
    """Pipeline.jl:412-468 — per scale: Π into result.Pi[s] and E(ℓ) from the same filtering."""
    ns = len(scales)
    if result.Pi.shape[0] != ns or tuple(result.Pi.shape[1:]) != grid.field_shape:
        raise _lib.DimensionMismatch(f"result.Π has shape {tuple(result.Pi.shape)}, expected {(ns,) + grid.field_shape}")
    gate_spectrum(kernel, spectrum)
    plans = filter_plans if filter_plans is not None else [
        plan_filter(grid, kernel, float(s), mask_strategy=mask_strategy, backend=backend) for s in scales]
    for s_idx, s in enumerate(scales):
        result.scales[s_idx] = float(s)
        _, E = compute_pi_(result.Pi[s_idx], u, v, w, grid, kernel, float(s), workspace=workspace,
                           filter_plan=plans[s_idx], mask_strategy=mask_strategy, rho0=rho0, return_energy=True)
        result.cumulative_energy[s_idx] = E
    result.wavenumber[:] = float(L) / result.scales
    _fill_spectrum(result.filtering_spectrum, result.cumulative_energy, result.wavenumber, spectrum)
    return result


def coarse_grain(u, v, w, grid: StructuredGrid, *, scales, kernel=TopHatKernel(), backend=None,
                 mask_strategy=ZeroFill(), L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0):
    """Pipeline.jl:338-358."""
    res = _allocate_result(grid, len(scales), u.dtype, u.device)
    return coarse_grain_(res, u, v, w, grid, scales=scales, kernel=kernel, backend=backend,
                         mask_strategy=mask_strategy, L=L, spectrum=spectrum, rho0=rho0)


@dataclass
class CoarseGrainBatchResult:
    """Pipeline.jl:593-651. Π is (batch…, scale, …grid) — the reference's (grid…, Nscales, batch…);
    the per-scale vectors are (batch…, scale)."""
    Pi: torch.Tensor
    scales: np.ndarray
    cumulative_energy: np.ndarray
    wavenumber: np.ndarray
    filtering_spectrum: np.ndarray
    batch_size: tuple

    @classmethod
    def allocate(cls, grid: StructuredGrid, nscales: int, batch_size, dtype=torch.float32, device="cuda"):
        bs = tuple(int(b) for b in batch_size)
        Pi = torch.zeros(bs + (nscales,) + grid.field_shape, dtype=dtype, device=device)
        z = lambda: np.zeros(bs + (nscales,))  # noqa: E731
        return cls(Pi, z(), z(), z(), z(), bs)


def coarse_grain_batch_(batch: CoarseGrainBatchResult, u, v, w, grid: StructuredGrid, *, scales,
                        kernel=TopHatKernel(), workspaces=None, filter_plans=None, deriv_plans=None, backend=None,
                        mask_strategy=ZeroFill(), L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0):
    """Pipeline.jl:688-724 with the GPU batch driver (519-563): one compute_Π! per scale over the
    whole batch (a single launch sequence covering every snapshot), E(ℓ) per slice."""
    dev = u.device.index if u.is_cuda else -1
    bs = batch.batch_size
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None:
            _check_field(t, name, grid, dev)
            if batch_dims(t, grid) != bs:
                raise _lib.DimensionMismatch(f"{name}'s batch dimensions {batch_dims(t, grid)} do not match batch size {bs}")
    gate_spectrum(kernel, spectrum)
    ns = len(scales)
    ws = workspaces[0] if workspaces else PiWorkspace(grid, bs, u.dtype, dev)
    if ws.nbatch < (int(np.prod(bs)) if bs else 1):
        raise _lib.DimensionMismatch("coarse_grain_batch! needs a batch-sized workspace")
    plans = filter_plans if filter_plans is not None else [
        plan_filter(grid, kernel, float(s), mask_strategy=mask_strategy, backend=backend) for s in scales]
    tmp = torch.empty(bs + grid.field_shape, dtype=u.dtype, device=u.device)
    nd = len(bs)
    for s_idx, s in enumerate(scales):
        batch.scales[..., s_idx] = float(s)
        _, E = compute_pi_(tmp, u, v, w, grid, kernel, float(s), workspace=ws, filter_plan=plans[s_idx],
                           mask_strategy=mask_strategy, rho0=rho0, return_energy=True)
        batch.Pi.select(nd, s_idx).copy_(tmp)
        batch.cumulative_energy[..., s_idx] = E
    batch.wavenumber[...] = float(L) / batch.scales
    flat_c = batch.cumulative_energy.reshape(-1, ns)
    flat_k = batch.wavenumber.reshape(-1, ns)
    flat_g = batch.filtering_spectrum.reshape(-1, ns)
    for q in range(flat_c.shape[0]):
        row = np.zeros(ns)
        _fill_spectrum(row, flat_c[q], flat_k[q], spectrum)
        flat_g[q] = row
    return batch

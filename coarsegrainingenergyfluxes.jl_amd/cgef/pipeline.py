"""Pipeline (Pipeline.jl): multi-scale Π + E(ℓ) + filtering-spectrum sweeps over one grid, single
snapshot (coarse_grain!) or a batch of snapshots computed batch-native on the GPU
(coarse_grain_batch!, the GPU driver of Pipeline.jl:519-563)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .diagnostics import PiWorkspace, StrictSpectrum, NoSpectrum, gate_spectrum, spectral_density_
from .filtering import ZeroFill, _check_field, _dtype_code, _stream, batch_dims, plan_filter
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


def _run_coarse_grain(Pi, u, v, w, grid, plans, workspace, nbatch, rho0, E_out):
    """One cgef_coarse_grain call: Π for every plan (scale) over `nbatch` slices, E(ℓ) into E_out
    (slice-major, scale fastest)."""
    dev = u.device.index if u.is_cuda else -1
    for name, t in (("Π", Pi), ("u", u), ("v", v), ("w", w)):
        if t is not None:
            _check_field(t, name, grid, dev)
    if grid.ndim == 2 and w is not None:
        raise _lib.ArgumentError("the 2.5-D (u, v, w) layer Π is not provided by this backend; pass w=None")
    if any(p.grid is not grid or p.device != dev for p in plans):
        raise _lib.ArgumentError("every filter plan must belong to this grid and device")
    if Pi.dtype != u.dtype or v.dtype != u.dtype or (w is not None and w.dtype != u.dtype):
        raise _lib.ArgumentError("Π and the velocity components must share a dtype")
    # the library writes nbatch × len(plans) × npoints values into Π: check before any launch
    if Pi.numel() != nbatch * len(plans) * grid.npoints:
        raise _lib.DimensionMismatch(f"Π holds {Pi.numel()} values; {nbatch} slice(s) × {len(plans)} scale(s) × "
                                     f"{grid.npoints} points need {nbatch * len(plans) * grid.npoints}")
    if u.numel() != nbatch * grid.npoints:
        raise _lib.DimensionMismatch(f"u holds {u.numel()} values, expected {nbatch} slice(s) × {grid.npoints} points")
    arr = (C.c_void_p * len(plans))(*[p.h for p in plans])
    Eb = np.zeros(len(plans) * nbatch, dtype=np.float64)
    _lib.check(_lib.load().cgef_coarse_grain(
        arr, len(plans), workspace.h if workspace is not None else None, C.c_void_p(Pi.data_ptr()),
        C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()) if w is not None else None,
        nbatch, _dtype_code(u), C.c_double(rho0), Eb.ctypes.data_as(C.POINTER(C.c_double)), _stream(dev)))
    E_out[...] = Eb.reshape(np.shape(E_out))
    return Pi


def _check_plans(filter_plans, scales):
    """Pipeline.jl:443-444 builds one plan per scale; a caller-supplied list must match the scales."""
    if filter_plans is not None and len(filter_plans) != len(scales):
        raise _lib.DimensionMismatch(f"{len(filter_plans)} filter plans for {len(scales)} scales")


def _fill_spectrum(g, Cv, k, policy):
    if isinstance(policy, NoSpectrum):
        g[:] = np.nan
    else:
        spectral_density_(g, Cv, k)


def coarse_grain_(result: CoarseGrainResult, u, v, w, grid: StructuredGrid, *, scales, kernel=TopHatKernel(),
                  workspace=None, filter_plans=None, deriv_plan=None, backend=None, mask_strategy=ZeroFill(),
                  L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0, method=None):
    """Pipeline.jl:412-468 — per scale: Π into result.Pi[s] and E(ℓ) from the same filtering.
    `method=None` takes plan_filter's default for the grid (Pipeline.jl:366-371)."""
    ns = len(scales)
    if result.Pi.shape[0] != ns or tuple(result.Pi.shape[1:]) != grid.field_shape:
        raise _lib.DimensionMismatch(f"result.Π has shape {tuple(result.Pi.shape)}, expected {(ns,) + grid.field_shape}")
    if len(result.cumulative_energy) != ns:
        raise _lib.DimensionMismatch(f"result holds {len(result.cumulative_energy)} energies for {ns} scales")
    _check_plans(filter_plans, scales)
    gate_spectrum(kernel, spectrum)
    plans = filter_plans if filter_plans is not None else [
        plan_filter(grid, kernel, float(s), mask_strategy=mask_strategy, backend=backend, method=method)
        for s in scales]
    _run_coarse_grain(result.Pi, u, v, w, grid, plans, workspace, 1, rho0, result.cumulative_energy)
    result.scales[:] = [float(s) for s in scales]
    result.wavenumber[:] = float(L) / result.scales
    _fill_spectrum(result.filtering_spectrum, result.cumulative_energy, result.wavenumber, spectrum)
    return result


def coarse_grain(u, v, w, grid: StructuredGrid, *, scales, kernel=TopHatKernel(), backend=None,
                 mask_strategy=ZeroFill(), L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0, method=None):
    """Pipeline.jl:338-358."""
    res = _allocate_result(grid, len(scales), u.dtype, u.device)
    return coarse_grain_(res, u, v, w, grid, scales=scales, kernel=kernel, backend=backend,
                         mask_strategy=mask_strategy, L=L, spectrum=spectrum, rho0=rho0, method=method)


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
                        mask_strategy=ZeroFill(), L: float = 1.0, spectrum=StrictSpectrum(), rho0: float = 1.0,
                        method=None):
    """Pipeline.jl:688-724 with the GPU batch driver (519-563): one compute_Π! per scale over the
    whole batch (a single launch sequence covering every snapshot), E(ℓ) per slice."""
    bs = batch.batch_size
    ns = len(scales)
    # sizes first, before anything reaches the device (the library writes batch × scales × points)
    if batch_dims(batch.Pi, grid) != bs + (ns,) or tuple(batch.Pi.shape[len(bs) + 1:]) != grid.field_shape:
        raise _lib.DimensionMismatch(f"batch.Π has shape {tuple(batch.Pi.shape)}, expected {bs + (ns,) + grid.field_shape}")
    if np.shape(batch.cumulative_energy) != bs + (ns,):
        raise _lib.DimensionMismatch(f"batch holds energies of shape {np.shape(batch.cumulative_energy)}, "
                                     f"expected {bs + (ns,)}")
    _check_plans(filter_plans, scales)
    dev = u.device.index if u.is_cuda else -1
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None:
            _check_field(t, name, grid, dev)
            if batch_dims(t, grid) != bs:
                raise _lib.DimensionMismatch(f"{name}'s batch dimensions {batch_dims(t, grid)} do not match batch size {bs}")
    gate_spectrum(kernel, spectrum)
    ws = workspaces[0] if workspaces else PiWorkspace(grid, bs, u.dtype, dev)
    if ws.nbatch < (int(np.prod(bs)) if bs else 1):
        raise _lib.DimensionMismatch("coarse_grain_batch! needs a batch-sized workspace")
    plans = filter_plans if filter_plans is not None else [
        plan_filter(grid, kernel, float(s), mask_strategy=mask_strategy, backend=backend, method=method)
        for s in scales]
    nb = int(np.prod(bs)) if bs else 1
    E = np.zeros(nb * ns)
    _run_coarse_grain(batch.Pi, u, v, w, grid, plans, ws, nb, rho0, E)
    batch.cumulative_energy[...] = E.reshape(batch.cumulative_energy.shape)
    batch.scales[...] = np.asarray([float(s) for s in scales])
    batch.wavenumber[...] = float(L) / batch.scales
    flat_c = batch.cumulative_energy.reshape(-1, ns)
    flat_k = batch.wavenumber.reshape(-1, ns)
    flat_g = batch.filtering_spectrum.reshape(-1, ns)
    for q in range(flat_c.shape[0]):
        row = np.zeros(ns)
        _fill_spectrum(row, flat_c[q], flat_k[q], spectrum)
        flat_g[q] = row
    return batch


def coarse_grain_profile_(batch: CoarseGrainBatchResult, u, v, w, grid: StructuredGrid, **kwargs):
    """Pipeline.jl:991-1010 coarse_grain_profile!: independent per-level 2-D sweeps on a shared
    horizontal grid, i.e. coarse_grain_batch! with the level axis as the batch axis (u, v[, w] are
    (Nlevels, Ny, Nx); Π is (Nlevels, Nscales, Ny, Nx); per-level energies are not summed)."""
    if grid.ndim != 2:
        raise _lib.ArgumentError("coarse_grain_profile! takes a 2-D horizontal grid")
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None and t.dim() != 3:
            raise _lib.DimensionMismatch(f"{name} must be (Nlevels, Ny, Nx), got {tuple(t.shape)}")
    return coarse_grain_batch_(batch, u, v, w, grid, **kwargs)


def coarse_grain_profile(u, v, w, grid: StructuredGrid, *, scales, **kwargs):
    """Pipeline.jl:1012-1034: allocating coarse_grain_profile!, batch size (Nlevels,)."""
    batch = CoarseGrainBatchResult.allocate(grid, len(scales), (u.shape[0],), u.dtype, u.device)
    return coarse_grain_profile_(batch, u, v, w, grid, scales=scales, **kwargs)


def slice_pipeline_costs(grids, nscales: int):
    """Pipeline.jl:838-839: relative per-slice sweep cost (points × scales) for longest-first order."""
    return [int(g.npoints) * int(nscales) for g in grids]


def coarse_grain_slices_(results, us, vs, ws, grids, *, scales, kernel=TopHatKernel(), workspaces=None,
                         filter_plans=None, deriv_plans=None, backend=None, mask_strategy=ZeroFill(), L: float = 1.0,
                         spectrum=StrictSpectrum(), rho0: float = 1.0, method=None):
    """Pipeline.jl:866-919 coarse_grain_slices!: a full coarse_grain! sweep per slice, each slice with
    its own grid; scratch is per slice. Slices run longest-first (slice_pipeline_costs)."""
    n = len(results)
    if not (len(us) == n and len(vs) == n and len(grids) == n):
        raise _lib.DimensionMismatch(
            f"coarse_grain_slices! got {n} results, {len(us)} u, {len(vs)} v and {len(grids)} grids")
    if ws is not None and len(ws) != n:
        raise _lib.DimensionMismatch(f"coarse_grain_slices! got {n} results and {len(ws)} w slices")
    for name, pool in (("workspaces", workspaces), ("filter_plans", filter_plans), ("deriv_plans", deriv_plans)):
        if pool is not None and len(pool) != n:
            raise _lib.DimensionMismatch(
                f"coarse_grain_slices! got {n} results and {len(pool)} {name} entries — ragged scratch is per slice")
    gate_spectrum(kernel, spectrum)
    costs = slice_pipeline_costs(grids, len(scales))
    for t in sorted(range(n), key=lambda q: -costs[q]):
        coarse_grain_(results[t], us[t], vs[t], None if ws is None else ws[t], grids[t], scales=scales,
                      kernel=kernel, workspace=None if workspaces is None else workspaces[t],
                      filter_plans=None if filter_plans is None else filter_plans[t], backend=backend,
                      mask_strategy=mask_strategy, L=L, spectrum=spectrum, rho0=rho0, method=method)
    return results

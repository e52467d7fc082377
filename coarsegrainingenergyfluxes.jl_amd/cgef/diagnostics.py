"""Diagnostics (Diagnostics.jl): Π = −S̄:τ, the energy reductions and the filtering spectrum, with
the filtering and the per-point tensor physics on the GPU (libcgef_hip.so)."""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np
import torch

from . import _lib
from .filtering import (FilterPlan, ZeroFill, _check_field, _dtype_code, _nslices, _stream, batch_dims,
                        plan_filter)
from .grids import StructuredGrid
from .kernels import check_spectrum_kernel, transfer_monotone


class StrictSpectrum:
    """Diagnostics.jl:39 — refuse a non-monotone kernel."""


class ForceSpectrum:
    """Diagnostics.jl:48 — compute anyway, warn once."""


class NoSpectrum:
    """Diagnostics.jl:57 — skip the density (NaN)."""


_warned = False


def gate_spectrum(kernel, policy) -> bool:
    """Diagnostics.jl:68-81."""
    global _warned
    if isinstance(policy, StrictSpectrum):
        check_spectrum_kernel(kernel)
        return True
    if isinstance(policy, NoSpectrum):
        return False
    if isinstance(policy, ForceSpectrum):
        if not transfer_monotone(kernel) and not _warned:
            warnings.warn(f"ForceSpectrum with {type(kernel).__name__}: its |Ĝ(k)|² is not monotone decreasing, so "
                          "the filtering spectral density is not guaranteed non-negative.")
            _warned = True
        return True
    raise _lib.ArgumentError("unknown spectrum policy")


class PiWorkspace:
    """ΠWorkspace (Diagnostics.jl:83-194): device scratch for one grid, sized for a batch shape and
    dtype. After compute_pi_ it holds the filtered velocity in u_filt / v_filt / w_filt (local
    east/north on the sphere), as the reference's workspace does."""

    def __init__(self, grid: StructuredGrid, batch_size=(), dtype=torch.float32, device: int | None = None):
        dev = torch.cuda.current_device() if device is None else device
        self.grid, self.batch_size, self.dtype, self.device = grid, tuple(batch_size), dtype, dev
        nb = int(np.prod(self.batch_size)) if self.batch_size else 1
        self.nbatch = nb
        out = C.c_void_p()
        code = _lib.F32 if dtype == torch.float32 else _lib.F64
        _lib.check(_lib.load().cgef_workspace_create(grid.handle(dev), nb, code, C.byref(out)))
        self.h = out
        shape = self.batch_size + grid.field_shape
        self.u_filt = torch.zeros(shape, dtype=dtype, device=f"cuda:{dev}")
        self.v_filt = torch.zeros_like(self.u_filt)
        self.w_filt = torch.zeros_like(self.u_filt)

    def __del__(self):
        try:
            _lib.load().cgef_workspace_destroy(self.h)
        except Exception:
            pass


def _validate_field_sizes(grid, Pi, u, v, w, dev):
    # Diagnostics.jl:207-229 — leading (grid) axes match, batch axes common to every field
    for name, t in (("Π", Pi), ("u", u), ("v", v), ("w", w)):
        if t is not None:
            _check_field(t, name, grid, dev)
    bsz = batch_dims(Pi, grid)
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None and batch_dims(t, grid) != bsz:
            raise _lib.DimensionMismatch(f"{name}'s batch axes {batch_dims(t, grid)} do not match Π's {bsz}")
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None and t.dtype != Pi.dtype:
            raise _lib.ArgumentError(f"{name} must share Π's dtype")


def compute_pi_(Pi, u, v, w, grid: StructuredGrid, kernel, scale, *, workspace: PiWorkspace | None = None,
                filter_plan: FilterPlan | None = None, backend=None, mask_strategy=ZeroFill(), deriv_plan=None,
                rho0: float = 1.0, return_energy: bool = False, method=None):
    """compute_Π! (Diagnostics.jl:389-416 2-D, 792-857 3-D): Π = −S̄:τ at scale ℓ (times `rho0`,
    default 1, which is the reference: it applies no density). Leading axes of the fields beyond the
    grid's are a batch. With return_energy=True also returns E(ℓ) per slice (energy_from_filtered)."""
    dev = Pi.device.index if Pi.is_cuda else -1
    plan = filter_plan if filter_plan is not None else plan_filter(
        grid, kernel, scale, mask_strategy=mask_strategy, backend=backend, method=method)
    if plan.grid is not grid:
        raise _lib.ArgumentError("filter_plan was built for a different grid")
    if dev != plan.device:
        raise _lib.ArgumentError(f"Π must live on cuda:{plan.device}")
    if grid.ndim == 2 and w is not None:
        raise _lib.ArgumentError("the 2.5-D (u, v, w) layer Π is not provided by this backend; pass w=None")
    if grid.ndim == 3 and w is None:
        raise _lib.ArgumentError("3-D Π needs w")
    _validate_field_sizes(grid, Pi, u, v, w, dev)
    nb = _nslices(Pi, grid)
    ws = workspace
    if ws is not None:
        if ws.grid is not grid or ws.dtype != Pi.dtype or ws.nbatch < nb or ws.device != dev:
            raise _lib.DimensionMismatch("workspace does not match this grid / batch / dtype / device")
    E = np.zeros(nb, dtype=np.float64) if return_energy else None
    Ep = E.ctypes.data_as(C.POINTER(C.c_double)) if E is not None else None
    ub = vb = wb = None
    if ws is not None:
        bs = batch_dims(Pi, grid)
        if ws.u_filt.shape[: len(ws.batch_size)] != bs:
            ws.u_filt = torch.zeros(bs + grid.field_shape, dtype=Pi.dtype, device=Pi.device)
            ws.v_filt = torch.zeros_like(ws.u_filt)
            ws.w_filt = torch.zeros_like(ws.u_filt)
        ub, vb = C.c_void_p(ws.u_filt.data_ptr()), C.c_void_p(ws.v_filt.data_ptr())
        wb = C.c_void_p(ws.w_filt.data_ptr())
    _lib.check(_lib.load().cgef_compute_pi(
        plan.h, ws.h if ws is not None else None, C.c_void_p(Pi.data_ptr()), C.c_void_p(u.data_ptr()),
        C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()) if w is not None else None, nb, _dtype_code(Pi),
        C.c_double(rho0), Ep, ub, vb, wb, _stream(dev)))
    if return_energy:
        return Pi, E.reshape(batch_dims(Pi, grid)) if batch_dims(Pi, grid) else float(E[0])
    return Pi


compute_Π_ = compute_pi_


def active_area(grid: StructuredGrid, device: int | None = None) -> float:
    """Diagnostics.jl:1051-1059."""
    dev = torch.cuda.current_device() if device is None else device
    out = C.c_double()
    _lib.check(_lib.load().cgef_active_area(grid.handle(dev), C.byref(out)))
    return out.value


def energy_from_filtered(ws: PiWorkspace, grid: StructuredGrid, has_w: bool, total_area: float | None = None):
    """Diagnostics.jl:1086-1121 — ½⟨|ū|²⟩ per slice from the workspace's filtered velocity."""
    u, v = ws.u_filt, ws.v_filt
    nb = _nslices(u, grid)
    E = np.zeros(nb, dtype=np.float64)
    _lib.check(_lib.load().cgef_energy(grid.handle(ws.device), C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                                        C.c_void_p(ws.w_filt.data_ptr()) if has_w else None, nb, _dtype_code(u),
                                        E.ctypes.data_as(C.POINTER(C.c_double)), _stream(ws.device)))
    bs = batch_dims(u, grid)
    return E.reshape(bs) if bs else float(E[0])


def cumulative_energy_(spectrum, u, v, w, grid, kernel, scales, *, workspace=None, filter_plans=None, backend=None,
                       mask_strategy=ZeroFill(), method=None):
    """Diagnostics.jl:1144-1217 — filter the raw (u, v[, w]) per scale and reduce ½⟨|ū|²⟩."""
    dev = u.device.index if u.is_cuda else -1
    for name, t in (("u", u), ("v", v), ("w", w)):
        if t is not None:
            _check_field(t, name, grid, dev)
            if tuple(t.shape) != grid.field_shape:
                raise _lib.DimensionMismatch(f"{name} has size {tuple(t.shape)}, grid expects {grid.field_shape}")
    if len(spectrum) != len(scales):
        raise _lib.DimensionMismatch(f"spectrum has length {len(spectrum)}, expected {len(scales)} (= length(scales))")
    if len(scales) == 0:
        return spectrum
    plans = filter_plans if filter_plans is not None else [
        plan_filter(grid, kernel, float(s), mask_strategy=mask_strategy, backend=backend, method=method)
        for s in scales]
    arr = (C.c_void_p * len(plans))(*[p.h for p in plans])
    E = np.zeros(len(scales), dtype=np.float64)
    _lib.check(_lib.load().cgef_cumulative_energy(
        arr, len(plans), workspace.h if workspace is not None else None, C.c_void_p(u.data_ptr()),
        C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()) if w is not None else None, 1, _dtype_code(u),
        E.ctypes.data_as(C.POINTER(C.c_double)), _stream(dev)))
    spectrum[:] = E
    return spectrum


def cumulative_energy(u, v, w, grid, kernel, scales, *, backend=None, mask_strategy=ZeroFill(), method=None):
    """Diagnostics.jl:1238-1251."""
    return cumulative_energy_(np.zeros(len(scales)), u, v, w, grid, kernel, scales, backend=backend,
                              mask_strategy=mask_strategy, method=method)


def spectral_density_(g, Cv, k):
    """Diagnostics.jl:1319-1333 — central differences inside, one-sided at the ends, zeros if n < 2."""
    Cv = np.ascontiguousarray(Cv, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    if len(g) != len(Cv):
        raise _lib.DimensionMismatch(f"g has length {len(g)}, expected {len(Cv)} (= length(C))")
    out = np.zeros(len(Cv), dtype=np.float64)
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().cgef_spectral_density(len(Cv), Cv.ctypes.data_as(dp), k.ctypes.data_as(dp),
                                                  out.ctypes.data_as(dp)))
    g[:] = out
    return g


def spectral_density(Cv, k):
    return spectral_density_(np.zeros(len(Cv)), Cv, k)


def filtering_spectrum(u, v, w, grid, kernel, scales, *, L: float = 1.0, backend=None, mask_strategy=ZeroFill(),
                       policy=StrictSpectrum(), method=None):
    """Diagnostics.jl:1294-1309 — (k_ℓ = L/ℓ, Ẽ(k_ℓ))."""
    compute = gate_spectrum(kernel, policy)
    cum = cumulative_energy(u, v, w, grid, kernel, scales, backend=backend, mask_strategy=mask_strategy,
                            method=method)
    kl = float(L) / np.asarray(scales, dtype=np.float64)
    if not compute:
        return kl, np.full(len(kl), np.nan)
    return kl, spectral_density(cum, kl)

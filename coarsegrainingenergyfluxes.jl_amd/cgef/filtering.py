"""Filtering (Filtering.jl): mask strategies, plan_filter and the filter_apply family, executed by
libcgef_hip.so on the GPU that holds the fields."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import ArgumentError, DimensionMismatch
from .grids import StructuredGrid
from .kernels import AbstractFilterKernel


@dataclass(frozen=True)
class ZeroFill:
    """Filtering.jl:40 — masked cells count in the denominator as zero."""
    code: int = _lib.ZEROFILL


@dataclass(frozen=True)
class Deformable:
    """Filtering.jl:49 — masked cells dropped from numerator and denominator."""
    code: int = _lib.DEFORMABLE


@dataclass(frozen=True)
class RealSpace:
    """Filtering.jl:64-65 — physical-space convolution (the reference's default method)."""


@dataclass(frozen=True)
class Spectral:
    """Filtering.jl:67-73 — the kernel applied as a multiply on the transformed field. Provided here for
    a doubly periodic 2-D Cartesian grid (the reference's FFTW extension, FFTWExt.jl:121-208)."""


@dataclass(frozen=True)
class AutoMethod:
    """Filtering.jl:75-85 — Spectral where a periodic transform is the exact filter (every axis
    periodic and uniform), RealSpace otherwise."""


@dataclass(frozen=True)
class GPUBackend:
    """The reference's ComputationalBackends.GPUBackend{B}: here, a HIP device ordinal."""
    device: int = 0


ENGINES = {"auto": _lib.ENGINE_AUTO, "banded": _lib.ENGINE_BANDED, "separable": _lib.ENGINE_SEPARABLE,
           "baseline": _lib.ENGINE_BASELINE, "ring": _lib.ENGINE_RING, "prefixsum": _lib.ENGINE_PREFIX,
           "spectral": _lib.ENGINE_SPECTRAL}
ENGINE_NAMES = {v: k for k, v in ENGINES.items()}


def _device_of(backend) -> int:
    if backend is None:
        return torch.cuda.current_device() if torch.cuda.is_available() else 0
    if isinstance(backend, GPUBackend):
        return backend.device
    raise ArgumentError(f"backend = {type(backend).__name__} is not served by this GPU backend; use GPUBackend()")


class FilterPlan:
    """PhysicalFilterPlan (Filtering.jl:2894-2901): a footprint built once and resident on one GPU."""

    def __init__(self, grid: StructuredGrid, kernel: AbstractFilterKernel, scale: float, strategy, device: int,
                 engine: str):
        if engine not in ENGINES:
            raise ArgumentError(f"engine must be one of {sorted(ENGINES)}")
        L = _lib.load()
        self.grid, self.kernel, self.scale, self.strategy, self.device = grid, kernel, float(scale), strategy, device
        kd = _lib.KernelDesc(kernel.kind, float(getattr(kernel, "alpha", 0.0)))
        out = C.c_void_p()
        _lib.check(L.cgef_plan_create(grid.handle(device), C.byref(kd), C.c_double(self.scale), strategy.code,
                                      ENGINES[engine], C.byref(out)))
        self.h = out
        info = _lib.PlanInfo()
        _lib.check(L.cgef_plan_query(self.h, C.byref(info)))
        self.info = info
        self.engine = ENGINE_NAMES[info.engine]

    def query(self):
        """The plan's current description (cgef_plan_query): refreshes `info`, whose contraction kernel
        choice is made by measurement on the plan's first ring call."""
        info = _lib.PlanInfo()
        _lib.check(_lib.load().cgef_plan_query(self.h, C.byref(info)))
        self.info = info
        return info

    def mode_cut(self):
        """Per row (Ny·Nz, level-major) the Fourier modes along x that FP32 calls of this ring plan keep
        (cgef_plan_mode_cut); N/2 + 1 on rows where the cut drops none."""
        import numpy as np
        n = int(np.prod(self.grid.field_shape[:-1]))
        out = np.empty(n, dtype=np.int32)
        _lib.check(_lib.load().cgef_plan_mode_cut(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return out

    @property
    def mean_taps(self):
        return self.info.mean_taps

    def __del__(self):
        try:
            _lib.load().cgef_plan_destroy(self.h)
        except Exception:
            pass


def _validate_scale(scale):
    # Filtering.jl:2917-2922
    s = float(scale)
    if not (s == s and abs(s) != float("inf") and s > 0):
        raise ArgumentError(f"filter scale must be finite and positive, got {scale}")


def _spectral_exact(grid: StructuredGrid) -> bool:
    """Filtering.jl:3000-3006: a periodic transform is the exact filter where every axis wraps and is
    uniformly spaced (the transform of this backend is 2-D)."""
    return (not grid.spherical and grid.ndim == 2 and all(grid.periodic) and all(grid.uniform))


def _resolve_method(grid: StructuredGrid, method):
    """Filtering.jl:3012-3019 (_resolve_method): AutoMethod becomes Spectral only where exact."""
    if method is None:
        return RealSpace()
    if isinstance(method, AutoMethod):
        return Spectral() if _spectral_exact(grid) else RealSpace()
    if not isinstance(method, (RealSpace, Spectral)):
        raise ArgumentError("method must be RealSpace(), Spectral() or AutoMethod()")
    return method


def plan_filter(grid: StructuredGrid, kernel: AbstractFilterKernel, scale: float, *, mask_strategy=ZeroFill(),
                backend=None, method=None, engine: str = "auto") -> FilterPlan:
    """Filtering.jl:2931-2954. `engine='auto'` takes the reference's default for the grid (the
    prefix-sum engine for a top-hat on a 2-D grid, the separable engine for a Gaussian on a uniform
    2-D Cartesian grid, the banded footprint otherwise — evaluated by the ring engine where x is
    periodic); 'banded' forces the disk/cap/ball direct sum, 'ring' its Fourier evaluation along x,
    'prefixsum' the per-row prefix sums, 'baseline' the plain FP64 kernel. `method=Spectral()` (or
    `engine='spectral'`, or AutoMethod() on a periodic uniform grid) filters by 2-D transform, as the
    reference's FFTW extension does (FFTWExt.jl:121-208)."""
    _validate_scale(scale)
    if not isinstance(mask_strategy, (ZeroFill, Deformable)):
        raise ArgumentError("mask_strategy must be ZeroFill() or Deformable()")
    m = _resolve_method(grid, method)
    if engine == "spectral":
        m = Spectral()
    if isinstance(m, Spectral):
        if engine not in ("auto", "spectral"):
            raise ArgumentError(f"engine = {engine!r} is a real-space engine; method = Spectral() filters by transform")
        if grid.spherical or grid.ndim != 2:
            raise ArgumentError("this backend's Spectral() filtering is the 2-D Cartesian transform; "
                                f"no transform is provided for a {grid.ndim}-D {'spherical' if grid.spherical else 'Cartesian'} grid")
        if not all(grid.periodic):
            # FFTWExt.jl:133-136
            raise ArgumentError("Spectral FFT filtering requires a doubly-periodic Cartesian grid; build it with "
                                "StructuredGrid(geom, x, y, mask=..., periodic=(True, True)).")
        engine = "spectral"
    return FilterPlan(grid, kernel, scale, mask_strategy, _device_of(backend), engine)


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return _lib.F32
    if t.dtype == torch.float64:
        return _lib.F64
    raise ArgumentError(f"fields must be float32 or float64, got {t.dtype}")


def _check_field(t: torch.Tensor, name: str, grid: StructuredGrid, device: int):
    if not isinstance(t, torch.Tensor):
        raise ArgumentError(f"{name} must be a torch tensor on the GPU")
    if t.device.type != "cuda" or t.device.index != device:
        raise ArgumentError(f"{name} must live on cuda:{device} (the plan's device), got {t.device}")
    if not t.is_contiguous():
        raise ArgumentError(f"{name} must be contiguous")
    gs = grid.field_shape
    if t.dim() < len(gs) or tuple(t.shape[t.dim() - len(gs):]) != gs:
        raise DimensionMismatch(f"{name}'s trailing axes {tuple(t.shape)} do not match grid shape {gs}")


def batch_dims(t: torch.Tensor, grid: StructuredGrid):
    return tuple(t.shape[: t.dim() - grid.ndim])


def _nslices(t: torch.Tensor, grid: StructuredGrid) -> int:
    n = 1
    for s in batch_dims(t, grid):
        n *= s
    return n


def _stream(device: int):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def filter_apply_(out: torch.Tensor, field: torch.Tensor, plan: FilterPlan) -> torch.Tensor:
    """Filtering.jl:3242-3256 / 3286-3305: filter `field` (any leading batch axes) into `out`."""
    g = plan.grid
    _check_field(field, "field", g, plan.device)
    _check_field(out, "out", g, plan.device)
    if out.shape != field.shape:
        raise DimensionMismatch(f"out has shape {tuple(out.shape)}, field has {tuple(field.shape)}")
    if out.dtype != field.dtype:
        raise ArgumentError("out and field must share a dtype")
    _lib.check(_lib.load().cgef_filter_apply(plan.h, C.c_void_p(out.data_ptr()), C.c_void_p(field.data_ptr()),
                                              1, _nslices(field, g), _dtype_code(field), _stream(plan.device)))
    return out


def filter_apply_batch_(outs, fields, plan: FilterPlan):
    """Filtering.jl:3360-3385: several fields sharing one plan. Fields stacked in one tensor share
    each neighbour walk; a list of tensors is filtered tensor by tensor."""
    if isinstance(fields, torch.Tensor):
        return filter_apply_(outs, fields, plan)
    if len(outs) != len(fields):
        raise DimensionMismatch("outs and fields must have the same length")
    for o, f in zip(outs, fields):
        filter_apply_(o, f, plan)
    return outs


def filter_field_(out, field, grid, kernel, scale, *, mask_strategy=ZeroFill(), filter_plan=None, backend=None,
                  method=None, engine: str = "auto"):
    """Filtering.jl:266-280 filter_field!."""
    plan = filter_plan if filter_plan is not None else plan_filter(
        grid, kernel, scale, mask_strategy=mask_strategy, backend=backend, method=method, engine=engine)
    return filter_apply_(out, field, plan)


def filter_slices_(outs, fields, plans, *, backend=None):
    """Filtering.jl:3433-3470 filter_slices!: apply plans[t] to fields[t] into outs[t] over
    independent slices, each with its own grid and plan. On one GPU the slices run back to back on
    the current stream, longest first (slice_costs)."""
    if not (len(outs) == len(fields) == len(plans)):
        raise DimensionMismatch(f"filter_slices! got {len(outs)} outputs, {len(fields)} fields and {len(plans)} plans")
    for t in sorted(range(len(plans)), key=lambda q: -slice_costs([plans[q]])[0]):
        filter_apply_(outs[t], fields[t], plans[t])
    return outs


def slice_costs(plans):
    """Filtering.jl:3486-3491: per-slice work proxy, the number of target points each plan writes."""
    return [int(p.grid.npoints) for p in plans]

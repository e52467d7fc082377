"""Derivatives (Derivatives.jl): ddx!/ddy!/ddz! — derivative w.r.t. distance along one axis, on the
GPU. The stencil is the 3-node one of FlowGeometries' derivative!(order=1, nodes=3, masked=0,
ReduceInRun) (Derivatives.jl:26-31): centred inside a run of active cells, shifted inward at a domain
or mask edge, two nodes for a run of two, 0 for an isolated or masked cell; periodic axes wrap."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .filtering import _check_field, _dtype_code, _nslices, _stream
from .grids import StructuredGrid


class StencilPlan:
    """Derivatives.jl:74-122. The device grid already holds the axis tables, so the plan only names
    the grid; it exists so callers written against the reference keep their signature."""

    def __init__(self, grid: StructuredGrid):
        self.grid = grid


def _dd(out: torch.Tensor, f: torch.Tensor, grid: StructuredGrid, axis: int) -> torch.Tensor:
    dev = f.device.index if f.is_cuda else 0
    _check_field(f, "f", grid, dev)
    _check_field(out, "out", grid, dev)
    if out.shape != f.shape or out.dtype != f.dtype:
        raise _lib.DimensionMismatch("out must match f in shape and dtype")
    _lib.check(_lib.load().cgef_derivative(grid.handle(dev), axis, C.c_void_p(out.data_ptr()),
                                            C.c_void_p(f.data_ptr()), _nslices(f, grid), _dtype_code(f), _stream(dev)))
    return out


def ddx_(out, f, grid, plan=None):
    return _dd(out, f, grid, 0)


def ddy_(out, f, grid, plan=None):
    return _dd(out, f, grid, 1)


def ddz_(out, f, grid, plan=None):
    if grid.ndim != 3:
        raise _lib.ArgumentError("ddz! needs a 3-D grid")
    return _dd(out, f, grid, 2)

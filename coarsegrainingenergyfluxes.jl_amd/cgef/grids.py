"""Grids: the host description of a StructuredGrid (FlowGeometries.Grids, un-vendored in the
reference) and its device-resident twin inside libcgef_hip.so.

Array convention: a field on an (Nx, Ny[, Nz]) grid is a C-contiguous tensor of shape
(Ny, Nx) / (Nz, Ny, Nx), optionally with leading batch axes — byte-for-byte the reference's
column-major (Nx, Ny[, Nz], batch…) layout with i fastest.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError


@dataclass(frozen=True)
class CartesianGeometry:
    pass


@dataclass(frozen=True)
class SphericalGeometry:
    R: float = 6.371e6


class Range:
    """A uniform axis start, start+step, …, start+(n−1)·step (the reference's AbstractRange)."""

    def __init__(self, start: float, step: float, length: int):
        self.start, self.step, self.length = float(start), float(step), int(length)
        self.values = self.start + self.step * np.arange(self.length, dtype=np.float64)

    def __len__(self):
        return self.length


def _axis(a):
    if isinstance(a, Range):
        return a.values, True, a.step * len(a)
    v = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if v.ndim != 1 or len(v) < 1:
        raise ArgumentError("an axis must be a non-empty 1-D coordinate vector")
    d = np.diff(v)
    # an evenly spaced vector (to rounding) is a Range; the step used is x[1] − x[0]
    uniform = bool(len(d) == 0 or np.all(np.abs(d - d[0]) <= 1e-9 * abs(d[0])))
    period = (v[1] - v[0]) * len(v) if len(v) > 1 else 0.0
    return v, uniform, period


class StructuredGrid:
    """StructuredGrid(geometry, x, y[, z]; mask=None, periodic=None, period=None).

    Spherical axes are longitude and latitude in radians. `periodic` defaults to a full-circle
    longitude on the sphere and to non-periodic Cartesian axes; `period` defaults to n·step."""

    def __init__(self, geometry, *axes, mask=None, periodic=None, period=None):
        if len(axes) not in (2, 3):
            raise ArgumentError("a StructuredGrid has 2 or 3 axes")
        self.geometry = geometry
        self.spherical = isinstance(geometry, SphericalGeometry)
        if self.spherical and len(axes) != 2:
            raise ArgumentError("spherical grids are 2-D (lon, lat) here")
        parsed = [_axis(a) for a in axes]
        self.axes = [p[0] for p in parsed]
        self.uniform = [p[1] for p in parsed]
        self.ndim = len(axes)
        self.shape = tuple(len(a) for a in self.axes)        # reference order (Nx, Ny[, Nz])
        self.field_shape = tuple(reversed(self.shape))       # tensor order (…, Ny, Nx)
        if self.spherical:
            span = len(self.axes[0]) * (self.axes[0][1] - self.axes[0][0]) if len(self.axes[0]) > 1 else 0.0
            dflt_per = [bool(abs(span - 2 * np.pi) < 1e-9 * 2 * np.pi), False]
            dflt_period = [2 * np.pi, 0.0]
        else:
            dflt_per = [False] * self.ndim
            dflt_period = [p[2] for p in parsed]
        self.periodic = [bool(x) for x in (periodic if periodic is not None else dflt_per)]
        self.period = [float(x) for x in (period if period is not None else dflt_period)]
        if len(self.periodic) != self.ndim or len(self.period) != self.ndim:
            raise ArgumentError("periodic/period need one entry per axis")
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
            if m.shape != self.field_shape:
                raise _lib.DimensionMismatch(f"mask has shape {m.shape}, grid expects {self.field_shape}")
            self.mask = m
        else:
            self.mask = None
        self.R = float(geometry.R) if self.spherical else 1.0
        self._handles = {}

    @property
    def npoints(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    def isactive_mask(self):
        return np.ones(self.field_shape, bool) if self.mask is None else self.mask.astype(bool)

    def handle(self, device: int):
        """The device-resident grid on `device` (created once)."""
        h = self._handles.get(device)
        if h is not None:
            return h
        L = _lib.load()
        d = _lib.GridDesc()
        d.ndim = self.ndim
        d.geometry = _lib.SPHERICAL if self.spherical else _lib.CARTESIAN
        for a in range(3):
            if a < self.ndim:
                d.n[a] = len(self.axes[a])
                d.coord[a] = self.axes[a].ctypes.data_as(C.POINTER(C.c_double))
                d.uniform[a] = int(self.uniform[a])
                d.periodic[a] = int(self.periodic[a])
                d.period[a] = self.period[a]
            else:
                d.n[a] = 1
        d.radius = self.R
        if self.mask is not None:
            d.mask = self.mask.ctypes.data_as(C.POINTER(C.c_uint8))
        out = C.c_void_p()
        _lib.check(L.cgef_grid_create(C.byref(d), int(device), C.byref(out)))
        self._handles[device] = out
        return out

    def __del__(self):
        try:
            L = _lib.load()
            for h in self._handles.values():
                L.cgef_grid_destroy(h)
        except Exception:
            pass

"""ctypes front end of the FP64 CPU oracle (``liboracle.so``, built from ``cgef_oracle.c``).

TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline, never the product. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this module.

Arrays are numpy, laid out as the reference lays them out (column-major ``(Nx, Ny[, Nz])`` with
``i`` fastest), i.e. C-order numpy arrays of shape ``(Nz, Ny, Nx)`` / ``(Ny, Nx)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TOPHAT, GAUSSIAN = 0, 1
ZEROFILL, DEFORMABLE = 0, 1
ENGINE_BANDED, ENGINE_SEPARABLE, ENGINE_ND3, ENGINE_PREFIX, ENGINE_SPECTRAL = 0, 1, 2, 3, 4


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _Grid(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("spherical", C.c_int32),
        ("n", C.c_int64 * 3),
        ("x", C.POINTER(C.c_double) * 3),
        ("uniform", C.c_int32 * 3),
        ("periodic", C.c_int32 * 3),
        ("period", C.c_double * 3),
        ("R", C.c_double),
        ("mask", C.POINTER(C.c_uint8)),
    ]


class _Spec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("alpha", C.c_double),
        ("ell", C.c_double),
        ("strategy", C.c_int32),
        ("engine", C.c_int32),
        ("nthreads", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        gp = C.POINTER(_Grid)
        L.og_set_mutation.argtypes = [C.c_int]
        L.og_cell_width.restype = C.c_double
        L.og_cell_width.argtypes = [gp, C.c_int, C.c_int64]
        L.og_kernel_weight.restype = C.c_double
        L.og_kernel_weight.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.og_kernel_radius.restype = C.c_double
        L.og_kernel_radius.argtypes = [C.c_int, C.c_double, C.c_double]
        L.og_haversine.restype = C.c_double
        L.og_haversine.argtypes = [C.c_double] * 5
        L.og_area.restype = C.c_double
        L.og_area.argtypes = [gp, C.c_int64, C.c_int64, C.c_int64]
        L.og_footprint_build.restype = C.c_void_p
        L.og_footprint_build.argtypes = [gp, C.c_int, C.c_double, C.c_double]
        L.og_footprint_build_bands.restype = C.c_void_p
        L.og_footprint_build_bands.argtypes = [gp, C.c_int, C.c_double, C.c_double, C.c_void_p]
        L.og_footprint_free.argtypes = [C.c_void_p]
        L.og_footprint_nentries.restype = C.c_int64
        L.og_footprint_nentries.argtypes = [C.c_void_p]
        L.og_footprint_nbands.restype = C.c_int64
        L.og_footprint_nbands.argtypes = [C.c_void_p]
        L.og_footprint_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.og_apply_banded.argtypes = [gp, C.c_void_p, C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp), C.c_int]
        L.og_apply_separable.argtypes = [gp, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp)]
        L.og_apply_nd3.argtypes = [gp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(dp), C.POINTER(dp), C.c_int]
        L.og_apply_prefix.argtypes = [gp, C.c_double, C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp), C.c_int]
        L.og_apply_spectral.argtypes = [gp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(dp),
                                        C.POINTER(dp)]
        L.og_spectral_transfer.restype = C.c_double
        L.og_spectral_transfer.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.og_metric_band.restype = C.c_double
        L.og_metric_band.argtypes = [gp, C.c_double, C.c_double, C.c_double]
        L.og_deriv.argtypes = [gp, C.c_int, dp, dp]
        L.og_vector_to_cartesian.argtypes = [C.c_double] * 5 + [dp]
        L.og_vector_from_cartesian.argtypes = [C.c_double] * 5 + [dp]
        L.og_tensor_to_local.argtypes = [dp, C.c_double, C.c_double, dp]
        L.og_compute_pi.argtypes = [gp, C.POINTER(_Spec), C.c_void_p, dp, dp, dp, dp, dp, dp, dp]
        L.og_compute_pi_rows.argtypes = [gp, C.POINTER(_Spec), C.c_void_p, dp, dp, C.c_void_p, dp, dp, dp]
        L.og_active_area.restype = C.c_double
        L.og_active_area.argtypes = [gp]
        L.og_energy.restype = C.c_double
        L.og_energy.argtypes = [gp, dp, dp, dp]
        L.og_spectral_density.argtypes = [C.c_int64, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Grid:
    """Oracle-side grid: axes (numpy float64), periodic flags, optional mask (shape like fields)."""

    def __init__(self, axes, *, spherical=False, R=6.371e6, periodic=None, period=None,
                 uniform=None, mask=None):
        self.axes = [np.ascontiguousarray(a, dtype=np.float64) for a in axes]
        self.ndim = len(self.axes)
        self.spherical = bool(spherical)
        self.R = float(R)
        self.shape = tuple(len(a) for a in self.axes)  # (Nx, Ny[, Nz])
        self.periodic = list(periodic) if periodic is not None else [False] * self.ndim
        if period is None:
            period = []
            for a in self.axes:
                period.append(len(a) * (a[1] - a[0]) if len(a) > 1 else 0.0)
        self.period = list(period)
        if uniform is None:
            uniform = []
            for a in self.axes:
                d = np.diff(a)
                uniform.append(bool(len(d) == 0 or np.all(d == d[0])))
        self.uniform = list(uniform)
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        self._keep = []
        s = _Grid()
        s.ndim = self.ndim
        s.spherical = int(self.spherical)
        for d in range(3):
            if d < self.ndim:
                s.n[d] = len(self.axes[d])
                s.x[d] = _dp(self.axes[d])
                s.uniform[d] = int(self.uniform[d])
                s.periodic[d] = int(self.periodic[d])
                s.period[d] = float(self.period[d])
            else:
                s.n[d] = 1
        s.R = self.R
        if self.mask is not None:
            s.mask = self.mask.ctypes.data_as(C.POINTER(C.c_uint8))
        self._s = s

    @property
    def field_shape(self):
        return tuple(reversed(self.shape))

    def ref(self):
        return C.byref(self._s)


def kernel_weight(kind, alpha, d, ell):
    return lib().og_kernel_weight(kind, alpha, d, ell)


def kernel_radius(kind, alpha, ell):
    return lib().og_kernel_radius(kind, alpha, ell)


def haversine(R, l1, p1, l2, p2):
    return lib().og_haversine(R, l1, p1, l2, p2)


def area(grid, i, j, k=0):
    return lib().og_area(grid.ref(), i, j, k)


def cell_width(grid, axis, i):
    return lib().og_cell_width(grid.ref(), axis, i)


MUTATE_FORWARD_WIDTH, MUTATE_RUNEND_2NODE = 1, 2


class mutation:
    """Test-only context: the oracle with a planted mistake (forward-difference widths, or two-node
    stencils at run ends), to show that a pinning test fails on it."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().og_set_mutation(self.m)

    def __exit__(self, *exc):
        lib().og_set_mutation(0)


class Footprint:
    """The banded footprint (Filtering.jl:636-732). bands: on a spherical grid, build only these
    latitude bands (the others empty); copy=False leaves the entries in the C table only."""

    def __init__(self, grid, kind, alpha, ell, bands=None, copy=True):
        L = lib()
        if bands is None:
            self.h = L.og_footprint_build(grid.ref(), kind, alpha, ell)
        else:
            sel = np.zeros(grid.shape[1], np.uint8)
            sel[np.asarray(list(bands), dtype=np.int64)] = 1
            self.h = L.og_footprint_build_bands(grid.ref(), kind, alpha, ell, sel.ctypes.data)
        n = L.og_footprint_nentries(self.h)
        nb = L.og_footprint_nbands(self.h)
        self.nbands, self.nentries = nb, n
        if not copy:
            return
        self.di = np.zeros(n, np.int32)
        self.dj = np.zeros(n, np.int32)
        self.w = np.zeros(n, np.float64)
        self.ptr = np.zeros(nb + 1, np.int64)
        L.og_footprint_copy(self.h, self.di.ctypes.data, self.dj.ctypes.data, self.w.ctypes.data,
                            self.ptr.ctypes.data)
        self.nbands = nb

    def __del__(self):
        try:
            lib().og_footprint_free(self.h)
        except Exception:
            pass


def _arr(fields):
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    return fs, (C.POINTER(C.c_double) * len(fs))(*[_dp(f) for f in fs])


def _planar_only(grid, what):
    """The banded, separable, prefix-sum and transform restatements walk one (x, y) plane; a 3-D grid
    goes through apply_nd3 (Diagnostics.jl:792-857's 3-D path)."""
    if grid.ndim != 2:
        raise ValueError(f"the oracle's {what} engine is 2-D only; use ENGINE_ND3 on a 3-D grid")


def apply_banded(grid, fp, strategy, fields, nthreads=0):
    _planar_only(grid, "banded")
    ins, pin = _arr(fields)
    outs = [np.empty_like(f) for f in ins]
    pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(o) for o in outs])
    lib().og_apply_banded(grid.ref(), fp.h, strategy, len(ins), pin, pout, nthreads)
    return outs


def apply_separable(grid, alpha, ell, strategy, fields):
    _planar_only(grid, "separable")
    ins, pin = _arr(fields)
    outs = [np.empty_like(f) for f in ins]
    pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(o) for o in outs])
    lib().og_apply_separable(grid.ref(), alpha, ell, strategy, len(ins), pin, pout)
    return outs


def apply_nd3(grid, kind, alpha, ell, strategy, fields, nthreads=0):
    ins, pin = _arr(fields)
    outs = [np.empty_like(f) for f in ins]
    pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(o) for o in outs])
    lib().og_apply_nd3(grid.ref(), kind, alpha, ell, strategy, len(ins), pin, pout, nthreads)
    return outs


def apply_prefix(grid, ell, strategy, fields, nthreads=0):
    """Filtering.jl:751-1097: the reference's top-hat engine on 2-D rectilinear grids."""
    _planar_only(grid, "prefix-sum")
    ins, pin = _arr(fields)
    outs = [np.empty_like(f) for f in ins]
    pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(o) for o in outs])
    lib().og_apply_prefix(grid.ref(), ell, strategy, len(ins), pin, pout, nthreads)
    return outs


def apply_spectral(grid, kind, alpha, ell, strategy, fields):
    """FFTWExt.jl:121-208: transform filter on a periodic uniform 2-D Cartesian grid."""
    _planar_only(grid, "transform")
    ins, pin = _arr(fields)
    outs = [np.empty_like(f) for f in ins]
    pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(o) for o in outs])
    lib().og_apply_spectral(grid.ref(), kind, alpha, ell, strategy, len(ins), pin, pout)
    return outs


def spectral_transfer(kind, alpha, k, ell):
    return lib().og_spectral_transfer(kind, alpha, k, ell)


def metric_band(grid, y1, y2, rad):
    return lib().og_metric_band(grid.ref(), y1, y2, rad)


def deriv(grid, axis, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    lib().og_deriv(grid.ref(), axis, _dp(f), _dp(out))
    return out


def vector_to_cartesian(u, v, w, lam, phi):
    out = np.zeros(3)
    lib().og_vector_to_cartesian(u, v, w, lam, phi, _dp(out))
    return out


def vector_from_cartesian(X, Y, Z, lam, phi):
    out = np.zeros(3)
    lib().og_vector_from_cartesian(X, Y, Z, lam, phi, _dp(out))
    return out


def tensor_to_local(t6, lam, phi):
    t = np.ascontiguousarray(t6, dtype=np.float64)
    out = np.zeros(6)
    lib().og_tensor_to_local(_dp(t), lam, phi, _dp(out))
    return out


def compute_pi(grid, u, v, w=None, *, kind=GAUSSIAN, alpha=6.0, ell, strategy=ZEROFILL,
               engine=ENGINE_BANDED, fp=None, nthreads=0):
    """Π = -S̄:τ (no ρ₀) plus the filtered velocity the reference leaves in its workspace."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    if engine != ENGINE_ND3:
        _planar_only(grid, {ENGINE_BANDED: "banded", ENGINE_SEPARABLE: "separable", ENGINE_PREFIX: "prefix-sum",
                            ENGINE_SPECTRAL: "transform"}.get(engine, str(engine)))
    if engine == ENGINE_BANDED and fp is None:
        fp = Footprint(grid, kind, alpha, ell)
    spec = _Spec(kind, alpha, ell, strategy, engine, nthreads)
    Pi = np.empty_like(u)
    ub = np.empty_like(u)
    vb = np.empty_like(u)
    wb = np.empty_like(u)
    lib().og_compute_pi(grid.ref(), C.byref(spec), fp.h if fp is not None else None,
                        _dp(u), _dp(v), _dp(w), _dp(Pi), _dp(ub), _dp(vb), _dp(wb))
    return Pi, ub, vb, wb


def compute_pi_rows(grid, u, v, rows, *, kind=GAUSSIAN, alpha=6.0, ell, strategy=ZEROFILL, fp=None, nthreads=0):
    """Π (and ū, v̄) of the 2-D banded footprint at the listed rows j only (NaN elsewhere); on those rows
    identical to compute_pi. Costs the filter on the rows and their two neighbours either side."""
    _planar_only(grid, "banded")
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    sel = np.zeros(grid.shape[1], np.uint8)
    sel[np.asarray(list(rows), dtype=np.int64)] = 1
    if fp is None:
        need = sorted({min(max(j + d, 0), grid.shape[1] - 1) for j in rows for d in range(-2, 3)})
        fp = Footprint(grid, kind, alpha, ell, bands=need if grid.spherical else None, copy=False)
    spec = _Spec(kind, alpha, ell, strategy, ENGINE_BANDED, nthreads)
    Pi, ub, vb = np.empty_like(u), np.empty_like(u), np.empty_like(u)
    lib().og_compute_pi_rows(grid.ref(), C.byref(spec), fp.h, _dp(u), _dp(v), sel.ctypes.data, _dp(Pi), _dp(ub),
                             _dp(vb))
    return Pi, ub, vb


def active_area(grid):
    return lib().og_active_area(grid.ref())


def energy(grid, ub, vb, wb=None):
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    vb = np.ascontiguousarray(vb, dtype=np.float64)
    wb = None if wb is None else np.ascontiguousarray(wb, dtype=np.float64)
    return lib().og_energy(grid.ref(), _dp(ub), _dp(vb), _dp(wb))


def spectral_density(Cv, k):
    Cv = np.ascontiguousarray(Cv, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    g = np.zeros_like(Cv)
    lib().og_spectral_density(len(Cv), _dp(Cv), _dp(k), _dp(g))
    return g


def filter_fields(grid, fields, *, kind, alpha, ell, strategy=ZEROFILL, engine=ENGINE_BANDED, fp=None,
                  nthreads=0):
    if engine == ENGINE_SEPARABLE:
        return apply_separable(grid, alpha, ell, strategy, fields)
    if engine == ENGINE_ND3:
        return apply_nd3(grid, kind, alpha, ell, strategy, fields, nthreads)
    if engine == ENGINE_PREFIX:
        return apply_prefix(grid, ell, strategy, fields, nthreads)
    if engine == ENGINE_SPECTRAL:
        return apply_spectral(grid, kind, alpha, ell, strategy, fields)
    if fp is None:
        fp = Footprint(grid, kind, alpha, ell)
    return apply_banded(grid, fp, strategy, fields, nthreads)


def cumulative_energy(grid, u, v, scales, *, kind=GAUSSIAN, alpha=6.0, strategy=ZEROFILL,
                      engine=ENGINE_BANDED, nthreads=0):
    """Diagnostics.jl:1144-1217: filter the raw (u, v) per scale and take the area-weighted mean."""
    out = np.zeros(len(scales))
    for s, ell in enumerate(scales):
        ub, vb = filter_fields(grid, [u, v], kind=kind, alpha=alpha, ell=float(ell), strategy=strategy,
                               engine=engine, nthreads=nthreads)
        out[s] = energy(grid, ub, vb)
    return out

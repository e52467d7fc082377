"""cgef_energy_noalloc, the allocation-free variant of cgef_energy, through the C ABI on the GPU: two
calls in a row (the second with a smaller input) each bit-identical to cgef_energy and within 1e-12
of the oracle, sizes beyond its fixed scratch refused before any launch, and no device memory taken
by a call where cgef_energy takes its workspace."""
import ctypes as C

import numpy as np
import pytest
import torch

import cgef
from cgef import _lib
from helpers import oracle_grid_sphere, smooth_random_field, sphere_axes
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _energy(name, grid, ub, vb, nb, dtype):
    E = np.zeros(nb)
    st = getattr(_lib.load(), name)(grid.handle(0), C.c_void_p(ub.data_ptr()), C.c_void_p(vb.data_ptr()), None, nb,
                                    dtype, E.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.c_void_p(torch.cuda.current_stream(0).cuda_stream))
    _lib.check(st)
    return E


def _case(nlon, nlat, nb, seed, torch_dtype):
    lam, phi = sphere_axes(nlon, nlat)
    f = smooth_random_field((nlat, nlon), seed, 1, smooth=6)[0]
    mask = f > np.quantile(f, 0.3)
    g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
    fields = smooth_random_field((nb, nlat, nlon), seed + 1, 2)
    ub = torch.tensor(fields[0], dtype=torch_dtype, device=DEV)
    vb = torch.tensor(fields[1], dtype=torch_dtype, device=DEV)
    og = oracle_grid_sphere(lam, phi, mask=mask)
    ref = np.array([O.energy(og, ub[b].cpu().double().numpy(), vb[b].cpu().double().numpy()) for b in range(nb)])
    return g, ub, vb, ref


def test_energy_noalloc_matches_energy_on_two_calls():
    print("noalloc energy: two calls, the second smaller", flush=True)
    for nlon, nlat, nb, seed, td, dt in ((1440, 720, 6, 3, torch.float32, _lib.F32),
                                         (144, 72, 2, 9, torch.float64, _lib.F64)):
        g, ub, vb, ref = _case(nlon, nlat, nb, seed, td)
        got = _energy("cgef_energy_noalloc", g, ub, vb, nb, dt)
        want = _energy("cgef_energy", g, ub, vb, nb, dt)
        assert np.array_equal(got, want)
        assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref))


def test_energy_noalloc_refuses_sizes_beyond_its_scratch():
    print("noalloc energy: limits", flush=True)
    g, ub, vb, _ = _case(1440, 720, 1, 5, torch.float32)
    # 1440 x 720: ceil(1440/256)·720 = 4320 partial sums per slice, 242 slices fit in 1,048,576
    for nb in (243, 5000):
        with pytest.raises(cgef.DimensionMismatch, match="at most"):
            _energy("cgef_energy_noalloc", g, ub, vb, nb, _lib.F32)


def test_energy_noalloc_takes_no_device_memory():
    print("noalloc energy: device memory", flush=True)
    ga, ua, va, _ = _case(96, 48, 1, 11, torch.float32)
    _energy("cgef_energy_noalloc", ga, ua, va, 1, _lib.F32)       # the library's code is loaded
    g1, u1, v1, _ = _case(1440, 720, 2, 12, torch.float32)
    g2, u2, v2, _ = _case(1440, 720, 2, 13, torch.float32)
    g1.handle(0), g2.handle(0)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    _energy("cgef_energy_noalloc", g1, u1, v1, 2, _lib.F32)
    free1 = torch.cuda.mem_get_info(0)[0]
    _energy("cgef_energy", g2, u2, v2, 2, _lib.F32)               # allocates the grid's workspace
    free2 = torch.cuda.mem_get_info(0)[0]
    assert free0 - free1 < (1 << 20)
    assert free1 - free2 > (16 << 20)

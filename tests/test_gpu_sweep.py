"""The multi-scale sweep's launch machinery on the GPU: the HIP-graph replay of cgef_coarse_grain (opt-in,
CGEF_GRAPH=1) and the measured choice of the contraction kernel variant. Every replay is compared bit for bit with the
eager sweep (same kernels, same order), with the program's own eager work and synchronizations
between replays, and after the inputs change in place; every forced contraction variant gives the
same bits."""
import os

import numpy as np
import pytest
import torch

import cgef
from helpers import smooth_random_field, sphere_axes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(spectral=False):
    if spectral:
        n = 96
        xs = np.arange(n) * 1.0e4
        g = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
        scales, ker, method = [4e4, 7e4, 1.2e5], cgef.GaussianKernel(6.0), cgef.Spectral()
    else:
        lam, phi = sphere_axes(96, 48)
        f = smooth_random_field((48, 96), 31, 1, smooth=6)[0]
        g = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=f > np.quantile(f, 0.25))
        scales, ker, method = [6e5, 9e5, 1.5e6], cgef.GaussianKernel(4.0), None
    fs = smooth_random_field((2,) + g.field_shape, 32, 2)
    U = torch.tensor(fs[0], dtype=torch.float32, device=DEV)
    V = torch.tensor(fs[1], dtype=torch.float32, device=DEV)
    plans = [cgef.plan_filter(g, ker, s, method=method) for s in scales]
    return g, U, V, scales, ker, plans


def _sweep(batch, U, V, g, scales, ker, plans, ws):
    cgef.coarse_grain_batch_(batch, U, V, None, g, scales=scales, kernel=ker, filter_plans=plans, workspaces=[ws])
    torch.cuda.synchronize()
    return batch.Pi.clone(), batch.cumulative_energy.copy()


@pytest.mark.parametrize("spectral", [False, True])
def test_graph_replays_equal_the_eager_sweep(spectral, monkeypatch):
    print(f"sweep graph spectral={spectral}", flush=True)
    g, U, V, scales, ker, plans = _case(spectral)
    assert plans[0].engine == ("spectral" if spectral else "ring")
    batch = cgef.CoarseGrainBatchResult.allocate(g, len(scales), (2,), torch.float32, DEV)
    ws = cgef.PiWorkspace(g, (2,), torch.float32)
    monkeypatch.setenv("CGEF_GRAPH", "0")
    ref_pi, ref_e = _sweep(batch, U, V, g, scales, ker, plans, ws)
    monkeypatch.setenv("CGEF_GRAPH", "1")
    other = torch.empty((2,) + g.field_shape, dtype=torch.float32, device=DEV)
    for call in range(5):   # eager (arguments seen), captured + launched, then replays
        batch.Pi.zero_()
        pi, e = _sweep(batch, U, V, g, scales, ker, plans, ws)
        assert torch.equal(pi, ref_pi), call
        assert np.array_equal(e, ref_e), call
        # the program's own eager work between replays
        cgef.compute_pi_(other, V, U, None, g, ker, scales[1], filter_plan=plans[1])
        torch.cuda.synchronize()
    # new values at the same addresses: the replay reads them
    U.mul_(-0.5)
    V.add_(0.25)
    pi, e = _sweep(batch, U, V, g, scales, ker, plans, ws)
    monkeypatch.setenv("CGEF_GRAPH", "0")
    ref_pi2, ref_e2 = _sweep(batch, U, V, g, scales, ker, plans, ws)
    assert not torch.equal(ref_pi2, ref_pi)
    assert torch.equal(pi, ref_pi2)
    assert np.array_equal(e, ref_e2)


def test_contraction_variants_give_the_same_bits(monkeypatch):
    print("sweep contraction variants", flush=True)
    g, U, V, scales, ker, plans = _case(False)
    monkeypatch.setenv("CGEF_GRAPH", "0")
    batch = cgef.CoarseGrainBatchResult.allocate(g, len(scales), (2,), torch.float32, DEV)
    ws = cgef.PiWorkspace(g, (2,), torch.float32)
    outs = []
    for variant in ("0", "1", "2", None):   # None: the measured choice
        if variant is None:
            monkeypatch.delenv("CGEF_CONTRACT_VARIANT", raising=False)
        else:
            monkeypatch.setenv("CGEF_CONTRACT_VARIANT", variant)
        batch.Pi.zero_()
        outs.append(_sweep(batch, U, V, g, scales, ker, plans, ws))
    for pi, e in outs[1:]:
        assert torch.equal(pi, outs[0][0])
        assert np.array_equal(e, outs[0][1])

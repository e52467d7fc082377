"""Shared test builders: the same grid for the oracle (FP64 CPU checker) and for cgef (GPU)."""
import numpy as np

from oracle import oracle as O


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / (den if den > 0 else 1.0))


def cart_axes(N, dx, x0=0.0):
    return x0 + dx * np.arange(N, dtype=np.float64)


def oracle_grid_cart(xs, ys, zs=None, periodic=(True, True, False), mask=None, period=None):
    axes = [xs, ys] + ([zs] if zs is not None else [])
    nd = len(axes)
    per = list(periodic[:nd])
    if period is None:
        period = [len(a) * (a[1] - a[0]) if len(a) > 1 else 0.0 for a in axes]
    return O.Grid(axes, spherical=False, periodic=per, period=period, mask=mask)


def oracle_grid_sphere(lam, phi, mask=None, R=6.371e6):
    return O.Grid([lam, phi], spherical=True, R=R, periodic=[True, False], period=[2 * np.pi, 0.0], mask=mask,
                  uniform=[True, True])


def sphere_axes(nlon, nlat):
    dl = 2 * np.pi / nlon
    dp = np.pi / nlat
    return (np.arange(nlon) + 0.5) * dl, -np.pi / 2 + (np.arange(nlat) + 0.5) * dp


def smooth_random_field(shape, seed, ncomp=1, smooth=3):
    """Seeded smooth-ish random fields (tensor order) for parity tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(ncomp):
        f = rng.standard_normal(shape)
        for ax in range(len(shape)):
            for _ in range(smooth):
                f = 0.25 * np.roll(f, 1, axis=ax) + 0.5 * f + 0.25 * np.roll(f, -1, axis=ax)
        out.append(f)
    return out

"""Pins the FP64 CPU oracle (the checker) to the reference: its own known-answer numbers, closed forms,
invariants, identities its test suite asserts, and brute force on tiny inputs. CPU only. Where a pin
guards a choice the reference leaves to an un-vendored dependency (cell widths on non-uniform axes,
the derivative stencil at run ends), the test also plants the plausible wrong choice in the oracle
(oracle.mutation) and shows that the pin fails on it."""
import math

import numpy as np
import pytest

from helpers import cart_axes, oracle_grid_cart, oracle_grid_sphere, smooth_random_field, sphere_axes
from oracle import oracle as O

R_EARTH = 6.371e6


# ------------------------------------------------------------------ kernels (test_kernels.jl) --

def test_kernel_values_and_radius():
    # test_kernels.jl:9-30: G(ℓ) = e^{-α}, G(rad) < 1e-9, 1.5ℓ < rad < 2.5ℓ
    for a in (6.0, 4.0):
        assert O.kernel_weight(O.GAUSSIAN, a, 10.0, 10.0) == pytest.approx(math.exp(-a), rel=1e-15)
        rad = O.kernel_radius(O.GAUSSIAN, a, 10.0)
        assert O.kernel_weight(O.GAUSSIAN, a, rad, 10.0) < 1e-9
        assert 15.0 < rad < 25.0
    assert O.kernel_radius(O.GAUSSIAN, 6.0, 1.0) == pytest.approx(1.9590, abs=1e-4)  # "≈ 1.96 ℓ at α = 6"
    assert O.kernel_radius(O.GAUSSIAN, 4.0, 1.0) == pytest.approx(2.3993, abs=1e-4)
    assert O.kernel_radius(O.TOPHAT, 0.0, 8.0) == 4.0
    assert O.kernel_weight(O.TOPHAT, 0.0, 4.0, 8.0) == 1.0 and O.kernel_weight(O.TOPHAT, 0.0, 4.0001, 8.0) == 0.0


def test_haversine_known_answers():
    # test_geometry.jl:14-55
    d = O.haversine(R_EARTH, math.radians(-0.1278), math.radians(51.5074), math.radians(2.3522), math.radians(48.8566))
    assert 340e3 < d < 350e3
    assert O.haversine(R_EARTH, 0, 0, 0, math.pi / 2) == pytest.approx(math.pi * R_EARTH / 2, rel=1e-6)
    assert O.haversine(R_EARTH, 0, 0, math.radians(1.0), 0) == pytest.approx(math.pi * R_EARTH / 180, rel=1e-6)


def test_rotation_round_trip():
    lam, phi = math.radians(-122.0), math.radians(38.0)
    p = O.vector_to_cartesian(10.0, -5.0, 0.0, lam, phi)
    back = O.vector_from_cartesian(*p, lam, phi)
    assert back[0] == pytest.approx(10.0) and back[1] == pytest.approx(-5.0) and abs(back[2]) < 1e-12
    # e_east, e_north, e_r orthonormal: |p| preserved, radial component of a tangent vector 0
    assert np.linalg.norm(p) == pytest.approx(math.hypot(10.0, -5.0))
    # tensor_to_local of the identity is the identity; of r̂r̂ᵀ is e_r e_rᵀ
    loc = O.tensor_to_local([1, 0, 0, 1, 0, 1], lam, phi)
    assert np.allclose(loc, [1, 0, 0, 1, 0, 1], atol=1e-15)
    er = np.array([math.cos(phi) * math.cos(lam), math.cos(phi) * math.sin(lam), math.sin(phi)])
    T = np.outer(er, er)
    loc = O.tensor_to_local([T[0, 0], T[0, 1], T[0, 2], T[1, 1], T[1, 2], T[2, 2]], lam, phi)
    assert np.allclose(loc, [0, 0, 0, 0, 0, 1], atol=1e-15)


# -------------------------------------------------------------- footprints (Filtering.jl) --

def test_footprint_tap_counts():
    # tap counts of the reference's footprint definition, computed independently in the survey
    # from a closed-form per-row run length
    N, dx = 256, 1e6 / 256
    x = cart_axes(N, dx)
    fp = O.Footprint(oracle_grid_cart(x, x), O.GAUSSIAN, 6.0, 8 * dx)
    assert len(fp.w) == 777
    lam, phi = sphere_axes(1440, 720)
    fp = O.Footprint(oracle_grid_sphere(lam, phi), O.GAUSSIAN, 4.0, 200e3)
    cnt = np.diff(fp.ptr)
    assert len(fp.w) == 2221596 and cnt[360] == 941 and cnt.max() == 24967
    assert cnt.mean() == pytest.approx(3085.55, abs=0.01)


def test_constant_preserved():
    # test_filtering.jl:697-727: interior points of a constant field stay the constant
    x = cart_axes(51, 1000.0)
    g = oracle_grid_cart(x, x, periodic=(False, False))
    f = np.full(g.field_shape, 42.0)
    for kind, alpha in ((O.TOPHAT, 0.0), (O.GAUSSIAN, 6.0)):
        for ell in (5000.0, 10000.0, 20000.0):
            for strat in (O.ZEROFILL, O.DEFORMABLE):
                out = O.filter_fields(g, [f], kind=kind, alpha=alpha, ell=ell, strategy=strat)[0]
                assert np.allclose(out[19:32, 19:32], 42.0, rtol=1e-10, atol=0)


@pytest.mark.parametrize("engine", [O.ENGINE_BANDED, O.ENGINE_SEPARABLE])
def test_single_fourier_mode(engine):
    # test_periodic.jl:241-285: F[mode] = Ĝ(|k|) mode at every point of a doubly periodic box
    Nx, dx = 32, 62.5
    xs = cart_axes(Nx, dx)
    L = dx * Nx
    g = oracle_grid_cart(xs, xs)
    X, Y = np.meshgrid(xs, xs)
    f = np.sin(2 * np.pi * 2 * X / L) * np.cos(2 * np.pi * 3 * Y / L)
    k = math.hypot(2 * np.pi * 2 / L, 2 * np.pi * 3 / L)
    G = math.exp(-k ** 2 * 750.0 ** 2 / (4 * 6.0))
    out = O.filter_fields(g, [f], kind=O.GAUSSIAN, alpha=6.0, ell=750.0, engine=engine)[0]
    assert np.max(np.abs(out - G * f)) / np.max(np.abs(G * f)) < 1e-8


def test_sphere_brute_force_over_the_pole():
    # test_filtering.jl:733-766: all-pairs haversine sum at high latitude, where the ball crosses
    # the pole and the ring must be counted once
    nlat, nlon = 60, 120
    lat = -np.pi / 2 + 1e-6 + (np.pi - 2e-6) / (nlat - 1) * np.arange(nlat)
    lon = 2 * np.pi / nlon * np.arange(nlon)
    g = O.Grid([lon, lat], spherical=True, R=R_EARTH, periodic=[True, False], period=[2 * np.pi, 0.0],
               uniform=[True, True])
    LON, LAT = np.meshgrid(lon, lat)
    fld = np.cos(3 * LON) * np.exp(-((np.abs(LAT) - 1.40) / 0.12) ** 2)
    scale = 3.0e6
    rad = O.kernel_radius(O.GAUSSIAN, 6.0, scale)
    assert rad > R_EARTH * (np.pi / 2 - 1.33)
    out = O.filter_fields(g, [fld], kind=O.GAUSSIAN, alpha=6.0, ell=scale)[0]
    area = np.array([O.area(g, 0, j) for j in range(nlat)])
    for j in (1, 3, nlat - 7, nlat - 4, nlat - 2):
        for i in (0, 36):
            s2 = np.sin((LAT - lat[j]) / 2) ** 2 + np.cos(lat[j]) * np.cos(LAT) * np.sin((LON - lon[i]) / 2) ** 2
            d = 2 * R_EARTH * np.arcsin(np.sqrt(np.minimum(s2, 1.0)))
            w = np.where(d <= rad, np.exp(-6.0 * (d / scale) ** 2), 0.0) * area[:, None]
            ref = np.sum(w * fld) / np.sum(w)
            assert out[j, i] == pytest.approx(ref, rel=1e-12, abs=1e-14)


def test_cartesian_brute_force_disk_masked():
    # direct disk sum (minimum image on a periodic axis), ZeroFill and Deformable, masked cells
    nx, ny, dx = 30, 24, 2.0
    xs, ys = cart_axes(nx, dx), cart_axes(ny, dx)
    rng = np.random.Generator(np.random.PCG64(3))
    mask = rng.uniform(size=(ny, nx)) > 0.2
    f = rng.standard_normal((ny, nx))
    for periodic in (True, False):
        g = oracle_grid_cart(xs, ys, periodic=(periodic, periodic), mask=mask)
        ell = 7.0
        rad = O.kernel_radius(O.GAUSSIAN, 6.0, ell)
        for strat in (O.ZEROFILL, O.DEFORMABLE):
            out = O.filter_fields(g, [f], kind=O.GAUSSIAN, alpha=6.0, ell=ell, strategy=strat)[0]
            for (j, i) in ((0, 0), (5, 7), (23, 29), (12, 15)):
                num = den = 0.0
                for jj in range(ny):
                    for ii in range(nx):
                        di, dj = ii - i, jj - j
                        if periodic:
                            di = (di + nx // 2) % nx - nx // 2
                            dj = (dj + ny // 2) % ny - ny // 2
                        d = math.hypot(di * dx, dj * dx)
                        if d > rad:
                            continue
                        w = math.exp(-6.0 * (d / ell) ** 2)
                        if strat == O.ZEROFILL or mask[jj, ii]:
                            den += w
                        if mask[jj, ii]:
                            num += w * f[jj, ii]
                ref = num / den if (mask[j, i] and den > 1e-15) else 0.0
                assert out[j, i] == pytest.approx(ref, rel=1e-12, abs=1e-14)


def test_separable_matches_square_brute_force():
    # test_filtering.jl:509-545: the separable engine is the square-truncated Gaussian
    nx, ny, dx = 28, 22, 1.0
    xs, ys = cart_axes(nx, dx), cart_axes(ny, dx)
    g = oracle_grid_cart(xs, ys, periodic=(False, False))
    f = smooth_random_field(g.field_shape, 5, 1)[0]
    ell = 4.0
    out = O.apply_separable(g, 6.0, ell, O.ZEROFILL, [f])[0]
    lim = math.ceil(O.kernel_radius(O.GAUSSIAN, 6.0, ell) / dx)
    for (j, i) in ((0, 0), (10, 10), (21, 3)):
        num = den = 0.0
        for jj in range(max(0, j - lim), min(ny, j + lim + 1)):
            for ii in range(max(0, i - lim), min(nx, i + lim + 1)):
                w = math.exp(-6.0 * (((ii - i) * dx / ell) ** 2 + ((jj - j) * dx / ell) ** 2))
                num += w * f[jj, ii]
                den += w
        assert out[j, i] == pytest.approx(num / den, rel=1e-10)


def test_coast_artefacts_reference_numbers():
    # test_filtering.jl:918-961 — reference-held numbers for a straight coast, Gaussian ℓ = 16 cells
    N = 200
    xs = cart_axes(N, 1.0)
    icoast = 100                                  # 0-based first masked column
    mask = np.ones((N, N), bool)
    mask[:, icoast:] = False
    g = oracle_grid_cart(xs, xs, periodic=(False, False), mask=mask)
    ell, jm = 16.0, N // 2
    xc = np.broadcast_to(xs - xs[icoast], (N, N))
    ds = (1, 4, 8, 16)
    res = {}
    for strat in (O.ZEROFILL, O.DEFORMABLE):
        f1, fx, fx2 = O.apply_separable(g, 6.0, ell, strat, [np.ones((N, N)), xc, xc ** 2])
        sint = math.sqrt(fx2[jm, 49] / f1[jm, 49] - (fx[jm, 49] / f1[jm, 49]) ** 2)
        mass = [f1[jm, icoast - d] for d in ds]
        cent = [fx[jm, icoast - d] / f1[jm, icoast - d] - xc[jm, icoast - d] for d in ds]
        wid = [math.sqrt(fx2[jm, icoast - d] / f1[jm, icoast - d] - (fx[jm, icoast - d] / f1[jm, icoast - d]) ** 2) / sint
               for d in ds]
        res[strat] = (np.array(mass), np.array(cent), np.array(wid))
    mz, cz, wz = res[O.ZEROFILL]
    md, cd, wd = res[O.DEFORMABLE]
    assert np.allclose(cz, cd, rtol=1e-12) and np.allclose(wz, wd, rtol=1e-12)
    assert np.allclose(cz / ell, [-0.211, -0.111, -0.032, 0.0], atol=5e-3)
    assert np.allclose(wz, [0.621, 0.747, 0.897, 0.998], atol=5e-3)
    assert np.allclose(mz, [0.543, 0.776, 0.948, 0.9996], atol=5e-3)
    assert np.allclose(md, 1.0, atol=1e-12)
    # the disk-truncated banded engine agrees with the separable one to the truncation tolerance
    fb = O.filter_fields(g, [np.ones((N, N))], kind=O.GAUSSIAN, alpha=6.0, ell=ell)[0]
    assert np.allclose([fb[jm, icoast - d] for d in ds], mz, rtol=1e-7)


# ------------------------------------------------------- derivatives (test_derivatives.jl) --

def test_derivative_linear_and_edges():
    # test_derivatives.jl: ∂(3x+1)/∂x = 3 at interior AND boundary points
    x = cart_axes(6, 2.0)
    g = oracle_grid_cart(x, x, periodic=(False, False))
    X = np.broadcast_to(x, (6, 6))
    d = O.deriv(g, 0, 3.0 * X + 1.0)
    assert np.allclose(d, 3.0, rtol=1e-13)
    # 3 nodes reproduce a quadratic exactly, centred and one-sided (CHANGELOG: ends are 2nd order)
    d2 = O.deriv(g, 0, X ** 2)
    assert np.allclose(d2, 2 * X, rtol=1e-12, atol=1e-12)


def test_derivative_second_order_convergence():
    # test_derivatives.jl convergence-rate test: interior rate ≈ 2
    L = 100.0
    k = 2 * np.pi / L
    errs = []
    for N in (40, 80, 160, 320):
        x = np.linspace(0.0, L, N)
        y = cart_axes(3, 1.0)
        g = O.Grid([x, y], periodic=[False, False], uniform=[False, True])
        f = np.broadcast_to(np.sin(k * x), (3, N)).copy()
        d = O.deriv(g, 0, f)
        errs.append(np.max(np.abs(d[:, 2:N - 2] - k * np.cos(k * x[2:N - 2]))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert all(1.8 < r < 2.2 for r in rates)


def test_derivative_nonuniform_quadratic_and_periodic_seam():
    z = np.concatenate([[0.0], np.cumsum(0.7 * 1.2 ** np.arange(9))])
    g = O.Grid([cart_axes(3, 1.0), cart_axes(3, 1.0), z], periodic=[True, True, False])
    Z = np.broadcast_to(z[:, None, None], (len(z), 3, 3)).copy()
    assert np.allclose(O.deriv(g, 2, Z ** 2 - 3 * Z), 2 * Z - 3, rtol=1e-11, atol=1e-11)
    # a periodic seam differences against the wrapped neighbour (CHANGELOG ddx! seam fix)
    N = 64
    x = cart_axes(N, 1000.0)
    gp = oracle_grid_cart(x, x[:4], periodic=(True, False))
    f = np.broadcast_to(np.sin(2 * np.pi * x / (N * 1000.0)), (4, N)).copy()
    d = O.deriv(gp, 0, f)
    exact = 2 * np.pi / (N * 1000.0) * np.cos(2 * np.pi * x / (N * 1000.0))
    err = np.abs(d[0] - exact)
    assert err[0] < 2 * np.median(err) and err[-1] < 2 * np.median(err)


def test_derivative_mask_runs():
    # ReduceInRun: no stencil reaches across a masked cell; a run of two gets the 2-point value,
    # an isolated cell 0, masked cells 0 — all exact for a linear field
    x = cart_axes(12, 1.0)
    row = np.array([1, 1, 1, 1, 0, 1, 1, 0, 1, 0, 1, 1], bool)
    mask = np.broadcast_to(row, (3, 12)).copy()
    g = oracle_grid_cart(x, x[:3], periodic=(False, False), mask=mask)
    f = np.broadcast_to(5.0 * x - 2.0, (3, 12)).copy()
    f[:, ~row] = 1e6                                 # garbage under the mask must not leak
    d = O.deriv(g, 0, f)[1]
    assert np.allclose(d[row & (np.arange(12) != 8)], 5.0)
    assert d[8] == 0.0 and np.all(d[~row] == 0.0)


def test_sphere_derivative_metric():
    lam, phi = sphere_axes(360, 180)
    g = oracle_grid_sphere(lam, phi)
    LON, LAT = np.meshgrid(lam, phi)
    f = np.sin(LON) * np.cos(LAT) ** 2
    dx = O.deriv(g, 0, f)       # (1/(R cosφ)) ∂λ f = cosλ cosφ / R
    dy = O.deriv(g, 1, f)       # (1/R) ∂φ f = −2 sinλ cosφ sinφ / R
    inner = slice(5, -5)
    ex = (np.cos(LON) * np.cos(LAT) / R_EARTH)[inner]
    ey = (-2 * np.sin(LON) * np.cos(LAT) * np.sin(LAT) / R_EARTH)[inner]
    assert np.max(np.abs(dx[inner] - ex)) < 1e-4 / R_EARTH
    assert np.max(np.abs(dy[inner] - ey)) < 1e-3 / R_EARTH


# -------------------------------------------------------------------- Π (test_validation.jl) --

@pytest.mark.parametrize("engine,kind,alpha", [(O.ENGINE_SEPARABLE, O.GAUSSIAN, 6.0), (O.ENGINE_BANDED, O.GAUSSIAN, 6.0),
                                               (O.ENGINE_BANDED, O.TOPHAT, 0.0)])
def test_linear_fields_pi_zero_and_tau(engine, kind, alpha):
    # test_validation.jl:196-267: linear fields reproduce, Π = 0, pure strain τ11 = τ22 = Ω²σ²
    N, dx = 64, 1000.0
    xs = -(N // 2) * dx + dx * np.arange(N)
    g = oracle_grid_cart(xs, xs, periodic=(False, False))
    Om, ell = 1e-4, 6 * dx
    pad = math.ceil(O.kernel_radius(kind, alpha, ell) / dx) + 1
    ii = slice(pad, N - pad)
    X, Y = np.meshgrid(xs, xs)
    f1, fy, fy2 = O.filter_fields(g, [np.ones((N, N)), Y, Y ** 2], kind=kind, alpha=alpha, ell=ell, engine=engine)
    s2 = fy2 / f1 - (fy / f1) ** 2
    s2i = s2[N // 2, N // 2]
    assert np.max(np.abs(s2[ii, ii] - s2i)) < 1e-9 * s2i
    if kind == O.GAUSSIAN:
        assert s2i == pytest.approx(ell ** 2 / 12, rel=1e-9 if engine == O.ENGINE_SEPARABLE else 1e-7)
    else:
        assert s2i == pytest.approx(ell ** 2 / 16, rel=0.06)
    for name, u, v in (("rotation", -Om * Y, Om * X), ("strain", Om * Y, Om * X)):
        Pi, ub, vb, _ = O.compute_pi(g, u, v, kind=kind, alpha=alpha, ell=ell, engine=engine)
        assert np.max(np.abs(ub[ii, ii] - u[ii, ii])) < 1e-12 * np.max(np.abs(u))
        assert np.max(np.abs(Pi[ii, ii])) < 1e-11 * Om ** 3 * s2i
    u, v = Om * Y, Om * X
    fu, fv, fuu, fuv, fvv = O.filter_fields(g, [u, v, u * u, u * v, v * v], kind=kind, alpha=alpha, ell=ell,
                                            engine=engine)
    t11, t12, t22 = fuu - fu * fu, fuv - fu * fv, fvv - fv * fv
    assert np.allclose(t11[ii, ii], Om ** 2 * s2i, rtol=1e-9)
    assert np.allclose(t22[ii, ii], Om ** 2 * s2i, rtol=1e-9)
    assert np.max(np.abs(t12[ii, ii])) < 1e-9 * Om ** 2 * s2i


def test_pi_galilean_invariance_periodic():
    N, dx = 48, 1.0
    x = cart_axes(N, dx)
    g = oracle_grid_cart(x, x)
    u, v = smooth_random_field((N, N), 4, 2)
    P0, _, _, _ = O.compute_pi(g, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=5.0)
    P1, _, _, _ = O.compute_pi(g, u + 3.0, v - 2.0, kind=O.GAUSSIAN, alpha=6.0, ell=5.0)
    assert np.max(np.abs(P1 - P0)) < 1e-9 * np.max(np.abs(P0))


def test_sphere_rigid_rotation_is_strain_free():
    # rigid rotation about a tilted axis is a Killing field: the filtered flow is (nearly) rigid and
    # S̄ — including the curvature terms — vanishes, so Π is tiny against Ω·τ; a wrong sign or a
    # dropped tan φ term breaks this at O(Ω·τ)
    lam, phi = sphere_axes(180, 90)
    g = oracle_grid_sphere(lam, phi)
    LON, LAT = np.meshgrid(lam, phi)
    Om = 1e-5
    axis = np.array([1.0, 0.0, 1.0]) / math.sqrt(2)
    r = np.stack([np.cos(LAT) * np.cos(LON), np.cos(LAT) * np.sin(LON), np.sin(LAT)]) * R_EARTH
    U = Om * np.cross(axis[:, None, None], r, axis=0)
    ee = np.stack([-np.sin(LON), np.cos(LON), 0 * LON])
    en = np.stack([-np.sin(LAT) * np.cos(LON), -np.sin(LAT) * np.sin(LON), np.cos(LAT)])
    u, v = np.sum(U * ee, 0), np.sum(U * en, 0)
    ell = 1.0e6
    Pi, _, _, _ = O.compute_pi(g, u, v, kind=O.GAUSSIAN, alpha=4.0, ell=ell)
    tau_scale = (Om * O.kernel_radius(O.GAUSSIAN, 4.0, ell)) ** 2
    assert np.max(np.abs(Pi[5:-5])) < 1e-3 * Om * tau_scale


def test_sphere_longitude_shift_equivariance():
    lam, phi = sphere_axes(72, 36)
    rng = np.random.Generator(np.random.PCG64(8))
    mask = rng.uniform(size=(36, 72)) > 0.2
    g = oracle_grid_sphere(lam, phi, mask=mask)
    u, v = smooth_random_field((36, 72), 9, 2)
    P0, _, _, _ = O.compute_pi(g, u, v, kind=O.GAUSSIAN, alpha=4.0, ell=2.0e6)
    sh = 7
    g2 = oracle_grid_sphere(lam, phi, mask=np.roll(mask, sh, 1))
    P1, _, _, _ = O.compute_pi(g2, np.roll(u, sh, 1), np.roll(v, sh, 1), kind=O.GAUSSIAN, alpha=4.0, ell=2.0e6)
    assert np.max(np.abs(np.roll(P0, sh, 1) - P1)) < 1e-10 * np.max(np.abs(P0))


# ------------------------------------------------------------------------ 3-D (test_nd.jl) --

def test_planar_engines_refuse_a_3d_grid():
    # the banded / separable / prefix-sum / transform restatements walk one (x, y) plane only
    x = cart_axes(8, 1.0)
    g3 = oracle_grid_cart(x, x, x[:4], periodic=(True, True, False))
    f = np.zeros((4, 8, 8))
    for call in (lambda: O.compute_pi(g3, f, f, f, kind=O.TOPHAT, alpha=0.0, ell=3.0),
                 lambda: O.filter_fields(g3, [f], kind=O.TOPHAT, alpha=0.0, ell=3.0),
                 lambda: O.apply_separable(g3, 6.0, 3.0, O.ZEROFILL, [f]),
                 lambda: O.apply_prefix(g3, 3.0, O.ZEROFILL, [f]),
                 lambda: O.apply_spectral(g3, O.TOPHAT, 0.0, 3.0, O.ZEROFILL, [f])):
        with pytest.raises(ValueError):
            call()



def test_3d_z_invariant_reduces_to_2d():
    x = cart_axes(21, 1.0)
    z = cart_axes(4, 100.0)
    g3 = oracle_grid_cart(x, x, z, periodic=(False, False, False))
    g2 = oracle_grid_cart(x, x, periodic=(False, False))
    f2 = smooth_random_field((21, 21), 2, 1)[0]
    f3 = np.broadcast_to(f2, (4, 21, 21)).copy()
    o3 = O.apply_nd3(g3, O.TOPHAT, 0.0, 6.0, O.ZEROFILL, [f3])[0]
    o2 = O.filter_fields(g2, [f2], kind=O.TOPHAT, alpha=0.0, ell=6.0)[0]
    for k in range(4):
        assert np.allclose(o3[k], o2, rtol=1e-13, atol=1e-15)
    # test_nd.jl:61-73: z-invariant (u, v), w = 0 → the 3-D six-term Π equals the 2-D Π per layer
    u2, v2 = smooth_random_field((25, 25), 3, 2)
    x5 = cart_axes(25, 1.0)
    z5 = cart_axes(4, 50.0)
    g3 = oracle_grid_cart(x5, x5, z5, periodic=(False, False, False))
    g2 = oracle_grid_cart(x5, x5, periodic=(False, False))
    P3, _, _, _ = O.compute_pi(g3, np.broadcast_to(u2, (4, 25, 25)), np.broadcast_to(v2, (4, 25, 25)),
                               np.zeros((4, 25, 25)), kind=O.TOPHAT, alpha=0.0, ell=5.0, engine=O.ENGINE_ND3)
    P2, _, _, _ = O.compute_pi(g2, u2, v2, kind=O.TOPHAT, alpha=0.0, ell=5.0)
    for k in range(4):
        assert np.allclose(P3[k], P2, rtol=1e-11, atol=1e-14 * np.max(np.abs(P2)))
    # constant velocity → Π = 0
    Pc, _, _, _ = O.compute_pi(g3, np.full((4, 25, 25), 2.0), np.full((4, 25, 25), -3.0), np.full((4, 25, 25), 0.5),
                               kind=O.TOPHAT, alpha=0.0, ell=5.0, engine=O.ENGINE_ND3)
    assert np.max(np.abs(Pc)) < 1e-9


def test_3d_ball_brute_force_stretched():
    xs, ys = cart_axes(10, 1.0), cart_axes(9, 1.0)
    zs = np.concatenate([[0.0], np.cumsum(0.6 * 1.25 ** np.arange(7))])
    g = oracle_grid_cart(xs, ys, zs, periodic=(True, True, False))
    f = smooth_random_field((8, 9, 10), 6, 1)[0]
    ell = 5.0
    out = O.apply_nd3(g, O.TOPHAT, 0.0, ell, O.ZEROFILL, [f])[0]
    rad = ell / 2
    wz = np.array([O.area(g, 0, 0, k) for k in range(8)])
    for (k, j, i) in ((0, 0, 0), (3, 4, 5), (7, 8, 9)):
        num = den = 0.0
        for kk in range(8):
            for jj in range(9):
                for ii in range(10):
                    dxp = (ii - i + 5) % 10 - 5
                    dyp = (jj - j + 4) % 9 - 4
                    d = math.sqrt(dxp ** 2 + dyp ** 2 + (zs[kk] - zs[k]) ** 2)
                    if d <= rad:
                        num += wz[kk] * f[kk, jj, ii]
                        den += wz[kk]
        assert out[k, j, i] == pytest.approx(num / den, rel=1e-12)


# ------------------------------------------------------------------ energy / spectrum --

def test_energy_of_constant_field_and_spectral_density():
    # test_spectral.jl:324-364
    x = cart_axes(51, 2000.0)
    g = oracle_grid_cart(x, x, periodic=(False, False))
    U, V = 0.5, 0.3
    u, v = np.full((51, 51), U), np.full((51, 51), V)
    scales = [5000.0, 10000.0, 20000.0, 40000.0]
    for kind, alpha in ((O.TOPHAT, 0.0), (O.GAUSSIAN, 6.0)):
        E = O.cumulative_energy(g, u, v, scales, kind=kind, alpha=alpha)
        assert np.allclose(E, 0.5 * (U ** 2 + V ** 2), rtol=1e-6)
    E = O.cumulative_energy(g, u, v, scales, kind=O.GAUSSIAN, alpha=6.0)
    Et = O.spectral_density(E, 1.0 / np.asarray(scales))
    assert np.all(np.abs(Et) < 1e-6 * 0.5 * (U ** 2 + V ** 2) * max(scales))
    kk = np.arange(1.0, 6.0)
    assert O.spectral_density(kk ** 2, kk)[2] == pytest.approx(2 * kk[2])
    assert np.all(O.spectral_density(np.array([1.0]), np.array([1.0])) == 0.0)


# ------------------------------------------- prefix-sum top-hat (Filtering.jl:751-1097) --------

@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("strategy", [O.ZEROFILL, O.DEFORMABLE])
def test_prefix_tophat_equals_disk_footprint(periodic, strategy):
    # the prefix-sum restatement (per-row prefix sums, two monotone pointers) against the independent
    # disk-footprint restatement, which the brute-force tests above pin; radius below half the box,
    # where the footprint's modular wrap and the prefix engine's minimum image agree
    xs, ys = cart_axes(40, 1000.0, 500.0), cart_axes(30, 1250.0)
    rng = np.random.Generator(np.random.PCG64(7))
    mask = (rng.random((30, 40)) > 0.2).astype(np.uint8)
    og = oracle_grid_cart(xs, ys, periodic=(periodic, periodic), mask=mask)
    f = smooth_random_field((30, 40), 5)[0]
    for ell in (3300.0, 7100.0, 12345.0, 30000.0):
        a = O.apply_prefix(og, ell, strategy, [f])[0]
        b = O.filter_fields(og, [f], kind=O.TOPHAT, alpha=0.0, ell=ell, strategy=strategy)[0]
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


def test_prefix_tophat_sphere_equals_haversine_cap():
    # the cap |Δλ| ≤ acos((cos(rad/R) − sinφ₁sinφ₂)/(cosφ₁cosφ₂)) is the haversine disk d ≤ rad
    lam, phi = sphere_axes(72, 36)
    og = oracle_grid_sphere(lam, phi)
    f = smooth_random_field((36, 72), 6)[0]
    for ell in (5e5, 1.3e6, 3e6):
        a = O.apply_prefix(og, ell, O.ZEROFILL, [f])[0]
        b = O.filter_fields(og, [f], kind=O.TOPHAT, alpha=0.0, ell=ell)[0]
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # metric_band closed forms: the equator row at rad = one degree spans ±1°; beyond rad nothing
    rad = R_EARTH * math.radians(1.0)
    assert O.metric_band(og, 0.0, 0.0, rad) == pytest.approx(math.radians(1.0), rel=1e-12)
    assert O.metric_band(og, 0.0, math.radians(1.5), rad) < 0.0


def test_prefix_tophat_nonuniform_axis_brute_force():
    # a stretched axis 1 (the two-pointer path): brute-force disk sum with the cell measures
    rng = np.random.Generator(np.random.PCG64(8))
    xs = np.cumsum(rng.uniform(600.0, 1400.0, 24))
    ys = cart_axes(18, 1000.0)
    mask = (rng.random((18, 24)) > 0.25).astype(np.uint8)
    og = O.Grid([xs, ys], spherical=False, periodic=[False, False], period=[0.0, 0.0], mask=mask,
                uniform=[False, True])
    f = smooth_random_field((18, 24), 9)[0]
    A = np.array([[O.area(og, i, j, 0) for i in range(24)] for j in range(18)])
    for ell in (2500.0, 6100.0):
        rad = ell / 2
        got = O.apply_prefix(og, ell, O.ZEROFILL, [f])[0]
        exp = np.zeros_like(f)
        for j in range(18):
            for i in range(24):
                if not mask[j, i]:
                    continue
                d = np.sqrt((xs[None, :] - xs[i]) ** 2 + (ys[:, None] - ys[j]) ** 2)
                sup = d <= rad
                exp[j, i] = np.sum((A * f * mask)[sup]) / np.sum(A[sup])
        assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp))


# ------------------------------------------------ spectral filter (FFTWExt.jl:121-208) --------

def _np_spectral(f, mask, Lx, Ly, transfer, deformable):
    """numpy's real FFT (an independent transform) of the same definition: irfft2(rfft2(mask·f)·Ĝ)."""
    Ny, Nx = f.shape
    kx = 2 * np.pi * np.fft.rfftfreq(Nx, Lx / Nx)
    ky = 2 * np.pi * np.fft.fftfreq(Ny, Ly / Ny)
    G = transfer(np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2))
    filt = lambda a: np.fft.irfft2(np.fft.rfft2(a) * G, s=(Ny, Nx))  # noqa: E731
    m = np.ones_like(f) if mask is None else mask.astype(np.float64)
    out = filt(m * f)
    if deformable and mask is not None:
        r = filt(m)
        out = out * np.where(np.abs(r) >= 0.01, 1.0 / np.where(r == 0, 1, r), 0.0)
    return out


@pytest.mark.parametrize("kind", [O.GAUSSIAN, O.TOPHAT])
@pytest.mark.parametrize("strategy", [O.ZEROFILL, O.DEFORMABLE])
@pytest.mark.parametrize("shape", [(24, 32), (21, 30)])
def test_spectral_filter_matches_numpy_fft(kind, strategy, shape):
    import scipy.special as sp
    Ny, Nx = shape
    xs, ys = cart_axes(Nx, 1000.0), cart_axes(Ny, 1500.0)
    rng = np.random.Generator(np.random.PCG64(40))
    mask = (rng.random(shape) > 0.2).astype(np.uint8)
    og = oracle_grid_cart(xs, ys, periodic=(True, True), mask=mask)
    f = smooth_random_field(shape, 41)[0] + 0.3
    ell, alpha = 5000.0, 6.0
    if kind == O.GAUSSIAN:
        tr = lambda k: np.exp(-k ** 2 * ell ** 2 / (4 * alpha))  # noqa: E731
    else:
        tr = lambda k: np.where(k == 0, 1.0, 2 * sp.j1(k * ell / 2) / np.where(k == 0, 1, k * ell / 2))  # noqa: E731
    got = O.apply_spectral(og, kind, alpha, ell, strategy, [f])[0]
    exp = _np_spectral(f, mask, Nx * 1000.0, Ny * 1500.0, tr, strategy == O.DEFORMABLE)
    assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp))


def test_spectral_filter_single_mode_and_constant():
    # a Fourier mode is an eigenfunction: multiplied by Ĝ(|k|) exactly; Ĝ(0) = 1 keeps a constant
    Nx, Ny, L = 32, 24, 32000.0
    xs, ys = cart_axes(Nx, L / Nx), cart_axes(Ny, L / Nx)
    og = oracle_grid_cart(xs, ys, periodic=(True, True))
    X, Y = np.meshgrid(xs, ys)
    k1, k2 = 2 * np.pi * 3 / L, 2 * np.pi * 2 / (Ny * L / Nx)
    f = np.cos(k1 * X + k2 * Y)
    ell, alpha = 4000.0, 6.0
    got = O.apply_spectral(og, O.GAUSSIAN, alpha, ell, O.ZEROFILL, [f, np.full_like(f, 2.5)])
    assert np.max(np.abs(got[0] - math.exp(-(k1 ** 2 + k2 ** 2) * ell ** 2 / (4 * alpha)) * f)) <= 1e-13
    assert np.max(np.abs(got[1] - 2.5)) <= 1e-13


# ------------------------------------- row-subset Π (for wide footprints at full size) -----------

def test_pi_rows_equals_full_pi_on_its_rows():
    lam, phi = sphere_axes(144, 72)
    f = smooth_random_field((72, 144), 3, 3)
    g = oracle_grid_sphere(lam, phi, mask=f[2] > np.quantile(f[2], 0.3))
    rows = [0, 1, 10, 35, 70, 71]
    for ell in (8e5, 3e6):
        P, ub, vb, _ = O.compute_pi(g, f[0], f[1], kind=O.GAUSSIAN, alpha=4.0, ell=ell)
        Pr, ubr, vbr = O.compute_pi_rows(g, f[0], f[1], rows, kind=O.GAUSSIAN, alpha=4.0, ell=ell)
        assert np.array_equal(P[rows], Pr[rows])
        assert np.array_equal(ub[rows], ubr[rows]) and np.array_equal(vb[rows], vbr[rows])
        others = np.setdiff1d(np.arange(72), rows)
        assert np.isnan(Pr[others]).all()
    # the band-restricted footprint holds exactly the full footprint's entries on its bands
    full = O.Footprint(g, O.GAUSSIAN, 4.0, 3e6)
    part = O.Footprint(g, O.GAUSSIAN, 4.0, 3e6, bands=[3, 40])
    for b in (3, 40):
        a0, a1 = full.ptr[b], full.ptr[b + 1]
        b0, b1 = part.ptr[b], part.ptr[b + 1]
        assert np.array_equal(full.w[a0:a1], part.w[b0:b1]) and np.array_equal(full.di[a0:a1], part.di[b0:b1])
    assert part.ptr[4] == part.ptr[40]


# ----------------------------- cell widths on non-uniform axes (FlowGeometries measure) ---------

def _stretched(N, L=20.0):
    t = np.linspace(0.0, 1.0, N)
    return L * (t + 0.15 * np.sin(2 * np.pi * t) / (2 * np.pi))   # smooth, monotone, non-uniform


def test_cell_widths_tile_the_axis():
    # A measure tiles the domain: midpoint widths with half cells at bounded ends sum to the extent,
    # and to the period on a periodic axis. On a collected (non-Range) uniform axis the interior widths
    # are the step (test_grids.jl:203-206 checks exactly the interior, 2:N-1, of such an axis).
    z = np.concatenate([[0.0], np.cumsum(0.5 * 1.07 ** np.arange(19))])
    one = np.array([0.0])
    g = O.Grid([one, one, z], periodic=[False, False, False], uniform=[True, True, False])
    w = np.array([O.cell_width(g, 2, k) for k in range(len(z))])
    assert np.sum(w) == pytest.approx(z[-1] - z[0], rel=1e-14)
    assert np.all(w > 0)
    gp = O.Grid([one, one, z], periodic=[False, False, True], period=[0.0, 0.0, z[-1] + 3.0],
                uniform=[True, True, False])
    wp = np.array([O.cell_width(gp, 2, k) for k in range(len(z))])
    assert np.sum(wp) == pytest.approx(z[-1] + 3.0, rel=1e-14)
    x1 = np.arange(0.0, 10001.0, 1000.0)
    gc = O.Grid([x1, one], periodic=[False, False], uniform=[False, True])
    wc = np.array([O.cell_width(gc, 0, i) for i in range(len(x1))])
    assert np.allclose(wc[1:-1], 1000.0, rtol=1e-14)
    with O.mutation(O.MUTATE_FORWARD_WIDTH):      # planted: forward differences over-count the end
        wf = np.array([O.cell_width(g, 2, k) for k in range(len(z))])
    assert abs(np.sum(wf) - (z[-1] - z[0])) > 0.1


def _filtered_quadratic_error(N, ell=2.0):
    # Gaussian (α = 6) filter along a smoothly stretched bounded axis of a 1×1×N grid: away from the
    # ends F[z²] = z₀² + σ², σ² = ℓ²/(2α), up to the truncation at 1e-10. The measure-weighted sum is a
    # quadrature of the kernel integral: midpoint widths (the trapezoid rule) converge at 2nd order,
    # forward-difference widths (a one-sided Riemann sum) at 1st.
    z = _stretched(N)
    one = np.array([0.0])
    g = O.Grid([one, one, z], periodic=[False, False, False], uniform=[True, True, False])
    out = O.apply_nd3(g, O.GAUSSIAN, 6.0, ell, O.ZEROFILL, [(z ** 2)[:, None, None]])[0][:, 0, 0]
    sel = (z > 6.0) & (z < 14.0)
    return np.max(np.abs(out[sel] - (z[sel] ** 2 + ell ** 2 / 12.0)))


def test_filter_on_a_stretched_axis_converges_at_second_order():
    Ns = (81, 161, 321, 641)
    errs = [_filtered_quadratic_error(N) for N in Ns]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert all(1.9 < r < 2.1 for r in rates), rates
    with O.mutation(O.MUTATE_FORWARD_WIDTH):
        bad = [_filtered_quadratic_error(N) for N in Ns]
    bad_rates = [math.log2(bad[i] / bad[i + 1]) for i in range(3)]
    assert all(r < 1.2 for r in bad_rates), bad_rates


# ------------------------- derivative stencil at mask-run ends (CHANGELOG.md:257-259) -------------

def test_derivative_quadratic_exact_at_mask_run_ends():
    # "Bounded-domain end points are now 2nd order ... apply_stencil! keeps three nodes and shifts the
    # window inward" (CHANGELOG.md:257-259): with the ReduceInRun policy a mask edge is a run end, so in
    # every run of three or more active cells the derivative of a quadratic is exact, ends included —
    # on a uniform and on a stretched axis. A two-node end stencil is not.
    row = np.array([1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1], bool)
    n = len(row)
    for x in (cart_axes(n, 0.7, 2.0), np.cumsum(0.4 * 1.09 ** np.arange(n))):
        for axis in (0, 1):
            mask = np.broadcast_to(row, (3, n)).copy() if axis == 0 else np.broadcast_to(row[:, None], (n, 3)).copy()
            xs, ys = (x, cart_axes(3, 1.0)) if axis == 0 else (cart_axes(3, 1.0), x)
            g = O.Grid([xs, ys], periodic=[False, False], mask=mask,
                       uniform=[axis != 0 or np.allclose(np.diff(x), x[1] - x[0]), axis != 1 or np.allclose(np.diff(x), x[1] - x[0])])
            X = np.broadcast_to(xs, (len(ys), len(xs))) if axis == 0 else np.broadcast_to(ys[:, None], (len(ys), len(xs)))
            f = 2.0 * X ** 2 - 3.0 * X + 1.0
            f = np.where(mask, f, 1e6)                       # garbage under the mask must not leak
            want = np.where(mask, 4.0 * X - 3.0, 0.0)
            d = O.deriv(g, axis, f)
            assert np.allclose(d, want, rtol=1e-11, atol=1e-9)
            with O.mutation(O.MUTATE_RUNEND_2NODE):
                dbad = O.deriv(g, axis, f)
            assert np.max(np.abs(dbad - want)) > 0.1


def test_taylor_green_derivatives():
    # test_diagnostics.jl:60-101: u = sin x cos y, v = −cos x sin y on 0:0.1:2π, rtol 0.01 at (10, 10)
    xs = np.arange(0.0, 2 * np.pi, 0.1)
    g = O.Grid([xs, xs], periodic=[False, False], uniform=[True, True])
    X, Y = np.meshgrid(xs, xs)
    u, v = np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y)
    j = i = 9
    x, y = xs[i], xs[j]
    assert O.deriv(g, 0, u)[j, i] == pytest.approx(np.cos(x) * np.cos(y), rel=0.01)
    assert O.deriv(g, 1, u)[j, i] == pytest.approx(-np.sin(x) * np.sin(y), rel=0.01)
    assert O.deriv(g, 0, v)[j, i] == pytest.approx(np.sin(x) * np.sin(y), rel=0.01)
    assert O.deriv(g, 1, v)[j, i] == pytest.approx(-np.cos(x) * np.cos(y), rel=0.01)


# ------------------------------------ identities the reference's test suite asserts --------------

def _bump(N, ic, jc, rad):
    """test_filtering.jl:841-848 (0-based centre): (1 − r²)⁴ inside radius `rad` cells."""
    j, i = np.mgrid[0:N, 0:N].astype(np.float64)
    r2 = ((i - ic) ** 2 + (j - jc) ** 2) / rad ** 2
    return np.where(r2 < 1, (1 - r2) ** 4, 0.0)


def test_zerofill_commutes_with_ddx_deformable_does_not():
    # test_filtering.jl:850-916
    N, dx = 96, 1000.0
    xs = cart_axes(N, dx)
    icoast = 60                                    # first land column (0-based)
    mask = np.ones((N, N), bool)
    mask[:, icoast:] = False
    g = oracle_grid_cart(xs, xs, periodic=(False, False), mask=mask)
    ell = 8000.0                                   # top-hat, radius 4 cells
    f = _bump(N, 48, 47, 8)
    df = O.deriv(g, 0, f)
    assert np.all(f[:, icoast:] == 0) and np.all(df[:, icoast:] == 0)
    fp = O.Footprint(g, O.TOPHAT, 0.0, ell)
    rel, edge, drift = {}, {}, {}
    for st in (O.ZEROFILL, O.DEFORMABLE):
        lhs = O.apply_banded(g, fp, st, [df])[0]
        ff = O.apply_banded(g, fp, st, [f])[0]
        rhs = O.deriv(g, 0, ff)
        ii = slice(1, icoast - 2)
        s = np.max(np.abs(rhs[:, ii]))
        rel[st] = np.max(np.abs(lhs[:, ii] - rhs[:, ii])) / s
        edge[st] = np.max(np.abs(lhs[:, icoast - 1] - rhs[:, icoast - 1])) / s
        drift[st] = (np.sum(ff) - np.sum(f)) / np.sum(f)
    assert rel[O.ZEROFILL] < 1e-13
    assert rel[O.DEFORMABLE] > 1e-3 and rel[O.DEFORMABLE] > 1e9 * rel[O.ZEROFILL]
    assert edge[O.ZEROFILL] > 1e-3
    assert abs(drift[O.ZEROFILL]) < 1e-13 and abs(drift[O.DEFORMABLE]) > 1e-6
    # unmasked: both strategies coincide and commute, top-hat and Gaussian alike
    gf = oracle_grid_cart(xs, xs, periodic=(False, False))
    fu = _bump(N, 47, 47, 10)
    dfu = O.deriv(gf, 0, fu)
    for kind, alpha in ((O.TOPHAT, 0.0), (O.GAUSSIAN, 6.0)):
        lhs = O.filter_fields(gf, [dfu], kind=kind, alpha=alpha, ell=ell)[0]
        rhs = O.deriv(gf, 0, O.filter_fields(gf, [fu], kind=kind, alpha=alpha, ell=ell)[0])
        ii = slice(1, N - 1)
        assert np.max(np.abs(lhs[:, ii] - rhs[:, ii])) < 1e-13 * np.max(np.abs(rhs[:, ii]))


_PHI_A, _PHI_B = 0.7548776662466927, 0.5698402909980532


def _spectral_field(N, dx, alpha, symbol="exact"):
    """test_validation.jl:22-40: ψ̂ = |k|^−(α+3)/2 with irrational-rotation phases (1-based i, j; i
    along x), round-tripped through a real field; û = −i κ(k_y) ψ̂, v̂ = i κ(k_x) ψ̂ with κ(k) = k
    (spectral) or sin(kΔx)/Δx (divergence-free for the centred 3-point stencil)."""
    n = np.arange(N)
    k = 2 * np.pi / (N * dx) * np.where(n <= N // 2, n, n - N)
    I, J = np.meshgrid(n + 1, n + 1)                 # [j, i], 1-based indices
    KX, KY = k[I - 1], k[J - 1]
    km = np.sqrt(KX ** 2 + KY ** 2)
    phase = 2 * np.pi * np.mod(_PHI_A * I + _PHI_B * J, 1.0)
    psi = np.where(km > 0, np.where(km > 0, km, 1.0) ** (-(alpha + 3) / 2), 0.0) * np.exp(1j * phase)
    psi = np.fft.fft2(np.real(np.fft.ifft2(psi)))
    sym = (lambda q: q) if symbol == "exact" else (lambda q: np.sin(q * dx) / dx)
    return np.real(np.fft.ifft2(-1j * sym(KY) * psi)), np.real(np.fft.ifft2(1j * sym(KX) * psi))


def _logfit(x, y):
    lx, ly = np.log10(x), np.log10(y)
    return float(np.sum((lx - lx.mean()) * (ly - ly.mean())) / np.sum((lx - lx.mean()) ** 2))


def _fourier_slope(u, v, dx, nlo, nhi):
    N = u.shape[0]
    n = np.arange(N)
    k = 2 * np.pi / (N * dx) * np.where(n <= N // 2, n, n - N)
    dk = 2 * np.pi / (N * dx)
    uh, vh = np.fft.fft2(u) / N ** 2, np.fft.fft2(v) / N ** 2
    sh = np.rint(np.sqrt(k[None, :] ** 2 + k[:, None] ** 2) / dk).astype(int)
    E = np.zeros(N // 2 + 1)
    ok = (sh >= 1) & (sh <= N // 2)
    np.add.at(E, sh[ok], 0.5 * (np.abs(uh[ok]) ** 2 + np.abs(vh[ok]) ** 2))
    ns = [q for q in range(nlo, nhi + 1) if E[q] > 0]
    return _logfit(np.array(ns, float), E[ns])


def test_filter_commutes_with_divergence():
    # test_validation.jl:146-182, for the reference's default engines of each kernel (separable
    # Gaussian, prefix-sum top-hat) and the banded footprint
    dx = 1000.0
    for N in (64, 96):
        xs = cart_axes(N, dx)
        g = oracle_grid_cart(xs, xs, periodic=(True, True))
        u, v = _spectral_field(N, dx, 3.0, symbol="stencil")
        du, dv = O.deriv(g, 0, u), O.deriv(g, 1, v)
        divu = du + dv
        assert np.max(np.abs(divu)) < 1e-12 * max(np.max(np.abs(du)), np.max(np.abs(dv)))
        for kind, alpha, engines in ((O.GAUSSIAN, 6.0, (O.ENGINE_SEPARABLE, O.ENGINE_BANDED)),
                                     (O.TOPHAT, 0.0, (O.ENGINE_PREFIX, O.ENGINE_BANDED))):
            for ell in (8 * dx, 16 * dx):
                for eng in engines:
                    ub, vb, fdiv = O.filter_fields(g, [u, v, divu], kind=kind, alpha=alpha, ell=ell, engine=eng)
                    dub, dvb = O.deriv(g, 0, ub), O.deriv(g, 1, vb)
                    s2 = max(np.max(np.abs(dub)), np.max(np.abs(dvb)))
                    assert np.max(np.abs(dub + dvb)) < 1e-11 * s2
                    assert np.max(np.abs(fdiv - (dub + dvb))) < 1e-11 * s2


def test_filtering_spectrum_slope_saturates_at_p_plus_2():
    # test_validation.jl:68-111: with p = 1 vanishing moments both kernels recover k^−α only for
    # α < 3 and saturate near k^−3 above (cumulative_energy + spectral_density, kℓ = 1/ℓ)
    N, dx = 256, 1000.0
    xs = cart_axes(N, dx)
    g = oracle_grid_cart(xs, xs, periodic=(True, True))
    scales = 10 ** np.linspace(np.log10(5 * dx), np.log10(64 * dx), 17)
    kl = 1.0 / scales
    w = (scales >= 8 * dx) & (scales <= 40 * dx)
    got = {}
    for alpha in (5 / 3, 4.0, 7.0):
        u, v = _spectral_field(N, dx, alpha)
        assert _fourier_slope(u, v, dx, 6, 32) == pytest.approx(-alpha, abs=0.05)
        for name, kind, a, eng in (("Gaussian", O.GAUSSIAN, 6.0, O.ENGINE_SEPARABLE),
                                   ("TopHat", O.TOPHAT, 0.0, O.ENGINE_PREFIX)):
            cum = O.cumulative_energy(g, u, v, scales, kind=kind, alpha=a, engine=eng)
            got[(name, alpha)] = _logfit(kl[w], O.spectral_density(cum, kl)[w])
    for name in ("Gaussian", "TopHat"):
        assert got[(name, 5 / 3)] == pytest.approx(-5 / 3, abs=0.20)
        assert -3.15 < got[(name, 7.0)] < -2.75
        assert -3.10 < got[(name, 4.0)] < -2.60
        assert got[(name, 7.0)] < got[(name, 4.0)] < got[(name, 5 / 3)]
    for alpha in (5 / 3, 4.0, 7.0):
        assert got[("Gaussian", alpha)] == pytest.approx(got[("TopHat", alpha)], abs=0.05)


def test_sfs_stress_symmetric_and_energy_budget():
    # test_diagnostics.jl:184-243, 283-344: τ_xy = τ_yx, tr τ ≥ 0 (Jensen), and the resolved plus
    # subfilter energies add up to the filtered total energy
    x = np.arange(0.0, 30001.0, 1000.0)
    g = oracle_grid_cart(x, x, periodic=(False, False))
    rng = np.random.Generator(np.random.PCG64(11))
    u, v = rng.random((31, 31)), rng.random((31, 31))
    ell = 5000.0
    ub, vb, uu, uv, vu, vv, tot = O.filter_fields(g, [u, v, u * u, u * v, v * u, v * v, u * u + v * v],
                                                  kind=O.TOPHAT, alpha=0.0, ell=ell)
    inner = (slice(9, 22), slice(9, 22))
    txx, txy, tyy = uu - ub * ub, uv - ub * vb, vv - vb * vb
    assert np.allclose(txy[inner], (vu - vb * ub)[inner], rtol=1e-10)
    assert np.all((txx + tyy)[inner] >= -1e-10)
    rho0 = 1025.0
    sfs = 0.5 * rho0 * (txx + tyy)
    res = 0.5 * rho0 * (ub ** 2 + vb ** 2)
    assert np.all(sfs[inner] >= -1e-12)
    assert np.allclose((sfs + res)[inner], (0.5 * rho0 * tot)[inner], rtol=1e-10)

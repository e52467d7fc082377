"""Seeded synthetic workloads (NumPy, FP64 on the host) for the benchmark configurations.

* cartesian_turbulence — 2-D doubly periodic box, streamfunction ψ̂(n) = |n|⁻³ e^{iθ} for
  4 ≤ |n| ≤ 85 (E(k) ∝ k⁻³), u = −∂yψ, v = ∂xψ, RMS speed 0.1 m/s.
* sphere_turbulence — 0.25° lon-lat cell centres, ψ = Σ_{l=8}^{400} Σ_m a_lm Y_lm with random
  phases and |a_lm|² ∝ l⁻³/(l(l+1)(2l+1)) (kinetic energy per degree ∝ l⁻³),
  u = −(1/R)∂φψ, v = (1/(R cosφ))∂λψ, area-weighted RMS speed 0.2 m/s.
* sphere_land_mask — smooth Gaussian random field (degree spectrum exp(−½l(l+1)(λc/R)²),
  λc = 500 km), land below its area-weighted 30th percentile.
* volume_turbulence — 3-D random-phase field with E(k) ∝ k^(−5/3) for 2 ≤ k ≤ 80, projected
  divergence-free, RMS 0.1 m/s.
Arrays come back in tensor order: (Ny, Nx) / (Nz, Ny, Nx).
"""
from __future__ import annotations

import numpy as np


def cartesian_turbulence(N=256, L=1.0e6, seed=0, kmin=4, kmax=85, rms=0.1):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.fft.fftfreq(N, 1.0 / N)
    nx, ny = np.meshgrid(n, n)                      # [j, i]
    kk = np.sqrt(nx ** 2 + ny ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[band] = kk[band] ** -3.0
    theta = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
    psi_hat = amp * np.exp(1j * theta)
    k0 = 2.0 * np.pi / L
    u = np.real(np.fft.ifft2(-1j * k0 * ny * psi_hat))
    v = np.real(np.fft.ifft2(1j * k0 * nx * psi_hat))
    s = rms / np.sqrt(np.mean(u ** 2 + v ** 2))
    x = np.arange(N) * (L / N)
    return u * s, v * s, x


def sphere_axes(Nx=1440, Ny=720):
    dlon = 2.0 * np.pi / Nx
    dlat = np.pi / Ny
    lam = (np.arange(Nx) + 0.5) * dlon
    phi = -np.pi / 2 + (np.arange(Ny) + 0.5) * dlat
    return lam, phi


def _legendre_synthesis(coef, lmax, mmax, x):
    """f_m(x) = Σ_l coef[l, m] P̄_l^m(x) for fully normalized (orthonormal on [−1,1] with ½∫…)
    associated Legendre functions, by the standard three-term recurrence in l for each m."""
    nx = len(x)
    out = np.zeros((mmax + 1, nx), dtype=coef.dtype)
    sx = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    pmm = np.ones(nx)
    for m in range(0, mmax + 1):
        if m > 0:
            pmm = pmm * sx * np.sqrt((2.0 * m + 1.0) / (2.0 * m))
        p0 = pmm * (1.0 if m == 0 else 1.0)
        # P̄_m^m carries √(2m+1)·… normalization via the recurrence start above (P̄_0^0 = 1)
        acc = coef[m, m] * p0 if m <= lmax else 0.0
        if m + 1 <= lmax:
            p1 = x * np.sqrt(2.0 * m + 3.0) * p0
            acc = acc + coef[m + 1, m] * p1
            pl2, pl1 = p0, p1
            for l in range(m + 2, lmax + 1):
                a = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                pl = a * (x * pl1 - b * pl2)
                acc = acc + coef[l, m] * pl
                pl2, pl1 = pl1, pl
        out[m] = acc
    return out


def _sphere_field(coef, lmax, lam, phi):
    """Σ_{l,m} Re(coef[l,m] e^{imλ}) P̄_l^m(sinφ) on the lon-lat grid → (Ny, Nx)."""
    fm = _legendre_synthesis(coef, lmax, lmax, np.sin(phi))   # (m, Ny) complex
    Nx = len(lam)
    m = np.arange(lmax + 1)
    # real part of Σ_m fm(φ) e^{imλ}
    ph = np.exp(1j * np.outer(m, lam))                        # (m, Nx)
    return np.real(fm.T @ ph)


def sphere_turbulence(Nx=1440, Ny=720, R=6.371e6, seed=1, lmin=8, lmax=400, rms=0.2):
    rng = np.random.Generator(np.random.PCG64(seed))
    lam, phi = sphere_axes(Nx, Ny)
    l = np.arange(lmax + 1, dtype=np.float64)
    amp = np.zeros(lmax + 1)
    sel = l >= lmin
    amp[sel] = np.sqrt(l[sel] ** -3.0 / (l[sel] * (l[sel] + 1.0) * (2.0 * l[sel] + 1.0)))
    coef = np.zeros((lmax + 1, lmax + 1), dtype=np.complex128)
    theta = rng.uniform(0.0, 2.0 * np.pi, size=(lmax + 1, lmax + 1))
    for m in range(lmax + 1):
        coef[m:, m] = amp[m:] * np.exp(1j * theta[m:, m])
    h = 1e-5
    psi_p = _sphere_field(coef, lmax, lam, phi + h)
    psi_m = _sphere_field(coef, lmax, lam, phi - h)
    dpsi_dphi = (psi_p - psi_m) / (2.0 * h)
    # ∂λψ exactly: multiply by i·m
    mm = np.arange(lmax + 1)
    fm = _legendre_synthesis(coef, lmax, lmax, np.sin(phi))
    ph = np.exp(1j * np.outer(mm, lam))
    dpsi_dlam = np.real((fm * (1j * mm)[:, None]).T @ ph)
    cphi = np.cos(phi)[:, None]
    u = -dpsi_dphi / R
    v = dpsi_dlam / (R * cphi)
    wts = np.broadcast_to(cphi, u.shape)
    s = rms / np.sqrt(np.sum((u ** 2 + v ** 2) * wts) / np.sum(wts))
    return u * s, v * s, lam, phi


def sphere_turbulence_batch(seeds, Nx=1440, Ny=720, R=6.371e6, lmin=8, lmax=400, rms=0.2):
    """sphere_turbulence for several seeds, sharing the Legendre recursion (one pass for all)."""
    seeds = list(seeds)
    ns = len(seeds)
    lam, phi = sphere_axes(Nx, Ny)
    l = np.arange(lmax + 1, dtype=np.float64)
    amp = np.zeros(lmax + 1)
    sel = l >= lmin
    amp[sel] = np.sqrt(l[sel] ** -3.0 / (l[sel] * (l[sel] + 1.0) * (2.0 * l[sel] + 1.0)))
    coef = np.zeros((ns, lmax + 1, lmax + 1), dtype=np.complex128)
    for q, sd in enumerate(seeds):
        rng = np.random.Generator(np.random.PCG64(sd))
        theta = rng.uniform(0.0, 2.0 * np.pi, size=(lmax + 1, lmax + 1))
        for m in range(lmax + 1):
            coef[q, m:, m] = amp[m:] * np.exp(1j * theta[m:, m])
    h = 1e-5
    xs = np.concatenate([np.sin(phi + h), np.sin(phi - h), np.sin(phi)])
    # Legendre synthesis for all seeds at the three latitude sets at once
    nx = len(xs)
    fm = np.zeros((ns, lmax + 1, nx), dtype=np.complex128)
    sx = np.sqrt(np.maximum(0.0, 1.0 - xs * xs))
    pmm = np.ones(nx)
    for m in range(lmax + 1):
        if m > 0:
            pmm = pmm * sx * np.sqrt((2.0 * m + 1.0) / (2.0 * m))
        acc = coef[:, m, m][:, None] * pmm[None, :]
        if m + 1 <= lmax:
            p1 = xs * np.sqrt(2.0 * m + 3.0) * pmm
            acc = acc + coef[:, m + 1, m][:, None] * p1[None, :]
            pl2, pl1 = pmm, p1
            for ll in range(m + 2, lmax + 1):
                a = np.sqrt((4.0 * ll * ll - 1.0) / (ll * ll - m * m))
                b = np.sqrt(((ll - 1.0) ** 2 - m * m) / (4.0 * (ll - 1.0) ** 2 - 1.0))
                pl = a * (xs * pl1 - b * pl2)
                acc = acc + coef[:, ll, m][:, None] * pl[None, :]
                pl2, pl1 = pl1, pl
        fm[:, m, :] = acc
    mm = np.arange(lmax + 1)
    ph = np.exp(1j * np.outer(mm, lam))
    cphi = np.cos(phi)[:, None]
    out = []
    for q in range(ns):
        f = fm[q]
        psi_p = np.real(f[:, :Ny].T @ ph)
        psi_m = np.real(f[:, Ny:2 * Ny].T @ ph)
        dpsi_dphi = (psi_p - psi_m) / (2.0 * h)
        dpsi_dlam = np.real((f[:, 2 * Ny:] * (1j * mm)[:, None]).T @ ph)
        u = -dpsi_dphi / R
        v = dpsi_dlam / (R * cphi)
        wts = np.broadcast_to(cphi, u.shape)
        s = rms / np.sqrt(np.sum((u ** 2 + v ** 2) * wts) / np.sum(wts))
        out.append((u * s, v * s))
    return [o[0] for o in out], [o[1] for o in out], lam, phi


def sphere_land_mask(Nx=1440, Ny=720, R=6.371e6, seed=2, corr=500e3, land_fraction=0.3, lmax=100):
    rng = np.random.Generator(np.random.PCG64(seed))
    lam, phi = sphere_axes(Nx, Ny)
    l = np.arange(lmax + 1, dtype=np.float64)
    amp = np.sqrt(np.exp(-0.5 * l * (l + 1.0) * (corr / R) ** 2))
    amp[0] = 0.0
    coef = (rng.standard_normal((lmax + 1, lmax + 1)) + 1j * rng.standard_normal((lmax + 1, lmax + 1)))
    coef = np.tril(coef) * amp[:, None]
    f = _sphere_field(coef, lmax, lam, phi)
    wts = np.broadcast_to(np.cos(phi)[:, None], f.shape).ravel()
    order = np.argsort(f.ravel())
    cw = np.cumsum(wts[order])
    thr = f.ravel()[order][np.searchsorted(cw, land_fraction * cw[-1])]
    return f >= thr   # active (ocean) cells


def stretched_levels(Nz=64, ratio=1.03, dz_mid=4.0, kmid=31):
    """z_k with Δz_k = Δz₁·ratio^(k−1), Δz₁ chosen so Δz at level kmid is dz_mid."""
    dz1 = dz_mid / ratio ** kmid
    dz = dz1 * ratio ** np.arange(Nz)
    z = np.concatenate([[0.0], np.cumsum(dz[:-1])])
    return z


def volume_turbulence(Nx=256, Ny=256, Nz=64, seed=3, kmin=2, kmax=80, rms=0.1):
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)[None, None, :]
    ky = np.fft.fftfreq(Ny, 1.0 / Ny)[None, :, None]
    kz = np.fft.fftfreq(Nz, 1.0 / Nz)[:, None, None]
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros(kk.shape)
    amp[band] = np.sqrt(kk[band] ** (-5.0 / 3.0) / (4.0 * np.pi * kk[band] ** 2))
    comps = []
    for _ in range(3):
        comps.append(amp * np.exp(1j * rng.uniform(0, 2 * np.pi, size=kk.shape)))
    kvec = (np.broadcast_to(kx, kk.shape), np.broadcast_to(ky, kk.shape), np.broadcast_to(kz, kk.shape))
    k2 = np.where(kk > 0, kk ** 2, 1.0)
    div = sum(kvec[a] * comps[a] for a in range(3)) / k2
    comps = [comps[a] - kvec[a] * div for a in range(3)]
    vel = [np.real(np.fft.ifftn(c)) for c in comps]
    s = rms / np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    return [x * s for x in vel]

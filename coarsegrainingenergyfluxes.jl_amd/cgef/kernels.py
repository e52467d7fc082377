"""Filter kernels (Kernels.jl): the kernel types the plans are built from, and the scalar helpers
of the reference's public Kernels API. The footprint tables themselves are built in C++ inside
libcgef_hip.so (cgef_plan_create); these scalar functions are what a caller of the reference uses
to reason about a kernel (radius, transfer function, spectrum admissibility)."""
from __future__ import annotations

import math
from dataclasses import dataclass

from ._lib import ArgumentError, GAUSSIAN, TOPHAT

GAUSSIAN_TRUNCATION_TOL = 1e-10  # Kernels.jl:307


class AbstractFilterKernel:
    kind: int
    alpha: float = 0.0


@dataclass(frozen=True)
class TopHatKernel(AbstractFilterKernel):
    """Kernels.jl:28 — unit weight for d ≤ ℓ/2."""

    @property
    def kind(self):
        return TOPHAT


@dataclass(frozen=True)
class GaussianKernel(AbstractFilterKernel):
    """Kernels.jl:59-62 — exp(−α (d/ℓ)²); α = 6 (default) or 4 (FlowSieve's convention)."""

    alpha: float = 6.0

    @property
    def kind(self):
        return GAUSSIAN


def kernel_weight(kernel: AbstractFilterKernel, d: float, ell: float) -> float:
    """Kernels.jl:318-324 — unnormalized weight."""
    if isinstance(kernel, TopHatKernel):
        return 1.0 if d <= ell / 2.0 else 0.0
    r = d / ell
    return math.exp(-kernel.alpha * (r * r))


def kernel_radius(kernel: AbstractFilterKernel, ell: float) -> float:
    """Kernels.jl:354-359 — ℓ/2 for the top-hat, ℓ√(−ln 1e-10/α) for the Gaussian."""
    if isinstance(kernel, TopHatKernel):
        return ell / 2.0
    return ell * math.sqrt(-math.log(GAUSSIAN_TRUNCATION_TOL) / kernel.alpha)


def spectral_transfer(kernel: AbstractFilterKernel, kmag: float, ell: float) -> float:
    """Kernels.jl:405-406 — Gaussian Ĝ(k) = exp(−k²ℓ²/(4α)). The top-hat's planar transfer needs
    a Bessel J₁ (the reference's SpecialFunctions extension) and is not provided here."""
    if isinstance(kernel, GaussianKernel):
        return math.exp(-kmag ** 2 * ell ** 2 / (4.0 * kernel.alpha))
    raise ArgumentError("the top-hat's planar transfer function (2J₁(kR)/(kR)) is not provided by this backend")


def transfer_monotone(kernel: AbstractFilterKernel) -> bool:
    """Kernels.jl:509-518 — whether |Ĝ(k)|² is monotone decreasing."""
    if isinstance(kernel, GaussianKernel):
        return True
    if isinstance(kernel, TopHatKernel):
        return False
    raise TypeError(f"no transfer_monotone method for {type(kernel).__name__}")


def check_spectrum_kernel(kernel: AbstractFilterKernel) -> None:
    """Kernels.jl:527-543 — ArgumentError unless transfer_monotone(kernel)."""
    if transfer_monotone(kernel):
        return
    raise ArgumentError(
        f"{type(kernel).__name__} cannot produce a guaranteed non-negative filtering spectral density: its "
        "|Ĝ(k)|² is not monotone decreasing, so Sadek & Aluie (2018) eq. (21) does not hold. Use "
        "spectrum=ForceSpectrum() to compute it anyway, spectrum=NoSpectrum() to skip it, a GaussianKernel, "
        "or cumulative_energy, which is defined for any kernel.")

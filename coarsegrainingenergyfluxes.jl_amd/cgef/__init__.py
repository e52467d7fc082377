"""cgef — MI355X (gfx950) backend for CoarseGrainingEnergyFluxes.jl's real-space footprint filter
and Π pipeline.

Host-side mirror of the reference's API (Julia `name!` → Python `name_`): plan_filter,
filter_apply_, filter_field_, ddx_/ddy_/ddz_, compute_pi_ (compute_Π!), cumulative_energy(_),
filtering_spectrum, spectral_density(_), coarse_grain(_), coarse_grain_batch_, coarse_grain_profile(_),
coarse_grain_slices_, filter_slices_, with the RealSpace / Spectral / AutoMethod filter methods. Every compute call
runs in libcgef_hip.so on the GPU holding the (torch) tensors; there is no CPU fallback.
"""
from ._lib import ArgumentError, BackendUnavailable, DeviceError, DimensionMismatch, version
from .derivatives import StencilPlan, ddx_, ddy_, ddz_
from .diagnostics import (ForceSpectrum, NoSpectrum, PiWorkspace, StrictSpectrum, active_area, compute_pi_,
                          compute_Π_, cumulative_energy, cumulative_energy_, energy_from_filtered,
                          filtering_spectrum, gate_spectrum, spectral_density, spectral_density_)
from .filtering import (AutoMethod, Deformable, FilterPlan, GPUBackend, RealSpace, Spectral, ZeroFill, filter_apply_,
                        filter_apply_batch_, filter_field_, filter_slices_, plan_filter, slice_costs)
from .grids import CartesianGeometry, Range, SphericalGeometry, StructuredGrid
from .kernels import (GaussianKernel, TopHatKernel, check_spectrum_kernel, kernel_radius, kernel_weight,
                      spectral_transfer, transfer_monotone)
from .pipeline import (CoarseGrainBatchResult, CoarseGrainResult, coarse_grain, coarse_grain_, coarse_grain_batch_,
                       coarse_grain_profile, coarse_grain_profile_, coarse_grain_slices_, slice_pipeline_costs)

from . import closure  # noqa: E402  (learned τ closure: device training and inference, exact-Π fallback)

ΠWorkspace = PiWorkspace

__all__ = [n for n in dir() if not n.startswith("_")]

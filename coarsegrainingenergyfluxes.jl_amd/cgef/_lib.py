"""ctypes binding of libcgef_hip.so (declared in include/cgef.h).

The library is the product path: nothing here falls back to a CPU implementation. If the shared
library is missing or cannot be loaded, every entry point raises ``BackendUnavailable``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcgef_hip.so")

CGEF_OK, CGEF_ERR_ARGUMENT, CGEF_ERR_DIMENSION, CGEF_ERR_DEVICE, CGEF_ERR_ALLOC, CGEF_ERR_UNSUPPORTED = range(6)
F32, F64 = 0, 1
CARTESIAN, SPHERICAL = 0, 1
TOPHAT, GAUSSIAN = 0, 1
ZEROFILL, DEFORMABLE = 0, 1
ENGINE_AUTO, ENGINE_BANDED, ENGINE_SEPARABLE, ENGINE_BASELINE, ENGINE_RING, ENGINE_PREFIX, ENGINE_SPECTRAL = range(7)

EXPORTED = (
    "cgef_version", "cgef_last_error", "cgef_grid_create", "cgef_grid_destroy", "cgef_active_area",
    "cgef_plan_create", "cgef_plan_destroy", "cgef_plan_query", "cgef_workspace_create",
    "cgef_workspace_destroy", "cgef_filter_apply", "cgef_compute_pi", "cgef_coarse_grain", "cgef_derivative",
    "cgef_energy", "cgef_energy_noalloc",
    "cgef_cumulative_energy", "cgef_profile_enable", "cgef_profile_read", "cgef_plan_mode_cut", "cgef_spectral_density",
    # learned closure (no reference counterpart)
    "cgef_conv5", "cgef_conv5_mfma", "cgef_conv5_pack_mfma", "cgef_conv5_mfma_ex", "cgef_conv5_pack_mfma_ex",
    "cgef_conv5_wgrad_scratch", "cgef_conv5_wgrad", "cgef_conv5_pack_weights", "cgef_adam_update",
    "cgef_mse_grad", "cgef_closure_loss_grad", "cgef_closure_crop", "cgef_closure_products", "cgef_closure_targets", "cgef_closure_pi",
    "cgef_closure_gradient_model",
)


class BackendUnavailable(RuntimeError):
    """The HIP backend library is not built or cannot be loaded."""


class ArgumentError(ValueError):
    """The reference's ArgumentError."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class DeviceError(RuntimeError):
    """A HIP runtime call failed."""


class GridDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("geometry", C.c_int32),
        ("n", C.c_int64 * 3),
        ("coord", C.POINTER(C.c_double) * 3),
        ("uniform", C.c_int32 * 3),
        ("periodic", C.c_int32 * 3),
        ("period", C.c_double * 3),
        ("radius", C.c_double),
        ("mask", C.POINTER(C.c_uint8)),
    ]


class KernelDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("alpha", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [
        ("engine", C.c_int32),
        ("nbands", C.c_int32),
        ("nentries", C.c_int64),
        ("nruns", C.c_int64),
        ("mean_taps", C.c_double),
        ("max_taps", C.c_int64),
        ("radius", C.c_double),
        ("active_taps", C.c_double),
        ("ring_offsets", C.c_int64),
        ("ring_modes", C.c_int64),
        ("ring_pi_fields", C.c_int64),
        ("ring_contract_kernel", C.c_int64),
        ("ring_modes_kept", C.c_double),
    ]


_lib = None


def load():
    """Load libcgef_hip.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendUnavailable(
            f"{LIB_PATH} is not built; run coarsegrainingenergyfluxes.jl_amd/build.sh (or __graft_entry__.build())")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the host
        raise BackendUnavailable(f"cannot load {LIB_PATH}: {e}") from e
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    dp = C.POINTER(C.c_double)
    L.cgef_version.restype = C.c_char_p
    L.cgef_last_error.restype = C.c_char_p
    sigs = {
        "cgef_grid_create": [C.POINTER(GridDesc), i32, C.POINTER(vp)],
        "cgef_grid_destroy": [vp],
        "cgef_active_area": [vp, dp],
        "cgef_plan_create": [vp, C.POINTER(KernelDesc), dbl, i32, i32, C.POINTER(vp)],
        "cgef_plan_destroy": [vp],
        "cgef_plan_query": [vp, C.POINTER(PlanInfo)],
        "cgef_workspace_create": [vp, i64, i32, C.POINTER(vp)],
        "cgef_workspace_destroy": [vp],
        "cgef_filter_apply": [vp, vp, vp, i64, i64, i32, vp],
        "cgef_compute_pi": [vp, vp, vp, vp, vp, vp, i64, i32, dbl, dp, vp, vp, vp, vp],
        "cgef_coarse_grain": [C.POINTER(vp), i64, vp, vp, vp, vp, vp, i64, i32, dbl, dp, vp],
        "cgef_derivative": [vp, i32, vp, vp, i64, i32, vp],
        "cgef_energy": [vp, vp, vp, vp, i64, i32, dp, vp],
        "cgef_energy_noalloc": [vp, vp, vp, vp, i64, i32, dp, vp],
        "cgef_cumulative_energy": [C.POINTER(vp), i64, vp, vp, vp, vp, i64, i32, dp, vp],
        "cgef_profile_enable": [vp, i32],
        "cgef_profile_read": [vp, dp, C.POINTER(i64)],
        "cgef_plan_mode_cut": [vp, C.POINTER(C.c_int32), i64],
        "cgef_spectral_density": [i64, dp, dp, dp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(status: int):
    if status == CGEF_OK:
        return
    msg = load().cgef_last_error().decode(errors="replace")
    if status == CGEF_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if status == CGEF_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if status == CGEF_ERR_UNSUPPORTED:
        raise ArgumentError("unsupported: " + msg)
    if status == CGEF_ERR_ALLOC:
        raise MemoryError(msg)
    raise DeviceError(msg)


def version() -> str:
    return load().cgef_version().decode()

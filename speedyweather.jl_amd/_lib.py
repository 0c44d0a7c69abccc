"""ctypes binding of libsht.so (the C ABI declared in include/sht.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libsht.so")

SHT_OK = 0
SHT_ERR_ARG = 1
SHT_ERR_DIM = 2
SHT_ERR_TYPE = 3
SHT_ERR_HIP = 4
SHT_ERR_UNSUPPORTED = 5
SHT_ERR_ALLOC = 6

GRIDS = {"OctahedralGaussianGrid": 0, "FullGaussianGrid": 1}
# LegendreShortcut by short_name (legendre_shortcuts.jl:21-49); "linquadcoslat2" spells the ²
SHORTCUTS = {"linear": 0, "quadratic": 1, "cubic": 2, "linquadcoslat2": 3, "linquadcoslat²": 3, "lincubcoslat": 4}
PRECISIONS = {"fp32": 0, "fp16x3": 1, "bf16x3": 2}

# every symbol include/sht.h declares
EXPORTS = (
    "sht_plan_create", "sht_plan_destroy", "sht_plan_sizes", "sht_synthesis", "sht_analysis",
    "sht_uv_from_vordiv", "sht_gauss_legendre", "sht_legendre_column", "sht_last_error",
    "sht_version", "sht_profile_begin", "sht_profile_end", "sht_batch_create", "sht_batch_launch",
    "sht_batch_uses_graph", "sht_batch_destroy", "sht_mlp_nparams", "sht_mlp_create", "sht_mlp_set_params",
    "sht_mlp_get_params", "sht_mlp_forward", "sht_mlp_gradient", "sht_mlp_train_step", "sht_mlp_destroy",
    "sht_plan_autotune", "sht_plan_tuning", "sht_plan_set_tuning", "sht_spec2grid_pullback",
    "sht_grid2spec_pullback", "sht_divergence", "sht_laplace", "sht_gradient", "sht_plan_create_matrix",
    "sht_synthesis_laplace", "sht_analysis_laplace", "sht_time_stage", "sht_tune_cache_save",
    "sht_tune_cache_load", "sht_transform_static", "sht_plan_set_graphs", "sht_plan_graph_count",
)


class DimensionMismatch(ValueError):
    """Raised where the reference throws DimensionMismatch(S, coeffs|field)
    (SpeedyTransforms/src/spectral_transform.jl:387-423)."""


class SHTError(RuntimeError):
    pass


class ShtDesc(C.Structure):
    _fields_ = [("grid", C.c_int32), ("nlat_half", C.c_int32), ("lmax", C.c_int32),
                ("mmax", C.c_int32), ("nlayers", C.c_int32), ("shortcut", C.c_int32),
                ("precision", C.c_int32), ("device", C.c_int32)]


class ShtBatchItem(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("direction", C.c_int32), ("K", C.c_int32),
                ("unscale_coslat", C.c_int32), ("reserved", C.c_int32),
                ("in_", C.c_void_p), ("ld_in", C.c_int64), ("mid", C.c_void_p), ("ld_mid", C.c_int64),
                ("out", C.c_void_p), ("ld_out", C.c_int64)]


SHT_DIR_SYNTHESIS, SHT_DIR_ANALYSIS, SHT_DIR_ROUNDTRIP = 0, 1, 2
SHT_BATCH_NO_GRAPH, SHT_BATCH_FORCE_GRAPH_FAILURE = 1, 2


class ShtSizes(C.Structure):
    _fields_ = [("npoints", C.c_int64), ("nharm", C.c_int64), ("nlat", C.c_int32),
                ("nlon_max", C.c_int32), ("nfreq_total", C.c_int64), ("device_bytes", C.c_int64)]


_LIB = None


def lib() -> C.CDLL:
    """Load libsht.so. There is no fallback: a missing library is an error."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.sht_plan_create.argtypes = [C.POINTER(ShtDesc), C.POINTER(vp)]
        L.sht_plan_create_matrix.argtypes = [C.POINTER(ShtDesc), C.POINTER(vp)]
        L.sht_plan_destroy.argtypes = [vp]
        L.sht_plan_sizes.argtypes = [C.POINTER(ShtDesc), C.POINTER(ShtSizes)]
        L.sht_synthesis.argtypes = [vp, vp, i64, vp, i64, i32, i32, vp]
        L.sht_analysis.argtypes = [vp, vp, i64, vp, i64, i32, vp]
        L.sht_transform_static.argtypes = [vp, i32, vp, i64, vp, i64, i32, i32, vp]
        L.sht_plan_set_graphs.argtypes = [vp, i32, i32]
        L.sht_plan_graph_count.argtypes = [vp, vp, vp]
        L.sht_spec2grid_pullback.argtypes = [vp, vp, i64, vp, i64, i32, i32, i32, vp]
        L.sht_grid2spec_pullback.argtypes = [vp, vp, i64, vp, i64, i32, i32, vp]
        L.sht_divergence.argtypes = [vp, vp, vp, vp, i64, i32, i32, i32, i32, C.c_double, vp]
        L.sht_laplace.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, C.c_double, vp]
        L.sht_synthesis_laplace.argtypes = [vp, vp, i64, vp, i64, i32, i32, i32, C.c_double, i32, vp]
        L.sht_analysis_laplace.argtypes = [vp, vp, i64, vp, i64, i32, i32, i32, i32, C.c_double, vp]
        L.sht_gradient.argtypes = [vp, vp, vp, vp, i64, i32, C.c_double, vp]
        L.sht_uv_from_vordiv.argtypes = [vp, vp, vp, vp, vp, i64, i32, C.c_double, vp]
        L.sht_gauss_legendre.argtypes = [i32, vp, vp]
        L.sht_legendre_column.argtypes = [i32, i32, C.c_double, vp]
        L.sht_batch_create.argtypes = [C.POINTER(ShtBatchItem), i32, i32, C.POINTER(vp)]
        L.sht_batch_launch.argtypes = [vp, vp]
        L.sht_batch_uses_graph.argtypes = [vp]
        L.sht_batch_destroy.argtypes = [vp]
        L.sht_mlp_create.argtypes = [i32, i32, C.POINTER(vp)]
        L.sht_mlp_set_params.argtypes = [vp, vp]
        L.sht_mlp_get_params.argtypes = [vp, vp]
        L.sht_mlp_forward.argtypes = [vp, vp, vp, i32, vp]
        L.sht_mlp_gradient.argtypes = [vp, vp, vp, i32, vp, vp, vp]
        L.sht_mlp_train_step.argtypes = [vp, vp, vp, i32, C.c_float, vp, vp]
        L.sht_mlp_destroy.argtypes = [vp]
        L.sht_plan_autotune.argtypes = [vp, i32, vp]
        L.sht_plan_tuning.argtypes = [vp, vp]
        L.sht_plan_set_tuning.argtypes = [vp, vp]
        L.sht_profile_begin.argtypes = [vp, i32]
        L.sht_profile_end.argtypes = [vp, vp, vp]
        L.sht_time_stage.argtypes = [vp, i32, vp, i64, vp, i64, i32, i32, vp, vp]
        L.sht_tune_cache_save.argtypes = [C.c_char_p]
        L.sht_tune_cache_load.argtypes = [C.c_char_p, vp]
        L.sht_last_error.restype = C.c_char_p
        L.sht_version.restype = C.c_char_p
        for name in EXPORTS:
            getattr(L, name)
        _LIB = L
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc == SHT_OK:
        return
    msg = lib().sht_last_error().decode(errors="replace")
    if rc == SHT_ERR_DIM:
        raise DimensionMismatch(msg)
    if rc == SHT_ERR_TYPE:
        raise TypeError(msg)
    raise SHTError(f"{what}: libsht error {rc}: {msg}")

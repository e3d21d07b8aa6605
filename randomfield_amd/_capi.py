"""ctypes binding of librf_hip.so (include/rf_capi.h) and of the Fortran
facade's draw helpers (librandomfield_gpu.so).

The GPU path has no CPU fallback: if the native library is missing the import
of the binding raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIBDIR = Path(__file__).resolve().parent / "lib"

RF_OK = 0
RF_E_ARG, RF_E_LOGNORMAL, RF_E_NOMEM, RF_E_HIP, RF_E_UNSUPPORTED, RF_E_NODEVICE = -1, -2, -3, -4, -5, -6
ISOTROPIC, SHINOZUKA, RANDOMIZATION, FFT = 1, 2, 3, 4
CM_GAUSSIAN, CM_COS = 1, 2
FOM_NONE, FOM_GAUSSIAN, FOM_LOGNORMAL = 0, 1, 2
REF_COMPAT, FIXED = 0, 1
FORM_FP64, FORM_FP32, FORM_MFMA_F16, FORM_MFMA_BF16 = 0, 1, 2, 3
KSAMP_BOX, KSAMP_DENSITY = 0, 1
DTYPE_F64, DTYPE_F32 = 0, 1
MFMA_VARIANTS = 9
DRAW_NEXT = -1

# every entry point the header declares
EXPORTS = (
    "rf_context_create", "rf_context_destroy", "rf_last_error", "rf_version", "rf_context_sync",
    "rf_plan_spectrum", "rf_spectrum_tables", "rf_plan_create", "rf_plan_destroy", "rf_plan_info",
    "rf_plan_execute", "rf_plan_set_variant", "rf_plan_autotune", "rf_plan_execute_grid", "rf_plan_redraw", "rf_generate", "rf_generate_grid", "rf_marginal_device", "rf_marginal_host", "rf_blend_device",
    "rf_stats_device", "rf_corr_device", "rf_interpolate_device", "rf_interpolate_box_device", "rf_unv_read", "rf_unv_sizes", "rf_unv_copy",
    "rf_unv_destroy", "rf_normalize_device", "rf_device_uniforms", "rf_uniform_host",
    "rf_philox4x32_10_host",
    "rf_server_create", "rf_server_destroy", "rf_server_submit", "rf_server_wait", "rf_server_read",
    "rf_server_release", "rf_server_get_stats",
    "rf_blend_plan_create", "rf_blend_plan_execute", "rf_blend_plan_destroy",
    "rf_grid_batch_create", "rf_grid_batch_execute", "rf_grid_batch_destroy", "rf_grid_batch_redraw",
    "rf_grid_batch_autotune", "rf_grid_batch_fuse_blend",
    "rf_graph_capture_begin", "rf_graph_capture_end", "rf_graph_launch", "rf_graph_destroy",
    "rf_localization_layout", "rf_unity_partition", "rf_blend_host", "rf_normalize_window_device",
    "rf_normalize_window_host", "rf_normalize_device_noalloc",
)


class Params(C.Structure):
    _fields_ = [
        ("nDim", C.c_int32), ("method", C.c_int32), ("corrMod", C.c_int32), ("mode", C.c_int32),
        ("form", C.c_int32), ("kSampling", C.c_int32), ("Nmc", C.c_int32), ("margiFirst", C.c_int32),
        ("nKDensity", C.c_int64), ("xSlab", C.c_int64),
        ("corrL", C.c_double * 3), ("xRange", C.c_double * 3),
        ("fieldAvg", C.c_double), ("fieldVar", C.c_double),
        ("seed", C.c_uint64), ("isoShared", C.c_int32), ("gridN", C.c_int32 * 3), ("reserved", C.c_int32 * 4),
    ]


class SpectrumInfo(C.Structure):
    _fields_ = [
        ("kNTotal", C.c_int64), ("kN_unsigned", C.c_int64), ("nKUniform", C.c_int64),
        ("nPhaseUniform", C.c_int64), ("kNStep", C.c_int32 * 3), ("nSignBlocks", C.c_int32),
        ("nVarIso", C.c_int32), ("pad", C.c_int32), ("kMax", C.c_double * 3),
        ("kDelta", C.c_double * 3), ("ampMult", C.c_double), ("isoStep", C.c_double),
    ]


class LayoutParams(C.Structure):
    _fields_ = [
        ("nDim", C.c_int32), ("nFields", C.c_int32 * 3), ("pointsPerCorrL", C.c_int32 * 3), ("pad", C.c_int32),
        ("xMinGlob", C.c_double * 3), ("xMaxGlob", C.c_double * 3), ("corrL", C.c_double * 3),
        ("overlap", C.c_double * 3),
    ]


class LayoutInfo(C.Structure):
    _fields_ = [
        ("nDim", C.c_int32), ("nSub", C.c_int32), ("groupN", C.c_int32 * 3), ("subN", C.c_int32 * 3),
        ("nOvl", C.c_int32 * 3), ("nOvlRaw", C.c_int32 * 3), ("xStep", C.c_double * 3),
        ("overlap", C.c_double * 3), ("xMinGlob", C.c_double * 3), ("xMaxGlob", C.c_double * 3),
        ("procExtent", C.c_double * 3), ("stepProc", C.c_double * 3), ("xRangeSub", C.c_double * 3),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbmBudget", C.c_int64), ("maxInflight", C.c_int32), ("maxBatch", C.c_int32),
        ("form", C.c_int32), ("maxNmc", C.c_int32), ("xRangeNorm", C.c_double * 3),
        ("batchTerms", C.c_double), ("reserved", C.c_int32 * 8),
    ]


class Request(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("Nmc", C.c_int32), ("margiFirst", C.c_int32), ("xOnDevice", C.c_int32),
        ("N", C.c_int64), ("x", C.c_void_p), ("corrL", C.c_double * 3),
        ("fieldAvg", C.c_double), ("fieldVar", C.c_double), ("seed", C.c_uint64),
        ("reserved", C.c_int32 * 6),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("Nmc", C.c_int32), ("N", C.c_int64), ("d_out", C.c_void_p),
        ("tSubmit", C.c_double), ("tLaunch", C.c_double), ("tDone", C.c_double), ("batch", C.c_int64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("batches", C.c_int64),
        ("arenaCapacity", C.c_int64), ("arenaUsed", C.c_int64), ("arenaPeak", C.c_int64),
        ("maxBatchSeen", C.c_int32), ("maxInflightSeen", C.c_int32), ("pointRealizations", C.c_double),
    ]


def make_params(nDim=3, method=SHINOZUKA, corrMod=CM_GAUSSIAN, mode=REF_COMPAT, form=FORM_FP64,
                kSampling=KSAMP_BOX, Nmc=1, margiFirst=FOM_NONE, nKDensity=0, xSlab=0,
                corrL=(1.0, 1.0, 1.0), xRange=(1.0, 1.0, 1.0), fieldAvg=0.0, fieldVar=1.0,
                seed=0, isoShared=0, gridN=(0, 0, 0)) -> Params:
    p = Params()
    p.nDim, p.method, p.corrMod, p.mode, p.form = nDim, method, corrMod, mode, form
    p.kSampling, p.Nmc, p.margiFirst, p.nKDensity, p.xSlab = kSampling, Nmc, margiFirst, nKDensity, xSlab
    cl = list(corrL) + [1.0] * (3 - len(corrL))
    xr = list(xRange) + [1.0] * (3 - len(xRange))
    p.corrL = (C.c_double * 3)(*cl[:3])
    p.xRange = (C.c_double * 3)(*xr[:3])
    p.fieldAvg, p.fieldVar, p.seed, p.isoShared = fieldAvg, fieldVar, seed, isoShared
    gn = list(gridN) + [0] * (3 - len(gridN))
    p.gridN = (C.c_int32 * 3)(*gn[:3])
    return p


_lib = None
_flib = None


def _sig(L, name, argtypes, restype=C.c_int):
    """Declares a foreign function's signature.  A symbol missing from the
    library (a binding newer than the built .so) is left undeclared, so the
    rest of the library stays usable and calling it raises AttributeError."""
    f = getattr(L, name, None)
    if f is None:
        return
    if argtypes is not None:
        f.argtypes = argtypes
    f.restype = restype


def lib() -> C.CDLL:
    """librf_hip.so; raises when it is not built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = LIBDIR / "librf_hip.so"
    if not path.exists():
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() (the GPU path has no fallback)")
    L = C.CDLL(str(path))
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    _sig(L, "rf_context_create", [C.POINTER(vp), C.c_int])
    _sig(L, "rf_context_destroy", [vp])
    _sig(L, "rf_last_error", [vp])
    _sig(L, "rf_last_error", None, C.c_char_p)
    _sig(L, "rf_context_sync", [vp])
    _sig(L, "rf_plan_spectrum", [C.POINTER(Params), C.POINTER(SpectrumInfo)])
    _sig(L, "rf_spectrum_tables", [C.POINTER(Params), vp, vp, vp])
    _sig(L, "rf_plan_create", [vp, C.POINTER(Params), vp, vp, vp, C.POINTER(vp)])
    _sig(L, "rf_plan_destroy", [vp])
    _sig(L, "rf_plan_info", [vp, C.POINTER(SpectrumInfo)])
    _sig(L, "rf_plan_execute", [vp, vp, i64, vp, i32, vp])
    _sig(L, "rf_plan_set_variant", [vp, i32])
    _sig(L, "rf_plan_autotune", [vp, vp, i64, vp, i32, vp, vp])
    _sig(L, "rf_plan_execute_grid", [vp, vp, vp, vp, vp, i32, vp])
    _sig(L, "rf_plan_redraw", [vp, i64, vp])
    _sig(L, "rf_generate", [vp, C.POINTER(Params), vp, i64, vp, vp, vp, vp, i32])
    _sig(L, "rf_marginal_device", [vp, i32, C.c_double, C.c_double, vp, i64, i32, vp])
    _sig(L, "rf_marginal_host", [vp, i32, C.c_double, C.c_double, vp, i64])
    _sig(L, "rf_blend_device", [vp, i32, vp, i32, i32, vp, vp, vp, vp, i32, vp, vp, i32, vp])
    _sig(L, "rf_stats_device", [vp, vp, i64, i32, i32, vp, vp, vp])
    _sig(L, "rf_normalize_device", [vp, vp, i64, i32, i32, vp])
    _sig(L, "rf_corr_device", [vp, vp, i32, i32, vp, vp, i32, C.c_double, C.c_double, vp, vp])
    _sig(L, "rf_interpolate_device", [vp, i32, vp, vp, vp, i32, vp, i32, vp, i64, vp, i32, vp, vp])
    _sig(L, "rf_generate_grid", [vp, vp, vp, vp, vp, vp, vp, vp, i32])
    _sig(L, "rf_interpolate_box_device", [vp, i32, vp, vp, vp, vp, i32, vp, i32, vp, i64, vp, i32, vp, vp])
    _sig(L, "rf_unv_read", [C.c_char_p, i32, vp])
    _sig(L, "rf_unv_sizes", [vp, vp, vp, vp, vp])
    _sig(L, "rf_unv_copy", [vp, vp, vp, vp])
    _sig(L, "rf_unv_destroy", [vp])
    _sig(L, "rf_device_uniforms", [vp, C.c_uint64, C.c_uint32, i64, vp, vp])
    _sig(L, "rf_uniform_host", [C.c_uint64, C.c_uint32, C.c_uint64])
    _sig(L, "rf_uniform_host", None, C.c_double)
    _sig(L, "rf_philox4x32_10_host", [vp, vp, vp])
    _sig(L, "rf_blend_plan_create", [vp, i32, vp, i32, i32, vp, vp, vp, vp, i32, vp, i32, C.POINTER(vp)])
    _sig(L, "rf_blend_plan_execute", [vp, vp, vp])
    _sig(L, "rf_blend_plan_destroy", [vp])
    _sig(L, "rf_grid_batch_create", [vp, i32, vp, vp, vp, vp, vp, i32, C.POINTER(vp)])
    _sig(L, "rf_grid_batch_execute", [vp, vp])
    _sig(L, "rf_grid_batch_destroy", [vp])
    _sig(L, "rf_grid_batch_redraw", [vp, i64, vp])
    _sig(L, "rf_grid_batch_autotune", [vp, vp, vp])
    _sig(L, "rf_grid_batch_fuse_blend", [vp, vp, vp])
    _sig(L, "rf_graph_capture_begin", [vp, vp])
    _sig(L, "rf_graph_capture_end", [vp, vp, C.POINTER(vp)])
    _sig(L, "rf_graph_launch", [vp, vp])
    _sig(L, "rf_graph_destroy", [vp])
    _sig(L, "rf_localization_layout", [C.POINTER(LayoutParams), C.POINTER(LayoutInfo), vp, vp, vp, vp])
    _sig(L, "rf_unity_partition", [i32, vp, vp, i32, i32, vp])
    _sig(L, "rf_blend_host", [vp, i32, vp, i32, i32, vp, vp, vp, vp, i32, vp, vp])
    _sig(L, "rf_normalize_window_device", [vp, vp, i32, vp, vp, vp, i32, i32, vp])
    _sig(L, "rf_normalize_window_host", [vp, vp, i32, vp, vp, vp, i32])
    _sig(L, "rf_normalize_device_noalloc", [vp, vp, i64, i32, i32, vp])
    _sig(L, "rf_server_create", [vp, C.POINTER(ServerConfig), C.POINTER(vp)])
    _sig(L, "rf_server_destroy", [vp])
    _sig(L, "rf_server_submit", [vp, C.POINTER(Request), C.POINTER(i64)])
    _sig(L, "rf_server_wait", [vp, i64, C.c_double, C.POINTER(RequestResult)])
    _sig(L, "rf_server_read", [vp, i64, vp])
    _sig(L, "rf_server_release", [vp, i64])
    _sig(L, "rf_server_get_stats", [vp, C.POINTER(ServerStats)])
    _lib = L
    return L


def flib() -> C.CDLL:
    """The Fortran facade (reference-order draws with the compiler's RNG)."""
    global _flib
    if _flib is not None:
        return _flib
    lib()  # the facade links librf_hip.so
    path = LIBDIR / "librandomfield_gpu.so"
    if not path.exists():
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
    F = C.CDLL(str(path))
    F.rf_fortran_draws.argtypes = [C.POINTER(Params), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    F.rf_fortran_seed_size.restype = C.c_int
    _flib = F
    return F


class RFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"status {code}: {msg}")
        self.code = code


def check(rc: int, ctx=None) -> None:
    if rc != RF_OK:
        msg = lib().rf_last_error(ctx).decode(errors="replace")
        raise RFError(rc, msg)

"""randomfield_amd -- Python mirror of the gfx950 spectral random-field path.

The product is the C ABI of ``lib/librf_hip.so`` (include/rf_capi.h) and the
Fortran facade ``lib/librandomfield_gpu.so``; this module wraps them for
tests, benchmarks and Python hosts.  Arrays follow the reference's
column-major layouts seen from C: points (N, nDim) row-major == nDim x N
column-major; fields (Nmc, N) row-major == N x Nmc column-major.

Mirror of the reference surface:
    create_random_field(...)  createRandomField / create_RF_Unstruct_Init
                              (calls_RF.f90:20-22, randomFieldND.f90:27-132)
    multivariate_transformation  multiVariateTransformation (randomFieldND.f90:139-178)
    reference_seed(seedStart)    calculate_random_seed (math_RF.f90:538-566)
    reference_draws(...)         the random_number sequence of one call
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from ._capi import (  # noqa: F401
    CM_COS, CM_GAUSSIAN, DRAW_NEXT, DTYPE_F32, DTYPE_F64, FIXED, FOM_GAUSSIAN, FOM_LOGNORMAL, FOM_NONE,
    FFT, FORM_FP32, FORM_FP64, FORM_MFMA_BF16, FORM_MFMA_F16, ISOTROPIC, KSAMP_BOX, KSAMP_DENSITY,
    MFMA_VARIANTS, RANDOMIZATION, REF_COMPAT, RF_E_ARG, RF_E_HIP, RF_E_LOGNORMAL, RF_E_NODEVICE, RF_E_NOMEM,
    RF_E_UNSUPPORTED, RF_OK, SHINOZUKA, EXPORTS, LayoutInfo, LayoutParams, Params, RFError, SpectrumInfo, check,
    flib, lib, make_params,
)

__all__ = [
    "Context", "FieldPlan", "make_params", "plan_spectrum", "spectrum_tables", "generate", "generate_grid",
    "create_random_field", "multivariate_transformation", "reference_seed", "reference_draws",
    "blend", "BlendPlan", "GridBatch", "Graph", "stats", "corr", "interpolate", "interpolate_box", "read_unv",
    "normalize", "localization_layout",
    "device_uniforms", "RFError", "MFMA_VARIANTS",
]


class _Arr:
    """Keeps a host array alive for the duration of the foreign call."""

    def __init__(self, a: np.ndarray):
        self.a = a
        self._as_parameter_ = C.c_void_p(a.ctypes.data)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return _Arr(a)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(int(a))


def _check_device(t, name: str, dtypes=("float32", "float64"), shape=None, device=None) -> None:
    """Refuses a buffer whose address must not reach a kernel: host arrays,
    CPU tensors, non-contiguous views, other dtypes or shapes, or a tensor on
    another device than the context's."""
    if isinstance(t, np.ndarray) or not hasattr(t, "data_ptr") or not hasattr(t, "is_cuda"):
        raise ValueError(f"{name} must be a device tensor, not {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} is not in device memory")
    dt = str(t.dtype).replace("torch.", "")
    if dt not in dtypes:
        raise ValueError(f"{name} has dtype {dt}; expected one of {', '.join(dtypes)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} is not contiguous")
    if shape is not None and tuple(int(v) for v in t.shape) != tuple(int(v) for v in shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {tuple(shape)}")
    if device is not None and t.device.index != device:
        raise ValueError(f"{name} is on device {t.device.index}, the context on device {device}")


def _dtype_code(dt) -> int:
    s = str(dt)
    if "float64" in s or s == "f8":
        return DTYPE_F64
    if "float32" in s or s == "f4":
        return DTYPE_F32
    raise ValueError(f"unsupported dtype {dt}")


def plan_spectrum(params: Params) -> SpectrumInfo:
    info = SpectrumInfo()
    check(lib().rf_plan_spectrum(C.byref(params), C.byref(info)))
    return info


def spectrum_tables(params: Params, kU: np.ndarray | None = None):
    """(kPoints (kNTotal, rows), SkVec (kNTotal,)) in normalised coordinates."""
    info = plan_spectrum(params)
    rows = 1 if params.method == ISOTROPIC else params.nDim
    k = np.zeros((info.kNTotal, rows))
    sk = np.zeros(info.kNTotal)
    if kU is not None:
        kU = np.ascontiguousarray(kU, dtype=np.float64)
    check(lib().rf_spectrum_tables(C.byref(params), _ptr(kU), _ptr(k), _ptr(sk)))
    return k, sk


def reference_seed(seedStart: int, n: int | None = None) -> np.ndarray:
    """calculate_random_seed: seed(k) = 72 + 18*seedStart + 37*(k-1)."""
    if n is None:
        n = flib().rf_fortran_seed_size()
    return np.array([72 + 18 * seedStart + 37 * k for k in range(n)], dtype=np.int32)


def reference_draws(params: Params, seed, calc=None):
    """Uniform draws of one createRandomField call, made by the Fortran facade
    with random_number after random_seed(PUT=seed): (kU, phU (Nmc, nPhase))."""
    info = plan_spectrum(params)
    seed = np.ascontiguousarray(np.atleast_1d(seed), dtype=np.int32)
    c = np.ones(params.Nmc, dtype=np.int8) if calc is None else np.ascontiguousarray(calc, dtype=np.int8)
    kU = np.zeros(max(1, info.nKUniform))
    phU = np.zeros((params.Nmc, info.nPhaseUniform))
    flib().rf_fortran_draws(C.byref(params), _ptr(seed), len(seed), _ptr(c), _ptr(kU), _ptr(phU))
    return (kU[: info.nKUniform] if info.nKUniform > 0 else None), phU


class Context:
    """rf_context: device binding and stream."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        self.device = int(device)
        self._children = []            # weak references to the objects made on this context
        check(lib().rf_context_create(C.byref(self._h), device))

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _adopt(self, obj) -> None:
        """Registers an object whose native handle lives on this context, so
        that close() releases it first (a native object must not outlive its
        context: its destroy call uses the context's device and stream)."""
        self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))

    def sync(self) -> None:
        check(lib().rf_context_sync(self._h), self._h)

    def close(self) -> None:
        if self._h:
            for r in reversed(self._children):      # newest first
                obj = r()
                if obj is not None:
                    obj.close()
            self._children = []
            lib().rf_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FieldPlan:
    """rf_field_plan: device-resident wavevectors, phases and amplitudes."""

    def __init__(self, ctx: Context, params: Params, kU=None, phU=None, calc=None):
        self.ctx = ctx
        self.params = params
        self._h = C.c_void_p()
        kU = None if kU is None else np.ascontiguousarray(kU, dtype=np.float64)
        phU = None if phU is None else np.ascontiguousarray(phU, dtype=np.float64)
        c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
        check(lib().rf_plan_create(ctx.handle, C.byref(params), _ptr(kU), _ptr(phU), _ptr(c),
                                   C.byref(self._h)), ctx.handle)
        ctx._adopt(self)

    @property
    def info(self) -> SpectrumInfo:
        info = SpectrumInfo()
        check(lib().rf_plan_info(self._h, C.byref(info)))
        return info

    def _check_points(self, x, out) -> int:
        dev = getattr(self.ctx, "device", None)
        _check_device(x, "x", ("float64",), device=dev)
        if x.dim() != 2 or int(x.shape[1]) != self.params.nDim:
            raise ValueError(f"x has shape {tuple(x.shape)}; expected (N, {self.params.nDim})")
        N = int(x.shape[0])
        _check_device(out, "out", shape=(self.params.Nmc, N), device=dev)
        return N

    def execute(self, x, out, stream=None) -> None:
        """x: device (N, nDim) float64; out: device (Nmc, N) float32/float64."""
        N = self._check_points(x, out)
        check(lib().rf_plan_execute(self._h, _ptr(x), N, _ptr(out), _dtype_code(out.dtype),
                                    _ptr(stream)), self.ctx.handle)

    def set_variant(self, variant: int) -> None:
        """MFMA launch shape 0 .. MFMA_VARIANTS-1 (waves per block, K-steps per
        LDS stage, MFMA issue order)."""
        check(lib().rf_plan_set_variant(self._h, int(variant)), self.ctx.handle)

    def autotune(self, x, out, stream=None) -> int:
        """Times every MFMA launch shape on these points, keeps the fastest."""
        N = self._check_points(x, out)
        chosen = C.c_int32(0)
        check(lib().rf_plan_autotune(self._h, _ptr(x), N, _ptr(out), _dtype_code(out.dtype),
                                     _ptr(stream), C.byref(chosen)), self.ctx.handle)
        return int(chosen.value)

    def execute_grid(self, origin, step, nstep, out, stream=None) -> None:
        """Structured grid (x fastest): out device (Nmc, prod(nstep))."""
        o = np.zeros(3); o[: len(origin)] = origin
        st = np.ones(3); st[: len(step)] = step
        ns = np.ones(3, dtype=np.int32); ns[: len(nstep)] = nstep
        _check_device(out, "out", shape=(self.params.Nmc, int(np.prod(ns, dtype=np.int64))),
                      device=getattr(self.ctx, "device", None))
        check(lib().rf_plan_execute_grid(self._h, _ptr(o), _ptr(st), _ptr(ns), _ptr(out),
                                         _dtype_code(out.dtype), _ptr(stream)), self.ctx.handle)

    def redraw(self, draw: int, stream=None) -> None:
        """Fresh realizations draw*Nmc .. draw*Nmc + Nmc - 1 of the plan's
        device phase stream (draw >= 0; 0 repeats a plan made without caller
        draws); phases, directions, DCT coefficients and operand tables are
        re-formed on the device."""
        check(lib().rf_plan_redraw(self._h, int(draw), _ptr(stream)), self.ctx.handle)

    def close(self) -> None:
        if self._h:
            lib().rf_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate(ctx: Context, params: Params, x: np.ndarray, kU=None, phU=None, calc=None,
             dtype=np.float64) -> np.ndarray:
    """rf_generate on host arrays: x (N, nDim) -> field (Nmc, N)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((params.Nmc, x.shape[0]), dtype=dtype)
    kU = None if kU is None else np.ascontiguousarray(kU, dtype=np.float64)
    phU = None if phU is None else np.ascontiguousarray(phU, dtype=np.float64)
    c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
    check(lib().rf_generate(ctx.handle, C.byref(params), _ptr(x), x.shape[0], _ptr(kU), _ptr(phU),
                            _ptr(c), _ptr(out), _dtype_code(out.dtype)), ctx.handle)
    return out


def generate_grid(ctx: Context, params: Params, origin, step, nstep, phU=None, calc=None,
                  dtype=np.float64) -> np.ndarray:
    """rf_generate_grid on host arrays: the grid of nstep points from origin at
    step (x fastest) -> field (Nmc, prod(nstep))."""
    o = np.zeros(3); o[: len(origin)] = origin
    st = np.ones(3); st[: len(step)] = step
    ns = np.ones(3, dtype=np.int32); ns[: len(nstep)] = nstep
    out = np.zeros((params.Nmc, int(np.prod(ns, dtype=np.int64))), dtype=dtype)
    phU = None if phU is None else np.ascontiguousarray(phU, dtype=np.float64)
    c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
    check(lib().rf_generate_grid(ctx.handle, C.byref(params), _ptr(o), _ptr(st), _ptr(ns), _ptr(phU), _ptr(c),
                                 _ptr(out), _dtype_code(out.dtype)), ctx.handle)
    return out


def create_random_field(ctx: Context, x: np.ndarray, *, method=SHINOZUKA, corrMod=CM_GAUSSIAN,
                        corrL=(1.0, 1.0, 1.0), xRange=None, Nmc=1, seedStart=0, mode=REF_COMPAT,
                        form=FORM_FP64, margiFirst=FOM_NONE, fieldAvg=0.0, fieldVar=1.0,
                        kSampling=KSAMP_BOX, nKDensity=0, calc=None, isoShared=0,
                        dtype=np.float64) -> np.ndarray:
    """createRandomField with the reference's draws for seedStart (Fortran
    random_number); xRange defaults to the points' bounding box."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nDim = x.shape[1]
    if xRange is None:
        ext = x.max(axis=0) - x.min(axis=0)
        xRange = [e if e > 0 else corrL[i] for i, e in enumerate(ext)]
    p = make_params(nDim=nDim, method=method, corrMod=corrMod, mode=mode, form=form,
                    kSampling=kSampling, Nmc=Nmc, margiFirst=margiFirst, nKDensity=nKDensity,
                    corrL=corrL[:nDim], xRange=xRange, fieldAvg=fieldAvg, fieldVar=fieldVar,
                    isoShared=isoShared)
    kU, phU = reference_draws(p, reference_seed(seedStart), calc)
    return generate(ctx, p, x, kU, phU, calc, dtype)


def multivariate_transformation(ctx: Context, margiFirst: int, fieldAvg: float, fieldVar: float,
                                field: np.ndarray) -> None:
    """In place on a host float64 array (GPU transform)."""
    assert field.dtype == np.float64 and field.flags.c_contiguous
    check(lib().rf_marginal_host(ctx.handle, margiFirst, fieldAvg, fieldVar, _ptr(field), field.size),
          ctx.handle)


def blend(ctx: Context, groupN, Nmc, origins, sizes, nOvl, faceMasks, mode, d_fields, d_group,
          dtype_code, stream=None) -> None:
    """rf_blend_device over device buffers (d_fields: list of device pointers/tensors)."""
    nDim = len(groupN)
    g = np.ascontiguousarray(groupN, dtype=np.int32)
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1)
    s = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
    n = np.ascontiguousarray(nOvl, dtype=np.int32)
    m = np.ascontiguousarray(faceMasks, dtype=np.int32)
    ptrs = (C.c_void_p * len(d_fields))(*[_ptr(f).value for f in d_fields])
    check(lib().rf_blend_device(ctx.handle, nDim, _ptr(g), Nmc, len(d_fields), _ptr(o), _ptr(s),
                                _ptr(n), _ptr(m), mode, C.cast(ptrs, C.c_void_p), _ptr(d_group),
                                dtype_code, _ptr(stream)), ctx.handle)


class BlendPlan:
    """rf_blend_plan: blend tables uploaded once; execute() is one launch."""

    def __init__(self, ctx: Context, groupN, Nmc, origins, sizes, nOvl, faceMasks, mode, d_fields, dtype_code):
        self.ctx = ctx
        self._fields = list(d_fields)
        g = np.ascontiguousarray(groupN, dtype=np.int32)
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1)
        s = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
        n = np.ascontiguousarray(nOvl, dtype=np.int32)
        m = np.ascontiguousarray(faceMasks, dtype=np.int32)
        ptrs = (C.c_void_p * len(self._fields))(*[_ptr(f).value for f in self._fields])
        self._h = C.c_void_p()
        check(lib().rf_blend_plan_create(ctx.handle, len(groupN), _ptr(g), Nmc, len(self._fields), _ptr(o), _ptr(s),
                                         _ptr(n), _ptr(m), mode, C.cast(ptrs, C.c_void_p), dtype_code,
                                         C.byref(self._h)), ctx.handle)
        ctx._adopt(self)

    def execute(self, d_group, stream=None) -> None:
        check(lib().rf_blend_plan_execute(self._h, _ptr(d_group), _ptr(stream)), self.ctx.handle)

    def close(self) -> None:
        if self._h:
            lib().rf_blend_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GridBatch:
    """rf_grid_batch: several structured-grid fields in three grouped launches."""

    def __init__(self, ctx: Context, plans, origins, steps, nsteps, outs, dtype_code):
        self.ctx = ctx
        self._plans, self._outs = list(plans), list(outs)
        n = len(self._plans)

        def pad3(rows, fill, dt):
            a = np.full((n, 3), fill, dtype=dt)
            for i, r in enumerate(rows):
                a[i, : len(r)] = r
            return a
        o, st, ns = pad3(origins, 0.0, np.float64), pad3(steps, 1.0, np.float64), pad3(nsteps, 1, np.int32)
        pp = (C.c_void_p * n)(*[pl._h.value for pl in self._plans])
        op = (C.c_void_p * n)(*[_ptr(t).value for t in self._outs])
        self._h = C.c_void_p()
        check(lib().rf_grid_batch_create(ctx.handle, n, C.cast(pp, C.c_void_p), _ptr(o), _ptr(st), _ptr(ns),
                                         C.cast(op, C.c_void_p), dtype_code, C.byref(self._h)), ctx.handle)
        ctx._adopt(self)

    def execute(self, stream=None) -> None:
        check(lib().rf_grid_batch_execute(self._h, _ptr(stream)), self.ctx.handle)

    def autotune(self, stream=None):
        """Times each stage with its wide tile kinds and with 64 x 64 tiles,
        keeps the faster; returns the widest kind used per stage."""
        chosen = (C.c_int32 * 3)()
        check(lib().rf_grid_batch_autotune(self._h, _ptr(stream), chosen), self.ctx.handle)
        return list(chosen)

    def fuse_blend(self, blend_plan, d_group) -> None:
        """Single-covered points go from the last contraction straight into
        d_group; blend_plan.execute(d_group) then forms the overlaps.
        d_group None undoes the fusion."""
        check(lib().rf_grid_batch_fuse_blend(self._h, blend_plan._h, _ptr(d_group)), self.ctx.handle)

    def redraw(self, draw_index: int = -1, stream=None) -> None:
        """New phases for every field: realizations draw_index*Nmc.. of each
        plan's device stream, or (-1) the next set from the batch's counter."""
        check(lib().rf_grid_batch_redraw(self._h, int(draw_index), _ptr(stream)), self.ctx.handle)

    def close(self) -> None:
        if self._h:
            lib().rf_grid_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Captures the library launches enqueued on `stream` inside the with
    block; launch() replays them."""

    def __init__(self, ctx: Context, stream):
        self.ctx, self.stream = ctx, stream
        self._h = C.c_void_p()

    def __enter__(self):
        check(lib().rf_graph_capture_begin(self.ctx.handle, _ptr(self.stream)), self.ctx.handle)
        return self

    def __exit__(self, et, ev, tb):
        rc = lib().rf_graph_capture_end(self.ctx.handle, _ptr(self.stream), C.byref(self._h))
        if self._h:
            self.ctx._adopt(self)
        if et is None:
            check(rc, self.ctx.handle)

    def launch(self, stream=None) -> None:
        check(lib().rf_graph_launch(self._h, _ptr(stream if stream is not None else self.stream)), self.ctx.handle)

    def close(self) -> None:
        if self._h:
            lib().rf_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stats(ctx: Context, d_field, N: int, Nmc: int, dtype_code: int):
    avg = np.zeros(Nmc)
    sd = np.zeros(Nmc)
    g = np.zeros(2)
    check(lib().rf_stats_device(ctx.handle, _ptr(d_field), N, Nmc, dtype_code, _ptr(avg), _ptr(sd),
                                _ptr(g)), ctx.handle)
    return avg, sd, g


def corr(ctx: Context, d_field, dtype_code: int, nstep, xStep, avg: float, std: float, realization: int = 0):
    """rf_corr_device: per-axis lag correlations (list of arrays) and corrL."""
    n = np.ascontiguousarray(nstep, dtype=np.int32)
    M = [int(v) // 2 for v in n]
    R = np.zeros(sum(M))
    cl = np.zeros(len(n))
    check(lib().rf_corr_device(ctx.handle, _ptr(d_field), dtype_code, len(n), _ptr(n),
                               _ptr(np.asarray(xStep, np.float64)), realization, avg, std, _ptr(R), _ptr(cl)),
          ctx.handle)
    return list(np.split(R, np.cumsum(M)[:-1])), cl


def interpolate(ctx: Context, nstep, xMin, xStep, Nmc: int, d_grid, d_nodes, d_out, stream=None) -> bool:
    """rf_interpolate_device: grid field (Nmc, prod(nstep)) to nodes (n, nDim)
    into d_out (Nmc, n); returns True if some node was outside the grid."""
    n = np.ascontiguousarray(nstep, dtype=np.int32)
    flag = C.c_int32(0)
    check(lib().rf_interpolate_device(ctx.handle, len(n), _ptr(n), _ptr(np.asarray(xMin, np.float64)),
                                      _ptr(np.asarray(xStep, np.float64)), Nmc, _ptr(d_grid),
                                      _dtype_code(d_grid.dtype), _ptr(d_nodes), int(d_nodes.shape[0]), _ptr(d_out),
                                      _dtype_code(d_out.dtype), _ptr(stream), C.byref(flag)), ctx.handle)
    return bool(flag.value)


def interpolate_box(ctx: Context, minPos, maxPos, xMinGlob, xStep, Nmc: int, d_box, d_nodes, d_out,
                    stream=None) -> bool:
    """rf_interpolate_box_device: d_box holds global grid points minPos..maxPos
    (1-based) of the grid starting at xMinGlob; as interpolate otherwise."""
    lo = np.ascontiguousarray(minPos, dtype=np.int32)
    hi = np.ascontiguousarray(maxPos, dtype=np.int32)
    flag = C.c_int32(0)
    check(lib().rf_interpolate_box_device(ctx.handle, len(lo), _ptr(lo), _ptr(hi),
                                          _ptr(np.asarray(xMinGlob, np.float64)), _ptr(np.asarray(xStep, np.float64)),
                                          Nmc, _ptr(d_box), _dtype_code(d_box.dtype), _ptr(d_nodes),
                                          int(d_nodes.shape[0]), _ptr(d_out), _dtype_code(d_out.dtype), _ptr(stream),
                                          C.byref(flag)), ctx.handle)
    return bool(flag.value)


def read_unv(path, nDim: int):
    """readUNV on one rank: (coordList (nNodes, nDim), connectList (nElem,
    maxConnect) with -1 padding, sizeList (nElem,), monoType)."""
    h = C.c_void_p()
    check(lib().rf_unv_read(str(path).encode(), nDim, C.byref(h)))
    try:
        nn, ne, mc, mono = C.c_int64(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
        check(lib().rf_unv_sizes(h, C.byref(nn), C.byref(ne), C.byref(mc), C.byref(mono)))
        coords = np.zeros((nn.value, nDim))
        conn = np.zeros((ne.value, mc.value), dtype=np.int32)
        sizes = np.zeros(ne.value, dtype=np.int32)
        check(lib().rf_unv_copy(h, _ptr(coords), _ptr(conn), _ptr(sizes)))
    finally:
        lib().rf_unv_destroy(h)
    return coords, conn, sizes, bool(mono.value)


def normalize(ctx: Context, d_field, N: int, Nmc: int, dtype_code: int, stream=None) -> None:
    check(lib().rf_normalize_device(ctx.handle, _ptr(d_field), N, Nmc, dtype_code, _ptr(stream)),
          ctx.handle)


def device_uniforms(ctx: Context, seed: int, tag: int, count: int, d_out, stream=None) -> None:
    check(lib().rf_device_uniforms(ctx.handle, seed, tag, count, _ptr(d_out), _ptr(stream)), ctx.handle)


def localization_layout(nFields, pointsPerCorrL, xMinGlob, xMaxGlob, corrL, overlap):
    """rf_localization_layout: the reference's subdomain geometry (one
    localization level).  Returns (LayoutInfo, origins (nSub, nDim), sizes
    (nSub, nDim), faceMasks (nSub,), xMinSub (nSub, nDim)); subdomain s is the
    reference's field s + 1 (last axis fastest)."""
    d = len(nFields)
    p = LayoutParams()
    p.nDim = d

    def fill(arr, vals, default):
        v = list(vals) + [default] * (3 - len(vals))
        for i in range(3):
            arr[i] = v[i]
    fill(p.nFields, nFields, 1)
    fill(p.pointsPerCorrL, pointsPerCorrL, 1)
    fill(p.xMinGlob, xMinGlob, 0.0)
    fill(p.xMaxGlob, xMaxGlob, 0.0)
    fill(p.corrL, corrL, 1.0)
    fill(p.overlap, overlap, 0.0)
    info = LayoutInfo()
    check(lib().rf_localization_layout(C.byref(p), C.byref(info), None, None, None, None))
    n = info.nSub
    org = np.zeros((n, d), dtype=np.int32)
    siz = np.zeros((n, d), dtype=np.int32)
    msk = np.zeros(n, dtype=np.int32)
    xms = np.zeros((n, d))
    check(lib().rf_localization_layout(C.byref(p), C.byref(info), _ptr(org), _ptr(siz), _ptr(msk), _ptr(xms)))
    return info, org, siz, msk, xms

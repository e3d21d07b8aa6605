"""ctypes binding of liboracle_rf.so (rf_oracle.h) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module; it is the checker, never the thing measured or shipped.
Arrays: points (N, nDim) == nDim x N column-major, fields (Nmc, N).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

REF_COMPAT, FIXED = 0, 1
ISOTROPIC, SHINOZUKA, RANDOMIZATION, FFT = 1, 2, 3, 4
KBOX, KDENSITY = 0, 1
PI_REF = 3.1415926535898


class Plan(C.Structure):
    _fields_ = [
        ("nDim", C.c_int32), ("method", C.c_int32), ("corrMod", C.c_int32), ("mode", C.c_int32),
        ("ksamp", C.c_int32), ("kNStep", C.c_int32 * 3), ("kN_unsigned", C.c_int64),
        ("kNTotal", C.c_int64), ("nKUniform", C.c_int64), ("nSignBlocks", C.c_int32),
        ("kMax", C.c_double * 3), ("kDelta", C.c_double * 3), ("ampMult", C.c_double),
        ("isoStep", C.c_double),
    ]


_L = None


def lib() -> C.CDLL:
    global _L
    if _L is None:
        path = HERE / "liboracle_rf.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.rfo_make_plan.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, vp, C.POINTER(Plan)]
        L.rfo_kpoints.argtypes = [C.POINTER(Plan), vp, vp]
        L.rfo_skvec.argtypes = [C.POINTER(Plan), vp, vp]
        L.rfo_ksign.argtypes = [C.c_int, vp]
        L.rfo_kmax.argtypes = [C.c_int, C.c_int, vp]
        L.rfo_shinozuka.argtypes = [C.POINTER(Plan), vp, C.c_int64, vp, vp, C.c_int, vp, vp, C.c_int64, vp, C.c_int]
        L.rfo_isotropic.argtypes = [C.POINTER(Plan), vp, C.c_int64, vp, vp, C.c_int, vp, vp, vp, C.c_int]
        L.rfo_marginal.argtypes = [C.c_int, C.c_double, C.c_double, vp, C.c_int64]
        L.rfo_overlap_npoints.argtypes = [C.c_int, vp, vp, vp, vp]
        L.rfo_localization_layout.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.rfo_unity_partition.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, vp]
        L.rfo_blend.argtypes = [C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp]
        L.rfo_normalize.argtypes = [vp, C.c_int64, C.c_int]
        L.rfo_stats.argtypes = [vp, C.c_int64, C.c_int, vp, vp, vp, vp]
        L.rfo_corr.argtypes = [vp, C.c_int, vp, C.c_double, C.c_double, vp]
        L.rfo_interp.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, C.c_int64, vp]
        L.rfo_interp.restype = C.c_int64
        L.rfo_corrL.argtypes = [vp, C.c_int, vp, vp, vp]
        L.rfo_iso_nvar.argtypes = [C.c_int]
        L.rfo_fft_plan.argtypes = [C.c_int, C.c_int, vp, vp, C.POINTER(Plan)]
        L.rfo_fft.argtypes = [C.POINTER(Plan), vp, C.c_int, vp, vp, vp, vp]
        _L = L
    return _L


class _Arr:
    """Keeps the array alive for the duration of the foreign call."""

    def __init__(self, a):
        self.a = a
        self._as_parameter_ = C.c_void_p(a.ctypes.data)


def _p(a):
    return None if a is None else _Arr(a)


def make_plan(nDim, method, mode, xRangeNorm, corrMod=1, ksamp=KBOX, nKDensity=0) -> Plan:
    p = Plan()
    xr = np.ascontiguousarray(list(xRangeNorm) + [1.0] * (3 - len(xRangeNorm)), dtype=np.float64)
    rc = lib().rfo_make_plan(nDim, method, corrMod, mode, ksamp, nKDensity, _p(xr), C.byref(p))
    if rc != 0:
        raise ValueError("rfo_make_plan rejected the parameters")
    return p


def fft_plan(nDim, xRangeNorm, gridN, corrMod=1) -> Plan:
    """Method 4 (FFT) plan on the grid of gridN points per axis."""
    p = Plan()
    xr = np.ascontiguousarray(list(xRangeNorm) + [1.0] * (3 - len(xRangeNorm)), dtype=np.float64)
    gn = np.ascontiguousarray(list(gridN) + [1] * (3 - len(gridN)), dtype=np.int32)
    if lib().rfo_fft_plan(nDim, corrMod, _p(xr), _p(gn), C.byref(p)) != 0:
        raise ValueError("rfo_fft_plan rejected the parameters")
    return p


def fft(plan: Plan, sk: np.ndarray, unif: np.ndarray, xStepNorm, Nmc: int, calc=None) -> np.ndarray:
    """Method 4 fields (Nmc, prod gridN), x fastest; unif (Nmc, 3 kN - 1)."""
    kN = plan.kNTotal
    unif = np.ascontiguousarray(unif, dtype=np.float64).reshape(Nmc, 3 * kN - 1)
    xs = np.ascontiguousarray(list(xStepNorm) + [1.0] * (3 - len(xStepNorm)), dtype=np.float64)
    c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
    out = np.zeros((Nmc, kN))
    lib().rfo_fft(C.byref(plan), _p(np.ascontiguousarray(sk, dtype=np.float64)), Nmc, _p(unif), _p(c), _p(xs),
                  _p(out))
    return out


def kpoints(plan: Plan, kU=None) -> np.ndarray:
    rows = 1 if plan.method == ISOTROPIC else plan.nDim
    k = np.zeros((plan.kNTotal, rows))
    kU = None if kU is None else np.ascontiguousarray(kU, dtype=np.float64)
    lib().rfo_kpoints(C.byref(plan), _p(kU), _p(k))
    return k


def skvec(plan: Plan, k: np.ndarray) -> np.ndarray:
    s = np.zeros(plan.kNTotal)
    lib().rfo_skvec(C.byref(plan), _p(np.ascontiguousarray(k)), _p(s))
    return s


def ksign(nDim: int) -> np.ndarray:
    rows = 1 << (nDim - 1)
    out = np.zeros(rows * nDim)
    lib().rfo_ksign(nDim, _p(out))
    return out.reshape(nDim, rows).T      # (row j, column i)


def shinozuka(plan: Plan, xNorm: np.ndarray, k: np.ndarray, sk: np.ndarray, phU: np.ndarray,
              calc=None, xSlab=0, threads=0) -> np.ndarray:
    """xNorm: (N, nDim) coordinates already divided by corrL."""
    xNorm = np.ascontiguousarray(xNorm, dtype=np.float64)
    N = xNorm.shape[0]
    Nmc = phU.shape[0]
    out = np.zeros((Nmc, N))
    c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
    lib().rfo_shinozuka(C.byref(plan), _p(xNorm), N, _p(np.ascontiguousarray(k)), _p(np.ascontiguousarray(sk)),
                        Nmc, _p(np.ascontiguousarray(phU, dtype=np.float64)), _p(c), xSlab, _p(out), threads)
    return out


def isotropic(plan: Plan, xNorm: np.ndarray, kRad: np.ndarray, sk: np.ndarray, unif: np.ndarray,
              calc=None, threads=0) -> np.ndarray:
    xNorm = np.ascontiguousarray(xNorm, dtype=np.float64)
    N = xNorm.shape[0]
    Nmc = unif.shape[0]
    out = np.zeros((Nmc, N))
    c = None if calc is None else np.ascontiguousarray(calc, dtype=np.uint8)
    lib().rfo_isotropic(C.byref(plan), _p(xNorm), N, _p(np.ascontiguousarray(kRad).reshape(-1)),
                        _p(np.ascontiguousarray(sk)), Nmc, _p(np.ascontiguousarray(unif, dtype=np.float64)),
                        _p(c), _p(out), threads)
    return out


def marginal(margi: int, avg: float, var: float, field: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(field, dtype=np.float64).copy()
    rc = lib().rfo_marginal(margi, avg, var, _p(f), f.size)
    if rc != 0:
        raise ValueError("lognormal with fieldAvg <= 0")
    return f


def localization_layout(nFields, pointsPerCorrL, xMinGlob, xMaxGlob, corrL, overlap) -> dict:
    """The reference's localization geometry (one level): per-axis values and
    per-subdomain origins (0-based group index) and first coordinates."""
    d = len(nFields)
    nsub = int(np.prod(nFields))
    g = np.zeros(13 * d)
    org = np.zeros(nsub * d, dtype=np.int32)
    xms = np.zeros(nsub * d)
    i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    n = lib().rfo_localization_layout(d, _p(i32(nFields)), _p(i32(pointsPerCorrL)), _p(f64(xMinGlob)),
                                      _p(f64(xMaxGlob)), _p(f64(corrL)), _p(f64(overlap)), _p(g), _p(org),
                                      _p(xms))
    assert n == nsub
    g = g.reshape(d, 13)
    keys = ["xStep", "overlap", "xMinGlob", "xMaxGlob", "procExtent", "stepProc", "groupN", "subN", "nOvl"]
    out = {k: g[:, i].copy() for i, k in enumerate(keys)}
    for k in ("groupN", "subN", "nOvl"):
        out[k] = out[k].astype(np.int64)
    out["origins"] = org.reshape(nsub, d)
    out["xMinSub"] = xms.reshape(nsub, d)
    return out


def overlap_npoints(overlap, corrL, xStep) -> np.ndarray:
    d = len(overlap)
    out = np.zeros(d, dtype=np.int32)
    lib().rfo_overlap_npoints(d, _p(np.asarray(overlap, np.float64)), _p(np.asarray(corrL, np.float64)),
                              _p(np.asarray(xStep, np.float64)), _p(out))
    return out


def unity_partition(xNStep, nOvl, faceMask: int, mode: int) -> np.ndarray:
    d = len(xNStep)
    xs = np.ascontiguousarray(xNStep, dtype=np.int32)
    no = np.ascontiguousarray(nOvl, dtype=np.int32)
    pu = np.zeros(int(np.prod(xs)))
    lib().rfo_unity_partition(d, _p(xs), _p(no), faceMask, mode, _p(pu))
    return pu


def blend(groupN, Nmc, origins, sizes, nOvl, faceMasks, mode, fields) -> np.ndarray:
    d = len(groupN)
    g = np.ascontiguousarray(groupN, dtype=np.int32)
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1)
    s = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
    no = np.ascontiguousarray(nOvl, dtype=np.int32)
    fm = np.ascontiguousarray(faceMasks, dtype=np.int32)
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    ptrs = (C.c_void_p * len(fs))(*[f.ctypes.data for f in fs])
    out = np.zeros((Nmc, int(np.prod(g))))
    lib().rfo_blend(d, _p(g), Nmc, len(fs), _p(o), _p(s), _p(no), _p(fm), mode, C.cast(ptrs, C.c_void_p), _p(out))
    return out


def normalize(field: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(field, dtype=np.float64).copy()
    lib().rfo_normalize(_p(f), f.shape[1], f.shape[0])
    return f


def stats(field: np.ndarray):
    f = np.ascontiguousarray(field, dtype=np.float64)
    Nmc, N = f.shape
    avg, sd = np.zeros(Nmc), np.zeros(Nmc)
    ga, gs = C.c_double(), C.c_double()
    lib().rfo_stats(_p(f), N, Nmc, _p(avg), _p(sd), C.byref(ga), C.byref(gs))
    return avg, sd, ga.value, gs.value


def iso_nvar(nDim: int) -> int:
    return lib().rfo_iso_nvar(nDim)


def corr(u: np.ndarray, nstep, avg: float, std: float) -> list:
    """Directional circular autocorrelation of one realization on a grid
    (x fastest): a list of nDim arrays of lags 0 .. n_a/2 - 1."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    n = np.ascontiguousarray(nstep, dtype=np.int32)
    M = [int(v) // 2 for v in n]
    R = np.zeros(sum(M))
    lib().rfo_corr(_p(u), len(n), _p(n), avg, std, _p(R))
    return list(np.split(R, np.cumsum(M)[:-1]))


def corrL(R: list, nstep, xStep) -> np.ndarray:
    n = np.ascontiguousarray(nstep, dtype=np.int32)
    flat = np.ascontiguousarray(np.concatenate(R), dtype=np.float64)
    out = np.zeros(len(n))
    lib().rfo_corrL(_p(flat), len(n), _p(n), _p(np.asarray(xStep, np.float64)), _p(out))
    return out


def interp(grid: np.ndarray, nstep, xMin, xStep, x: np.ndarray):
    """Multilinear interpolation of grid (Nmc, prod(nstep)) to nodes x (n, nDim):
    returns (out (Nmc, n), number of nodes outside the grid)."""
    g = np.ascontiguousarray(grid, dtype=np.float64)
    Nmc = g.shape[0] if g.ndim == 2 else 1
    xs = np.ascontiguousarray(x, dtype=np.float64)
    n = np.ascontiguousarray(nstep, dtype=np.int32)
    out = np.zeros((Nmc, xs.shape[0]))
    bad = lib().rfo_interp(len(n), _p(n), _p(np.asarray(xMin, np.float64)), _p(np.asarray(xStep, np.float64)),
                           Nmc, _p(g), _p(xs), xs.shape[0], _p(out))
    return out, int(bad)

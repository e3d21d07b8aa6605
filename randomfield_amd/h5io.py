"""HDF5 / XDMF output of structured fields (librf_h5.so, include/rf_h5.h).

write_simple mirrors write_Simple_pHDF5_Str (writeResultFile_RF.f90:1062-1247)
on one rank and write_attributes mirrors write_HDF5_attributes (:1758-1833).
Fields are host arrays, x fastest (a device tensor is copied back first).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "lib" / "librf_h5.so"
RF_H5_OK, RF_H5_E_ARG, RF_H5_E_IO = 0, -1, -7


class H5Attrs(C.Structure):
    _fields_ = [
        ("nb_procs", C.c_int32), ("nDim", C.c_int32), ("Nmc", C.c_int32), ("method", C.c_int32),
        ("seedStart", C.c_int32), ("corrMod", C.c_int32), ("margiFirst", C.c_int32),
        ("localizationLevel", C.c_int32), ("kNStep", C.c_int32 * 3), ("nFields", C.c_int32 * 3),
        ("xMinGlob", C.c_double * 3), ("xMaxGlob", C.c_double * 3), ("xStep", C.c_double * 3),
        ("corrL", C.c_double * 3), ("overlap", C.c_double * 3), ("procExtent", C.c_double * 3),
        ("kMax_out", C.c_double * 3),
    ]


class H5Box(C.Structure):
    _fields_ = [
        ("nDim", C.c_int32), ("Nmc", C.c_int32), ("xMinGlob", C.c_double * 3), ("xMaxGlob", C.c_double * 3),
        ("xStep", C.c_double * 3), ("xNStep", C.c_int32 * 3), ("minPos", C.c_int32 * 3), ("maxPos", C.c_int32 * 3),
    ]


class H5Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rf_h5 error {code}: {msg}")
        self.code = code


_L = None


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: build it with randomfield_amd.build (needs the HDF5 C library)")
        L = C.CDLL(str(LIB))
        vp, ip = C.c_void_p, C.c_int32
        L.rf_h5_write_simple.argtypes = [C.c_char_p, C.c_char_p, ip, vp, vp, vp, vp, vp, vp, vp, vp,
                                         C.c_char_p, C.c_char_p, ip]
        L.rf_h5_write_attributes.argtypes = [C.c_char_p, C.POINTER(H5Attrs)]
        L.rf_h5_write_xmf.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, ip, vp, vp, vp, C.c_char_p]
        L.rf_h5_read_box_info.argtypes = [C.c_char_p, ip, vp, vp, C.POINTER(H5Box)]
        L.rf_h5_read_box.argtypes = [C.c_char_p, C.POINTER(H5Box), vp]
        L.rf_h5_last_error.restype = C.c_char_p
        _L = L
    return _L


def _check(rc: int) -> None:
    if rc != RF_H5_OK:
        raise H5Error(rc, lib().rf_h5_last_error().decode())


def _i32(v, n):
    return np.ascontiguousarray(list(v)[:n], dtype=np.int32)


def _f64(v, n):
    return np.ascontiguousarray(list(v)[:n], dtype=np.float64)


def write_simple(folder, name, field, xNStep, xStep, xMinGlob, xNStepGlob, origin=None, minPos=None,
                 maxPos=None) -> tuple[str, str]:
    """Writes <folder>/h5/<name>.h5 ("samples": realization 1 of field over
    the global grid xNStepGlob, window minPos..maxPos placed at origin, all
    1-based) and <folder>/xmf/<name>.xmf; returns the two paths."""
    nDim = len(xNStep)
    f = field.detach().cpu().numpy() if hasattr(field, "detach") else np.asarray(field)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    if f.size < int(np.prod(xNStep)):
        raise ValueError("field holds fewer than prod(xNStep) values")
    origin = [1] * nDim if origin is None else origin
    minPos = [1] * nDim if minPos is None else minPos
    maxPos = list(xNStep) if maxPos is None else maxPos
    keep = [_i32(minPos, nDim), _i32(maxPos, nDim), _i32(origin, nDim), _i32(xNStep, nDim), _f64(xStep, nDim),
            _f64(xMinGlob, nDim), _i32(xNStepGlob, nDim), f]
    h5 = C.create_string_buffer(4096)
    xmf = C.create_string_buffer(4096)
    ptrs = [a.ctypes.data for a in keep]
    _check(lib().rf_h5_write_simple(str(folder).encode(), str(name).encode(), nDim, *ptrs, h5, xmf, 4096))
    return h5.value.decode(), xmf.value.decode()


def write_attributes(h5path, *, nDim, Nmc, method, seedStart, corrMod, margiFirst, xMinGlob, xMaxGlob, xStep,
                     corrL, overlap=None, procExtent=None, kMax_out=None, kNStep=None, nFields=None, nb_procs=1,
                     localizationLevel=0) -> None:
    """write_HDF5_attributes on the file root of h5path."""
    a = H5Attrs()
    a.nb_procs, a.nDim, a.Nmc, a.method = nb_procs, nDim, Nmc, method
    a.seedStart, a.corrMod, a.margiFirst, a.localizationLevel = seedStart, corrMod, margiFirst, localizationLevel
    z = [0.0] * nDim
    for fld, v in (("xMinGlob", xMinGlob), ("xMaxGlob", xMaxGlob), ("xStep", xStep), ("corrL", corrL),
                   ("overlap", overlap or z), ("procExtent", procExtent or z), ("kMax_out", kMax_out or z)):
        setattr(a, fld, (C.c_double * 3)(*(list(v)[:nDim] + [0.0] * (3 - nDim))))
    for fld, v in (("kNStep", kNStep or [0] * nDim), ("nFields", nFields or [1] * nDim)):
        setattr(a, fld, (C.c_int32 * 3)(*(list(v)[:nDim] + [0] * (3 - nDim))))
    _check(lib().rf_h5_write_attributes(str(h5path).encode(), C.byref(a)))


def read_box_info(h5path, nDim: int, xMinLoc, xMaxLoc) -> H5Box:
    """Attributes and the bounding box of [xMinLoc, xMaxLoc] (interpolateToMesh's read)."""
    b = H5Box()
    lo, hi = _f64(xMinLoc, nDim), _f64(xMaxLoc, nDim)
    _check(lib().rf_h5_read_box_info(str(h5path).encode(), nDim, lo.ctypes.data, hi.ctypes.data, C.byref(b)))
    return b


def read_box(h5path, box: H5Box) -> np.ndarray:
    """The box of the samples dataset, x fastest (flat)."""
    n = int(np.prod([box.maxPos[i] - box.minPos[i] + 1 for i in range(box.nDim)]))
    out = np.zeros(n)
    _check(lib().rf_h5_read_box(str(h5path).encode(), C.byref(box), out.ctypes.data))
    return out


def interpolate_to_mesh(ctx, h5path, coordList, xMinLoc=None, xMaxLoc=None) -> np.ndarray:
    """interpolateToMesh (sample_RF.f90:872-1106): the samples of the file at
    the mesh nodes coordList (nNodes, nDim), read over the nodes' bounding
    box and interpolated on the device in FP64.  Returns (nNodes,); nodes
    outside the grid raise ValueError."""
    import torch

    from . import interpolate_box
    x = np.ascontiguousarray(coordList, dtype=np.float64)
    nDim = x.shape[1]
    lo = x.min(axis=0) if xMinLoc is None else np.asarray(xMinLoc, np.float64)
    hi = x.max(axis=0) if xMaxLoc is None else np.asarray(xMaxLoc, np.float64)
    box = read_box_info(h5path, nDim, lo, hi)
    data = read_box(h5path, box)
    d_box = torch.from_numpy(data).cuda()
    d_x = torch.from_numpy(x).cuda()
    out = torch.empty((1, x.shape[0]), dtype=torch.float64, device="cuda")
    mn = [box.minPos[i] for i in range(nDim)]
    mx = [box.maxPos[i] for i in range(nDim)]
    if interpolate_box(ctx, mn, mx, [box.xMinGlob[i] for i in range(nDim)], [box.xStep[i] for i in range(nDim)],
                       1, d_box, d_x, out):
        raise ValueError("mesh nodes outside the grid of " + str(h5path))
    return out[0].cpu().numpy()

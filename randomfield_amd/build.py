"""Build the native parts in-tree.

    librf_hip.so          HIP kernels + C ABI (hipcc, gfx950 code objects)
    librandomfield_gpu.so Fortran facade over the C ABI (amdflang)
    librf_h5.so           HDF5 / XDMF writer of structured fields (host, g++)
    rf_peaks              measured MFMA / transcendental ceilings (bench.py --peaks;
                          source rf_peaks.hip beside bench.py)
    tests/fortran/bin/*   Fortran drivers of the facade used by the test suite
    oracle/liboracle_rf.so  the CPU oracle (test infrastructure, gcc)

Run ``python -m randomfield_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "randomfield_amd"
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "lib"
OBJ = PKG / "lib" / "obj"
ORACLE = ROOT / "oracle"
FTESTS = ROOT / "tests" / "fortran"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("RF_FC", "/opt/rocm/bin/amdflang")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# sources compiled as several translation units (one per value of a macro)
SPLIT = {"rf_mfma.hip": 7}
SPLIT_MACRO = {"rf_mfma.hip": "RF_MFMA_PART"}


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_hip(jobs: int = 8, incremental: bool = False) -> Path:
    """librf_hip.so; incremental=True recompiles only the sources newer than
    their object (or older than any header), build_all() always compiles all."""
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    headers = list(CSRC.glob("*.h")) + list((ROOT / "include").glob("*.h"))
    newest_header = max((h.stat().st_mtime for h in headers), default=0.0)
    objs = []
    cmds = []
    for s in srcs:
        for part in range(SPLIT.get(s.name, 0) or 1):
            o = OBJ / (s.name + (f".{part}" if s.name in SPLIT else "") + ".o")
            objs.append(o)
            if incremental and o.exists() and o.stat().st_mtime > max(s.stat().st_mtime, newest_header):
                continue
            lang = ["-x", "hip"] if s.suffix == ".hip" else []
            defs = [f"-D{SPLIT_MACRO[s.name]}={part}"] if s.name in SPLIT else []
            cmds.append([HIPCC, *HIP_FLAGS, *defs, *lang, "-c", str(s), "-o", str(o)])
    for stale in OBJ.glob("*.o"):
        if stale not in objs:
            stale.unlink()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "librf_hip.so"
    tmp = LIB / "librf_hip.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_fortran() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    mod = LIB / "mod"
    mod.mkdir(exist_ok=True)
    out = LIB / "librandomfield_gpu.so"
    tmp = LIB / "librandomfield_gpu.so.tmp"
    _run([FC, "-O2", "-fPIC", "-shared", "-module-dir", str(mod),
          str(FSRC / "randomField_gpu.f90"), "-o", str(tmp),
          f"-L{LIB}", "-lrf_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    # Fortran drivers of the facade (test programs)
    bindir = FTESTS / "bin"
    bindir.mkdir(parents=True, exist_ok=True)
    for src in sorted(FTESTS.glob("*.f90")):
        exe = bindir / src.stem
        _run([FC, "-O2", "-I", str(mod), str(src), "-o", str(exe),
              f"-L{LIB}", "-lrandomfield_gpu", "-lrf_hip", f"-Wl,-rpath,{LIB}",
              "-Wl,-rpath,$ORIGIN/../../../randomfield_amd/lib"])
    return out


H5_INC = Path(os.environ.get("RF_HDF5_INC", "/opt/conda/include"))
H5_LIB = Path(os.environ.get("RF_HDF5_LIB", "/opt/conda/lib"))


def build_h5() -> Path | None:
    """librf_h5.so: the HDF5/XDMF writer (host only).  Skipped, with a note,
    where the HDF5 C library is absent."""
    if not (H5_INC / "hdf5.h").exists() or not (H5_LIB / "libhdf5.so").exists():
        sys.stderr.write(f"note: HDF5 not found under {H5_INC} / {H5_LIB}: librf_h5.so not built\n")
        return None
    out = LIB / "librf_h5.so"
    tmp = LIB / "librf_h5.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{H5_INC}",
          str(PKG / "h5" / "rf_h5.cpp"), "-o", str(tmp), f"-L{H5_LIB}", "-lhdf5", f"-Wl,-rpath,{H5_LIB}"])
    os.replace(tmp, out)
    return out


def build_peaks() -> Path:
    """lib/rf_peaks: measured ceilings of the MFMA and transcendental units
    (bench.py reports them beside the datasheet peaks)."""
    out = LIB / "rf_peaks"
    tmp = LIB / "rf_peaks.tmp"
    _run([HIPCC, *HIP_FLAGS, str(ROOT / "rf_peaks.hip"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    out = ORACLE / "liboracle_rf.so"
    tmp = ORACLE / "liboracle_rf.so.tmp"
    _run([CC, "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
          str(ORACLE / "rf_oracle.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 8) -> None:
    build_oracle()
    build_hip(jobs)
    build_peaks()
    build_h5()
    if shutil.which(FC) or Path(FC).exists():
        build_fortran()
    else:
        raise RuntimeError(f"Fortran compiler not found: {FC}")


if __name__ == "__main__":
    build_all()
    print("built:", LIB / "librf_hip.so", LIB / "librandomfield_gpu.so", ORACLE / "liboracle_rf.so")

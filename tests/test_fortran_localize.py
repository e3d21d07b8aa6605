"""The Fortran facade's localization step (tests/fortran/rf_localize.f90),
written against the reference's calling surface, checked against the oracle:
geometry (round_basic_inputs, set_global_extremes, prepare_Localization),
the partition of unity, every field, the blend and normalize_randField."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "fortran" / "bin" / "rf_localize"


class _Reader:
    def __init__(self, raw):
        self.raw, self.off = raw, 0

    def take(self, dtype, count):
        a = np.frombuffer(self.raw[self.off: self.off + np.dtype(dtype).itemsize * count], dtype)
        self.off += a.nbytes
        return a


def _run(tmp_path, mode):
    out = tmp_path / f"loc{mode}.bin"
    r = subprocess.run([str(EXE), str(mode), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = _Reader(out.read_bytes())
    d = {}
    nDim = int(rd.take(np.int32, 1)[0])
    d["nFields"] = rd.take(np.int32, 3)
    d["Nmc"] = int(rd.take(np.int32, 1)[0])
    for k in ("xStep", "overlap", "xMinGlob", "xMaxGlob", "procExtent", "procStart", "stepProc"):
        d[k] = rd.take(np.float64, 3)
    for k in ("xNStep_Proc", "xNStep_Group", "origin_Group"):
        d[k] = rd.take(np.int32, 3)
    nTot = int(np.prod(d["nFields"]))
    d["subdivisionCoords"] = rd.take(np.float64, 3 * nTot).reshape(nTot, 3)
    d["xMinFiles"] = rd.take(np.float64, 3 * nTot).reshape(nTot, 3)
    d["xMin_Group"] = rd.take(np.float64, 3)
    nP = int(np.prod(d["xNStep_Proc"]))
    nG = int(np.prod(d["xNStep_Group"]))
    d["pu"] = rd.take(np.float64, nP)
    if mode >= 2:
        Nmc = d["Nmc"]
        d["fields"] = rd.take(np.float64, nP * Nmc * nTot).reshape(nTot, Nmc, nP)
        d["group"] = rd.take(np.float64, nG * Nmc).reshape(Nmc, nG)
        d["groupNorm"] = rd.take(np.float64, nG * Nmc).reshape(Nmc, nG)
        d["minP"] = rd.take(np.int32, 3)
        d["maxP"] = rd.take(np.int32, 3)
    assert nDim == 3
    return d


def test_geometry_and_partition(tmp_path):
    """Box [0,6]^3, corrL 1, 2 points per corrL, overlap 2 corrL, 2 x 2 x 2:
    xStep 0.5, overlap 4 steps; 12 - (4 + 6) = 2 -> 1 step per subdomain;
    extent 1 + 8 - 2 = 7 steps (3.5), stride 5 steps (2.5), 8 points, group
    13 points; 3 overlap points."""
    d = _run(tmp_path, 1)
    ref = O.localization_layout([2, 2, 2], [2, 2, 2], [0, 0, 0], [6, 6, 6], [1, 1, 1], [2, 2, 2])
    assert d["xStep"].tolist() == [0.5] * 3 and d["overlap"].tolist() == [2.0] * 3
    assert d["procExtent"].tolist() == [3.5] * 3 and d["stepProc"].tolist() == [2.5] * 3
    assert d["xMaxGlob"].tolist() == [6.0] * 3 and d["procStart"].tolist() == [0.0] * 3
    assert d["xNStep_Proc"].tolist() == [8] * 3 and d["xNStep_Group"].tolist() == [13] * 3
    assert d["origin_Group"].tolist() == [1, 1, 1]
    for k in ("xStep", "overlap", "xMinGlob", "xMaxGlob", "procExtent", "stepProc"):
        assert np.allclose(d[k], ref[k], rtol=0, atol=1e-12), k
    assert d["xNStep_Group"].tolist() == ref["groupN"].tolist()
    # field order: last axis fastest
    assert np.allclose(d["subdivisionCoords"], ref["xMinSub"], rtol=0, atol=1e-12)
    assert d["subdivisionCoords"][1].tolist() == [0.0, 0.0, 2.5]
    pu = O.unity_partition([8, 8, 8], ref["nOvl"].astype(np.int32), 63, O.REF_COMPAT)
    assert np.array_equal(d["pu"], pu)


def test_unity_partition_entry_matches_oracle():
    """rf_unity_partition (host) against the oracle in both modes and on
    subsets of tapered faces, bit for bit."""
    import ctypes as C
    for sizes, novl in (((8, 8, 8), (3, 3, 3)), ((12, 5), (4, 2)), ((9,), (0,))):
        d = len(sizes)
        for mode in (R.REF_COMPAT, R.FIXED):
            for mask in (0, 1, 2, 3, (1 << 2 * d) - 1, 0b100101 & ((1 << 2 * d) - 1)):
                got = np.zeros(int(np.prod(sizes)))
                n = np.ascontiguousarray(sizes, np.int32)
                o = np.ascontiguousarray(novl, np.int32)
                R.check(R.lib().rf_unity_partition(d, R._ptr(n), R._ptr(o), mask, mode, R._ptr(got)))
                assert np.array_equal(got, O.unity_partition(sizes, novl, mask, mode)), (sizes, mode, mask)
    assert R.lib().rf_unity_partition(2, None, None, 0, 0, None) == R.RF_E_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 3])
def test_localization_pipeline(tmp_path, mode):
    d = _run(tmp_path, mode)
    Nmc = d["Nmc"]
    nP = [8, 8, 8]
    omode = O.REF_COMPAT if mode == 2 else O.FIXED
    xs = d["xStep"]
    # every field: the oracle on the field's grid with the draws of seed
    # seedStart + field number (the Fortran random_number stream)
    # (x, y, z) local indices, x fastest
    idx = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")[::-1], -1).reshape(-1, 3)
    po = O.make_plan(3, O.SHINOZUKA, omode, list(d["procExtent"]))
    k = O.kpoints(po)
    sk = O.skvec(po, k)
    for i in range(8):
        x = d["xMinFiles"][i] + idx * xs
        p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=omode, Nmc=Nmc, xRange=tuple(d["procExtent"]))
        _, phU = R.reference_draws(p, R.reference_seed(100 + i + 1))
        ref = O.shinozuka(po, x, k, sk, phU)
        err = np.max(np.abs(d["fields"][i] - ref)) / np.max(np.abs(ref))
        assert err <= 1e-12, (i, err)
    # blend of the Fortran fields against the oracle's add_RF_to_Group + divide
    origins = np.rint((d["xMinFiles"] - d["xMin_Group"]) / xs).astype(np.int32)
    nOvl = [3, 3, 3]
    if mode == 2:
        masks = np.full(8, 63, np.int32)
    else:
        masks = np.array([sum((1 << 2 * a) * (origins[s, a] > 0) + (1 << 2 * a + 1) * (origins[s, a] + 8 < 13)
                              for a in range(3)) for s in range(8)], np.int32)
    refg = O.blend([13, 13, 13], Nmc, origins, [nP] * 8, nOvl, masks, omode, list(d["fields"]))
    assert np.max(np.abs(d["group"] - refg)) <= 1e-14 * np.max(np.abs(refg)) * 10
    # normalize_randField: window mean then RMS of realization 1, applied to all
    g1 = d["group"][0].reshape(13, 13, 13)            # [z][y][x]
    lo, hi = d["minP"] - 1, d["maxP"]
    assert d["maxP"].tolist() == [10, 10, 10]
    win = g1[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    avg = win.mean()
    sd = np.sqrt(((win - avg) ** 2).mean())
    refn = (d["group"] - avg) / sd
    assert np.allclose(d["groupNorm"], refn, rtol=1e-12, atol=1e-12)

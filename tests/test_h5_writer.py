"""HDF5 / XDMF output (librf_h5.so) read back with the HDF5 distribution's
own h5dump, an independent reader: dataset name, type, rank, dimension order
(HDF5 lists the Fortran dimensions last first), values of the written window
and the fill elsewhere, the attribute set of write_HDF5_attributes with its
types, and the XDMF text of write_XMF_Elements (writeResultFile_RF.f90)."""
import re
import shutil
import subprocess
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

from randomfield_amd import h5io

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not h5io.LIB.exists() or not Path(H5DUMP).exists(),
                                reason="HDF5 C library or h5dump absent")


def _dump(*args) -> str:
    return subprocess.run([H5DUMP, *map(str, args)], capture_output=True, text=True, check=True).stdout


def _read(path):
    txt = _dump("-d", "/samples", "-y", "-w", "0", "-m", "%.17g", path)
    dims = tuple(int(v) for v in re.search(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", txt).group(1).split(","))
    dtype = re.search(r"DATATYPE\s+(\S+)", txt).group(1)
    body = txt.split("DATA {", 1)[1].rsplit("}", 3)[0]
    vals = np.array([float(v) for v in re.split(r"[,\s]+", body.strip()) if v])
    return dtype, dims, vals.reshape(dims)


def test_whole_field_3d(tmp_path):
    n = (7, 5, 4)
    f = np.random.default_rng(1).standard_normal(int(np.prod(n)))
    h5, xmf = h5io.write_simple(tmp_path, "field", f, n, (0.5, 0.25, 2.0), (-1.0, 0.0, 3.0), n)
    assert h5 == str(tmp_path / "h5" / "field.h5") and xmf == str(tmp_path / "xmf" / "field.xmf")
    dtype, dims, data = _read(h5)
    assert dtype == "H5T_IEEE_F64LE" and dims == (4, 5, 7)
    assert np.array_equal(data, f.reshape(4, 5, 7))      # x fastest: C order (z, y, x)


def test_window_at_offset_2d_and_3d(tmp_path):
    rng = np.random.default_rng(2)
    n, g = (7, 5, 4), (10, 6, 5)
    f = rng.standard_normal(int(np.prod(n)))
    mn, mx, org = (2, 1, 2), (6, 5, 4), (3, 2, 1)
    h5, _ = h5io.write_simple(tmp_path, "w3", f, n, (1, 1, 1), (0, 0, 0), g, origin=org, minPos=mn, maxPos=mx)
    _, dims, data = _read(h5)
    assert dims == (5, 6, 10)
    want = np.zeros((5, 6, 10))
    F = f.reshape(4, 5, 7)
    want[org[2] - 1:org[2] - 1 + mx[2] - mn[2] + 1, org[1] - 1:org[1] - 1 + mx[1] - mn[1] + 1,
         org[0] - 1:org[0] - 1 + mx[0] - mn[0] + 1] = F[mn[2] - 1:mx[2], mn[1] - 1:mx[1], mn[0] - 1:mx[0]]
    assert np.array_equal(data, want)
    f2 = rng.standard_normal(6 * 3)
    h5, _ = h5io.write_simple(tmp_path, "w2", f2, (6, 3), (1, 1), (0, 0), (6, 3))
    _, dims, data = _read(h5)
    assert dims == (3, 6) and np.array_equal(data, f2.reshape(3, 6))
    with pytest.raises(h5io.H5Error) as e:
        h5io.write_simple(tmp_path, "bad", f, n, (1, 1, 1), (0, 0, 0), g, origin=(5, 2, 1))
    assert e.value.code == h5io.RF_H5_E_ARG


def test_attributes(tmp_path):
    n = (4, 3, 2)
    h5, _ = h5io.write_simple(tmp_path, "a", np.zeros(24), n, (0.5,) * 3, (0.0,) * 3, n)
    h5io.write_attributes(h5, nDim=3, Nmc=4, method=2, seedStart=7, corrMod=1, margiFirst=2,
                          xMinGlob=(0, 0, 0), xMaxGlob=(1.5, 1.0, 0.5), xStep=(0.5, 0.5, 0.5), corrL=(1, 2, 3),
                          overlap=(2, 2, 2), procExtent=(1.5, 1.0, 0.5), kMax_out=(7.1, 7.2, 7.3),
                          kNStep=(14, 15, 16), nFields=(1, 1, 1), nb_procs=1, localizationLevel=0)
    txt = _dump("-A", h5)
    found = re.findall(r'ATTRIBUTE "(\w+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+(SCALAR|SIMPLE)', txt)
    attrs = {n: t for n, t, _ in found}
    space = {n: s for n, _, s in found}
    assert all(space[k] == "SCALAR" for k in ("nDim", "Nmc", "method")) and space["xStep"] == "SIMPLE"
    names = {"nb_procs", "nDim", "Nmc", "method", "seedStart", "corrMod", "margiFirst", "localizationLevel",
             "kNStep", "nFields", "xMinGlob", "xMaxGlob", "xStep", "corrL", "overlap", "procExtent", "kMax_out"}
    assert set(attrs) == names
    for k in ("nb_procs", "nDim", "Nmc", "method", "seedStart", "corrMod", "margiFirst", "localizationLevel"):
        assert attrs[k] == "H5T_STD_I32LE"
    assert attrs["kNStep"] == "H5T_STD_I32LE" and attrs["kMax_out"] == "H5T_IEEE_F64LE"

    def val(name):
        blk = re.search(r'ATTRIBUTE "%s" \{.*?DATA \{(.*?)\}' % name, txt, re.S).group(1)
        return [float(v) for v in re.findall(r"\(\d+\): ([^\n]+)", blk)[0].split(",")]
    assert val("Nmc") == [4.0] and val("seedStart") == [7.0] and val("margiFirst") == [2.0]
    assert val("kNStep") == [14.0, 15.0, 16.0] and val("corrL") == [1.0, 2.0, 3.0]
    assert val("kMax_out") == [7.1, 7.2, 7.3]


def test_xmf_text(tmp_path):
    n = (9, 5, 3)
    _, xmf = h5io.write_simple(tmp_path, "g", np.zeros(135), n, (0.25, 0.5, 1.0), (-1.0, 2.0, 0.5), n)
    txt = Path(xmf).read_text()
    root = ET.fromstring(txt)
    assert root.tag == "Xdmf" and root.get("Version") == "2.0"
    item = root.find("Domain/DataItem")
    assert item.get("Name") == "RF_1" and item.get("Dimensions") == "3 5 9" and item.get("Precision") == "8"
    assert item.text.strip() == "../h5/g.h5:/samples"
    grid = root.find("Domain/Grid/Grid")
    assert grid.find("Topology").get("TopologyType") == "3DCoRectMesh"
    assert grid.find("Topology").get("Dimensions") == "3 5 9"
    geo = grid.find("Geometry")
    assert geo.get("GeometryType") == "ORIGIN_DXDYDZ"
    org, step = geo.findall("DataItem")
    assert [float(v) for v in org.text.split()] == [0.5, 2.0, -1.0]     # last axis first
    assert [float(v) for v in step.text.split()] == [1.0, 0.5, 0.25]
    # F25.10 records: four blanks before the first value, one before the others
    lines = org.text.strip("\n").split("\n")
    assert lines[0] == "    " + f"{0.5:25.10f}" and lines[1] == " " + f"{2.0:25.10f}"
    att = grid.find("Attribute")
    assert att.get("Center") == "Node" and att.find("DataItem").text.strip() == '/Xdmf/Domain/DataItem[@Name="RF_1"]'
    # 2-D geometry and no XDMF in 1-D
    _, x2 = h5io.write_simple(tmp_path, "g2", np.zeros(15), (5, 3), (0.1, 0.2), (0, 0), (5, 3))
    assert ET.parse(x2).getroot().find("Domain/Grid/Grid/Geometry").get("GeometryType") == "ORIGIN_DXDY"
    _, x1 = h5io.write_simple(tmp_path, "g1", np.zeros(5), (5,), (0.1,), (0,), (5,))
    assert not Path(x1).exists()


def test_header_symbols_exported():
    hdr = (Path(__file__).resolve().parent.parent / "include" / "rf_h5.h").read_text()
    decl = set(re.findall(r"\b(rf_h5_[a-z0-9_]+)\s*\(", hdr))
    assert decl == {"rf_h5_write_simple", "rf_h5_write_attributes", "rf_h5_write_xmf", "rf_h5_last_error",
                    "rf_h5_read_box_info", "rf_h5_read_box"}
    for name in decl:
        assert hasattr(h5io.lib(), name), name


def test_read_box(tmp_path):
    """interpolateToMesh's read: attributes, the bounding box (floor / ceiling
    + 1, clamped to the grid) and the box values, x fastest."""
    n, step, x0 = (11, 7, 5), (0.5, 0.25, 1.0), (-1.0, 0.0, 2.0)
    f = np.random.default_rng(3).standard_normal(int(np.prod(n)))
    h5, _ = h5io.write_simple(tmp_path, "b", f, n, step, x0, n)
    xmax = [x0[i] + (n[i] - 1) * step[i] for i in range(3)]
    h5io.write_attributes(h5, nDim=3, Nmc=1, method=2, seedStart=1, corrMod=1, margiFirst=1, xMinGlob=x0,
                          xMaxGlob=xmax, xStep=step, corrL=(1, 1, 1))
    lo, hi = (0.1, 0.3, 1.0), (2.0, 1.4, 9.0)
    b = h5io.read_box_info(h5, 3, lo, hi)
    assert b.nDim == 3 and b.Nmc == 1 and list(b.xNStep) == list(n)
    mn = [int(np.floor((lo[i] - x0[i]) / step[i])) + 1 for i in range(3)]
    mx = [min(int(np.ceil((hi[i] - x0[i]) / step[i])) + 1, n[i]) for i in range(3)]
    mn = [max(v, 1) for v in mn]
    assert list(b.minPos) == mn and list(b.maxPos) == mx and mn[2] == 1 and mx[2] == n[2]
    got = h5io.read_box(h5, b)
    F = f.reshape(n[2], n[1], n[0])
    want = F[mn[2] - 1:mx[2], mn[1] - 1:mx[1], mn[0] - 1:mx[0]].reshape(-1)
    assert np.array_equal(got, want)
    with pytest.raises(h5io.H5Error):
        h5io.read_box_info(h5, 2, lo[:2], hi[:2])          # nDim differs from the file

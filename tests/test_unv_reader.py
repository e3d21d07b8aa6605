"""readUNV (readUNV_RF.f90:14-915) on synthetic universal files: node
coordinates (2411 / 781, D exponents), brick connectivity (2412 / 780, eight
labels per record), other datasets skipped by their tags, blank records,
mixed node counts (monoType false, -1 padding), refused element types."""
import numpy as np
import pytest

import randomfield_amd as R


def _fmt(v):
    return f"{v:25.16E}".replace("E", "D")


def _unv(path, coords, elems, tags=("2411", "2412"), fe=None, extra_blank=False):
    out = ["    -1", "   164", "         1  SI - mks (Newton)  2",
           "  1.0000000000000000D+00  1.0000000000000000D+00  1.0000000000000000D+00",
           "  2.7314999999999998D+02", "    -1"]
    out += ["    -1", f"  {tags[0]}"]
    for i, c in enumerate(coords):
        out.append(f"{i + 1:10d}{1:10d}{1:10d}{11:10d}")
        out.append("".join(_fmt(v) for v in c))
        if extra_blank and i == 1:
            out.append("")
    out += ["    -1", "    -1", f"  {tags[1]}"]
    for i, e in enumerate(elems):
        out.append(f"{i + 1:10d}{(fe or {})[i] if fe and i in fe else 115:10d}{2:10d}{1:10d}{7:10d}{len(e):10d}")
        for k in range(0, len(e), 8):
            out.append("".join(f"{v:10d}" for v in e[k:k + 8]))
    out += ["    -1"]
    path.write_text("\n".join(out) + "\n")


def _brick_mesh(nx, ny, nz, h=(0.5, 0.25, 1.0), x0=(-1.0, 0.3, 2.0)):
    idx = lambda i, j, k: 1 + i + (nx + 1) * (j + (ny + 1) * k)
    coords = [(x0[0] + i * h[0], x0[1] + j * h[1], x0[2] + k * h[2])
              for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)]
    elems = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                elems.append([idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                              idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)])
    return np.array(coords), elems


@pytest.mark.parametrize("tags", [("2411", "2412"), ("781", "780")])
def test_brick_mesh(tmp_path, tags):
    coords, elems = _brick_mesh(3, 2, 2)
    f = tmp_path / "m.unv"
    _unv(f, coords, elems, tags=tags, extra_blank=True)
    c, conn, sizes, mono = R.read_unv(f, 3)
    assert np.array_equal(c, coords)
    assert mono and conn.shape == (12, 8) and np.array_equal(conn, np.array(elems)) and (sizes == 8).all()
    c2, *_ = R.read_unv(f, 2)
    assert np.array_equal(c2, coords[:, :2])


def test_mixed_node_counts(tmp_path):
    coords, elems = _brick_mesh(2, 1, 1)
    elems = [elems[0], list(range(1, 21))]              # 8 and 20 labels (three records)
    f = tmp_path / "mixed.unv"
    _unv(f, coords, elems)
    _, conn, sizes, mono = R.read_unv(f, 3)
    assert not mono and conn.shape == (2, 20)
    assert conn[0, :8].tolist() == elems[0] and (conn[0, 8:] == -1).all()
    assert conn[1].tolist() == elems[1] and sizes.tolist() == [8, 20]


def test_refusals(tmp_path):
    coords, elems = _brick_mesh(1, 1, 1)
    f = tmp_path / "beam.unv"
    _unv(f, coords, elems + [[1, 2]], fe={1: 11})
    with pytest.raises(R.RFError) as e:
        R.read_unv(f, 3)
    assert e.value.code == R.RF_E_UNSUPPORTED
    with pytest.raises(R.RFError) as e:
        R.read_unv(tmp_path / "absent.unv", 3)
    assert e.value.code == R.RF_E_ARG
    g = tmp_path / "empty.unv"
    g.write_text("    -1\n   164\n    -1\n")
    with pytest.raises(R.RFError):
        R.read_unv(g, 3)

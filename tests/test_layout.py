"""The localization geometry (topography_RF.f90:19-255, localization_RF.f90:
127-192,746) restated by rf_localization_layout, against numbers derived by
hand from the reference's rules and against the oracle's literal restatement
of the reference's floating-point steps.  Host only."""
import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O


def test_cfg4_reference_layout_by_hand():
    """512 x 512 x 256 points at spacing 0.25 (box 127.75 x 127.75 x 63.75),
    corrL 1, 4 points per corrL, overlap 2 corrL, 4 x 4 x 2 subdomains.
      xStep 0.25; overlap nint(2/0.25) = 8 steps = 2 corrL
      x, y: overlap total 8*3 + 2*(8-1) = 38 steps; 511 - 38 = 473 -> up to a
            multiple of 4 -> 476, 119 per subdomain; extent 119 + 16 - 2 = 133
            steps (33.25), stride 119 + 8 = 127 steps (31.75); 134 points;
            global box grown to 38 + 476 = 514 steps = 128.5 -> 515 points
      z:    8 + 14 = 22 steps; 255 - 22 = 233 -> 234, 117 per subdomain; extent
            131 steps (32.75), stride 125 steps (31.25); 132 points; 257 points
      overlap points nint((2 - 0.5)/0.25) + 1 = 7 (= 133 - 127 + 1)"""
    info, org, siz, msk, xms = R.localization_layout((4, 4, 2), (4, 4, 4), (0, 0, 0), (127.75, 127.75, 63.75),
                                                     (1, 1, 1), (2, 2, 2))
    assert info.nSub == 32
    assert list(info.xStep) == [0.25] * 3 and list(info.overlap) == [2.0] * 3
    assert list(info.subN) == [134, 134, 132]
    assert list(info.groupN) == [515, 515, 257]
    assert list(info.nOvl) == [7, 7, 7]
    assert list(info.procExtent) == [33.25, 33.25, 32.75]
    assert list(info.stepProc) == [31.75, 31.75, 31.25]
    assert list(info.xMaxGlob) == [128.5, 128.5, 64.0]
    assert list(info.xRangeSub) == [33.25, 33.25, 32.75]
    # field order: last axis fastest (field 2 is the second z subdomain)
    assert org[0].tolist() == [0, 0, 0] and org[1].tolist() == [0, 0, 125] and org[2].tolist() == [0, 127, 0]
    assert org[31].tolist() == [381, 381, 125]
    assert (siz == [134, 134, 132]).all()
    # FIXED taper faces: those with a neighbour
    assert msk[0] == 0b101010 and msk[31] == 0b010101
    assert xms[9].tolist() == [31.75, 0.0, 31.25]
    # the last subdomain ends on the group grid's last point
    assert (org[31] + siz[31] - 1 == np.array(info.groupN)).all() is False or True
    assert (org[31] + siz[31]).tolist() == list(info.groupN)


def test_overlap_rounding_and_minimum():
    # 0.3 corrL at xStep 0.25: nint(1.2) = 1 step < 2 -> 2 steps = 0.5 corrL;
    # overlap points nint((0.5 - 0.5)/0.25) + 1 = 1
    info, *_ = R.localization_layout((3, 2), (4, 4), (0, 0), (20, 10), (1, 1), (0.3, 0.3))
    assert list(info.overlap)[:2] == [0.5, 0.5] and list(info.nOvl)[:2] == [1, 1]
    # 1.4 corrL: nint(5.6) = 6 steps = 1.5 corrL; 5 overlap points
    info, *_ = R.localization_layout((3, 2), (4, 4), (0, 0), (20, 10), (1, 1), (1.4, 1.4))
    assert list(info.overlap)[:2] == [1.5, 1.5] and list(info.nOvl)[:2] == [5, 5]
    # corrL 2, 5 points per corrL: xStep 0.4; overlap 1 corrL = 5 steps; 4 points
    info, *_ = R.localization_layout((2,), (5,), (0,), (30,), (2,), (1.0,))
    assert info.xStep[0] == pytest.approx(0.4) and info.overlap[0] == pytest.approx(1.0)
    assert info.nOvl[0] == 4


def test_unsplit_axis_has_no_overlap():
    # nFields 1: overlap 0, the subdomain spans the rounded box, no taper
    info, org, siz, msk, _ = R.localization_layout((1, 3), (2, 2), (0.1, 0), (9.9, 12), (1, 1), (2, 2))
    assert info.overlap[0] == 0.0 and info.nOvl[0] == 0 and info.nOvlRaw[0] == -1
    # box rounded to steps of 0.5: [0, 10] -> 21 points
    assert info.xMinGlob[0] == 0.0 and info.xMaxGlob[0] == 10.0 and info.subN[0] == 21
    assert (msk & 0b11 == 0).all()
    assert (org[:, 0] == 0).all()


def test_box_rounding_tolerance():
    # areEqual tolerance 1e-9: 10.0000000001 / 0.25 rounds to 40 steps, not 41
    info, *_ = R.localization_layout((1,), (4,), (0.0,), (10.0000000001,), (1,), (0,))
    assert info.subN[0] == 41
    info, *_ = R.localization_layout((1,), (4,), (0.0,), (10.001,), (1,), (0,))
    assert info.subN[0] == 42


@pytest.mark.parametrize("seed", range(40))
def test_layout_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    d = int(rng.integers(1, 4))
    nF = rng.integers(1, 6, size=d)
    ppc = rng.integers(1, 9, size=d)
    corrL = rng.choice([0.5, 1.0, 2.0, 0.8, 1.25], size=d)
    xmin = np.round(rng.uniform(-20, 20, size=d), 3)
    xmax = xmin + np.round(rng.uniform(0, 80, size=d), 3)
    ovl = np.round(rng.uniform(0, 4, size=d), 2)
    info, org, siz, msk, xms = R.localization_layout(nF, ppc, xmin, xmax, corrL, ovl)
    ref = O.localization_layout(nF, ppc, xmin, xmax, corrL, ovl)
    assert np.allclose(info.xStep[:d], ref["xStep"], rtol=1e-15)
    assert np.allclose(info.overlap[:d], ref["overlap"], rtol=1e-12)
    for k in ("xMinGlob", "xMaxGlob", "procExtent", "stepProc"):
        assert np.allclose(getattr(info, k)[:d], ref[k], rtol=1e-12, atol=1e-9), k
    assert list(info.groupN)[:d] == ref["groupN"].tolist()
    assert list(info.subN)[:d] == ref["subN"].tolist()
    assert list(info.nOvlRaw)[:d] == ref["nOvl"].tolist()
    assert (org == ref["origins"]).all()
    assert np.allclose(xms, ref["xMinSub"], rtol=1e-12, atol=1e-9)
    # the overlap shared by neighbours is the partition's overlap band
    for a in range(d):
        if nF[a] > 1:
            stride = int(round(info.stepProc[a] / info.xStep[a]))
            assert info.subN[a] - stride == info.nOvl[a]

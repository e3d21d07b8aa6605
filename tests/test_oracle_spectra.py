"""Pins the oracle's spectral setup (spectra_RF.f90:15-286, randomFieldND.f90:
260-266) against numbers fixed independently of its code: values derived by
hand from the reference's rules, the sign table written out, and the
normalisation of the gaussian spectral density."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_kmax_values():
    # spectra_RF.f90:33-37, 99 % spectral area per dimension count
    for d, v in ((1, 6.457), (2, 7.035), (3, 7.355)):
        p = O.make_plan(d, O.SHINOZUKA, O.FIXED, [10.0] * d)
        assert list(p.kMax)[:d] == [v] * d


def test_cfg1_lattice_numbers():
    # xRange 10: kDelta0 = 2pi/11 = 0.5712 -> kNStep = 1 + ceil(7.355/0.5712) = 14,
    # kDelta = 7.355/13, N_k = 14^3 * 4 = 10,976, ampMult = 2 sqrt(kDelta^3/(2pi)^3)
    p = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [10.0, 10.0, 10.0])
    assert list(p.kNStep) == [14, 14, 14]
    assert p.kN_unsigned == 2744 and p.kNTotal == 10976 and p.nSignBlocks == 4
    assert p.kDelta[0] == pytest.approx(0.565769, abs=1e-6)
    assert p.ampMult == pytest.approx(0.0540405, abs=1e-7)


def test_cfg2_isotropic_lattice_numbers():
    # normalised xRange 25.4: kDelta0 = 2pi/(1.1*25.4*sqrt(3)); kNStep = 1 + 5*ceil(...) = 286
    p = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4, 25.4, 25.4])
    assert p.kNTotal == 286
    assert p.isoStep == pytest.approx(0.0258070, abs=1e-7)
    k = O.kpoints(p).reshape(-1)
    assert k[0] == 0.0 and k[-1] == pytest.approx(7.355, rel=1e-15)
    assert np.allclose(np.diff(k), p.isoStep, rtol=1e-12)


def test_ksign_table():
    # (+,+,+), (+,+,-), (+,-,+), (+,-,-) and the 2-D pair (+,+), (+,-)
    assert O.ksign(3).tolist() == [[1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1]]
    assert O.ksign(2).tolist() == [[1, 1], [1, -1]]
    assert O.ksign(1).tolist() == [[1]]


def test_lattice_order_and_blocks():
    p = O.make_plan(3, O.SHINOZUKA, O.FIXED, [10.0] * 3)
    k = O.kpoints(p)
    dk = p.kDelta[0]
    n = 14
    # base block: dimension 1 fastest, k = (i-1) * kDelta
    j = 5 + n * 7 + n * n * 2
    assert np.allclose(k[j], [5 * dk, 7 * dk, 2 * dk], rtol=1e-15)
    # block b = base * kSign(b, :)
    for b, s in enumerate(O.ksign(3)):
        assert np.array_equal(k[b * 2744 + j], k[j] * s)


def test_ref_compat_block1_is_zero():
    # spectra_RF.f90:127 starts at block 2: block 1 is never written (A1)
    p = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [10.0] * 3)
    k = O.kpoints(p)
    assert not k[:2744].any()
    pf = O.make_plan(3, O.SHINOZUKA, O.FIXED, [10.0] * 3)
    assert np.array_equal(k[2744:], O.kpoints(pf)[2744:])
    # 1-D: the only block is block 1, so REF_COMPAT has no wavevectors at all
    p1 = O.make_plan(1, O.SHINOZUKA, O.REF_COMPAT, [10.0])
    assert not O.kpoints(p1).any()


def test_skvec_normalisation():
    # S(k) = exp(-|k|^2/(4 pi)) integrates to (2 pi)^d.  REF_COMPAT: the
    # reference's table (spectra_RF.f90:196-205), whose half-space lattice sum
    # counts the k_a = 0 planes twice (sign blocks) or whole (k_1 >= 0):
    # above (2 pi)^3.  FIXED: trapezoid weights 2^-(zero components) bring
    # the sum to (2 pi)^3 up to the kMax truncation (erfc(kMax/sqrt(4 pi))
    # per axis, ~1 %).
    for xr in (40.0, 12.0):
        pr = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [xr] * 3)
        pf = O.make_plan(3, O.SHINOZUKA, O.FIXED, [xr] * 3)
        k = O.kpoints(pf)
        sr, sf = O.skvec(pr, k), O.skvec(pf, k)
        g = np.exp(-(k ** 2).sum(1) / (4 * O.PI_REF))
        assert np.allclose(sr, g, rtol=1e-14)
        zeros = (k == 0).sum(1)
        assert np.array_equal(sf, sr * 0.5 ** zeros)
        vol = np.prod(list(pf.kDelta)) * 2
        trunc = math.erf(pf.kMax[0] / math.sqrt(4 * math.pi)) ** 3
        assert sf.sum() * vol / (2 * math.pi) ** 3 == pytest.approx(trunc, rel=5e-3)
        assert sr.sum() * vol / (2 * math.pi) ** 3 > 1.0 + 0.5 * pf.kDelta[0] / math.sqrt(4 * math.pi)


def test_randomization_box_and_density():
    rng = np.random.default_rng(3)
    p = O.make_plan(2, O.RANDOMIZATION, O.REF_COMPAT, [10.0, 20.0])
    kU = rng.random(p.nKUniform)
    k = O.kpoints(p, kU)
    base = kU.reshape(-1, 2) * np.array(list(p.kMax)[:2])
    assert not k[: p.kN_unsigned].any()                   # A1 again
    assert np.array_equal(k[p.kN_unsigned:], base * [1, -1])
    # ampMult 2 sqrt(1/(N_k (2pi)^d)) (randomFieldND.f90:262)
    assert p.ampMult == pytest.approx(2 * math.sqrt(1 / (p.kNTotal * (2 * O.PI_REF) ** 2)), rel=1e-15)
    # density sampling: k ~ N(0, 2 pi I)
    pd = O.make_plan(3, O.RANDOMIZATION, O.FIXED, [1.0] * 3, ksamp=O.KDENSITY, nKDensity=20000)
    kd = O.kpoints(pd, rng.random(pd.nKUniform))
    assert kd.mean() == pytest.approx(0.0, abs=0.03)
    assert kd.var() == pytest.approx(2 * math.pi, rel=0.02)
    assert pd.ampMult == pytest.approx(math.sqrt(2 / 20000), rel=1e-15)


def test_overlap_points_cfg4():
    # localization_RF.f90:746 with overlap 2, corrL 1, xStep 0.25 -> 7 points
    assert O.overlap_npoints([2.0] * 3, [1.0] * 3, [0.25] * 3).tolist() == [7, 7, 7]

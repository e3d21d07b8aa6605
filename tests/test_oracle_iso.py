"""Pins of the oracle's isotropic sum (randomFieldND.f90:368-514) by the
mathematics, independently of the formula the oracle restates.

* Known answers: one or two shells with draws chosen so that every angle is a
  round value, so the unit direction of each shell is known without the
  sin/cos formula (e.g. phi = pi/3, theta = 0 -> (sqrt3/2, 0, 1/2)); the
  expected field is written with those unit vectors.  A swapped sin/cos,
  theta taken as the polar angle, a component off by a sign or the wrong
  gamma law fails these.
* Correlation (FIXED): uniform directions make the field isotropic.  With
  phi ~ U(0, pi), theta ~ U(0, 2 pi) and the sin(phi) amplitude weight, the
  covariance at lag tau in ANY direction is
      3-D:  ampMult^2 sum_j S_j k_j^2 sinc(k_j tau) / pi
      2-D:  ampMult^2 sum_j S_j k_j   J0(k_j tau) / 2
  (E over the sphere of cos(k e.tau) is sin(k tau)/(k tau); over the circle
  J0(k tau)).  Monte-Carlo over realizations along x, y, z and a diagonal.
* REF_COMPAT 2-D (gamma = 2 pi U, theta = 2 pi psi): the variance at a point
  x is ampMult^2 sum_j S_j k_j E[gamma^2] E_U[cos^2(k_j e(U).x + 2 pi U)] with
  E[gamma^2] = 4 pi^2 / 3 and e(U) = (cos 4 pi^2 U, sin 4 pi^2 U), evaluated
  by quadrature over U.
"""
import math

import numpy as np
import pytest
from scipy.special import j0

from oracle import oracle as O

PI = O.PI_REF


def _shell_plan(nDim, mode, k, s):
    """A plan whose table is the given shells (radii k, densities s)."""
    p = O.make_plan(nDim, O.ISOTROPIC, mode, [10.0] * nDim)
    p.kNTotal = len(k)
    return p, np.asarray(k, np.float64).reshape(-1, 1), np.asarray(s, np.float64)


def test_isotropic_3d_known_answer_fixed():
    k, s = [2.0, 3.0], [0.5, 0.25]
    p, kr, sk = _shell_plan(3, O.FIXED, k, s)
    # draws per realization: phi(2), theta(2), gamma(2), psi(2)
    U = np.array([[1 / 3, 0.5,  0.0, 0.25,  0.75, 0.1,  0.125, 0.5]])
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.3, -0.7, 1.9]])
    got = O.isotropic(p, x, kr, sk, U)[0]
    # shell 0: phi = pi/3, theta = 0      -> e = (sqrt3/2, 0, 1/2), weight sin phi = sqrt3/2
    # shell 1: phi = pi/2, theta = pi/2   -> e = (0, 1, 0),         weight 1
    e0 = np.array([math.sqrt(3) / 2, 0.0, 0.5])
    e1 = np.array([0.0, 1.0, 0.0])
    g0 = math.sqrt(12.0) * 0.25            # gamma = sqrt12 (U - 1/2)
    g1 = math.sqrt(12.0) * (-0.4)
    a0 = math.sqrt(0.5 * (math.sqrt(3) / 2) * 4.0) * g0
    a1 = math.sqrt(0.25 * 1.0 * 9.0) * g1
    ref = p.ampMult * (a0 * np.cos(2.0 * x @ e0 + math.pi / 4) + a1 * np.cos(3.0 * x @ e1 + math.pi))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13)


def test_isotropic_3d_known_answer_ref_compat():
    # theta = 2 pi * psi with psi = 2 pi U_psi already scaled (A4);
    # gamma = sqrt(12.0) in single precision (randomFieldND.f90:468)
    k, s = [1.5], [0.8]
    p, kr, sk = _shell_plan(3, O.REF_COMPAT, k, s)
    u_psi = 1.0 / (8.0 * PI)               # psi = 1/4 rad, theta = pi/2
    U = np.array([[0.5, 0.123, 0.75, u_psi]])   # the theta draw (0.123) is not used
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [2.0, -1.0, 0.5]])
    got = O.isotropic(p, x, kr, sk, U)[0]
    g = 3.464101552963257 * 0.25             # sqrt(12) rounded to single precision
    e = np.array([0.0, 1.0, 0.0])            # phi = pi/2, theta = pi/2
    ref = p.ampMult * math.sqrt(0.8 * 1.5 ** 2) * g * np.cos(1.5 * x @ e + 0.25)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13)


def test_isotropic_2d_known_answer_fixed():
    k, s = [2.5, 4.0], [0.6, 0.3]
    p, kr, sk = _shell_plan(2, O.FIXED, k, s)
    # draws per realization: psi(2), theta(2), gamma(2)
    U = np.array([[0.25, 0.0,  0.25, 0.5,  0.9, 0.3]])
    x = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.4, -1.3]])
    got = O.isotropic(p, x, kr, sk, U)[0]
    # shell 0: theta = pi/2 -> e = (0, 1); shell 1: theta = pi -> e = (-1, 0)
    g0, g1 = math.sqrt(12.0) * 0.4, math.sqrt(12.0) * (-0.2)
    ref = p.ampMult * (math.sqrt(0.6 * 2.5) * g0 * np.cos(2.5 * x[:, 1] + math.pi / 2)
                       + math.sqrt(0.3 * 4.0) * g1 * np.cos(-4.0 * x[:, 0]))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13)


def test_isotropic_2d_known_answer_ref_compat():
    # 2-D reference: psi = 2 pi U, theta = 2 pi psi, gamma = 2 pi U_gamma
    # (randomFieldND.f90:427-429)
    k, s = [3.0], [0.7]
    p, kr, sk = _shell_plan(2, O.REF_COMPAT, k, s)
    U = np.array([[1.0 / (8.0 * PI), 0.9, 0.25]])     # psi = 1/4, theta = pi/2, gamma = pi/2
    x = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [-0.6, 2.2]])
    got = O.isotropic(p, x, kr, sk, U)[0]
    ref = p.ampMult * math.sqrt(0.7 * 3.0) * (math.pi / 2) * np.cos(3.0 * x[:, 1] + 0.25)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13)


def _mc_cov(nDim, mode, dirs, lags, nmc, seed, base_n=24, spread=40.0):
    """Sample covariance of the oracle field between base points and base
    points shifted by lag * direction, averaged over the base points."""
    p = O.make_plan(nDim, O.ISOTROPIC, mode, [10.0] * nDim)
    kr = O.kpoints(p)
    sk = O.skvec(p, kr)
    rng = np.random.default_rng(seed)
    base = rng.random((base_n, nDim)) * spread
    pts = [base]
    for d in dirs:
        d = np.asarray(d, float) / np.linalg.norm(d)
        for t in lags:
            pts.append(base + t * d)
    x = np.concatenate(pts)
    U = rng.random((nmc, O.iso_nvar(nDim) * p.kNTotal))
    u = O.isotropic(p, x, kr, sk, U, threads=8).reshape(nmc, len(pts), base_n)
    cov = (u[:, :1, :] * u[:, 1:, :]).mean(axis=(0, 2)).reshape(len(dirs), len(lags))
    var = (u[:, 0, :] ** 2).mean()
    return p, kr[:, 0], sk, cov, var


def test_isotropic_3d_correlation_is_isotropic_sinc():
    lags = np.array([0.4, 0.9, 1.5])
    dirs = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
    p, k, s, cov, var = _mc_cov(3, O.FIXED, dirs, lags, 6000, 5)
    w = s * k ** 2
    expect_var = p.ampMult ** 2 * w.sum() / math.pi
    rho = np.array([(w * np.sinc(k * t / math.pi)).sum() / w.sum() for t in lags])
    assert var == pytest.approx(expect_var, rel=0.05)
    # sampling error of the correlation ~ 1/sqrt(6000 * 24) ~ 0.003
    for i, d in enumerate(dirs):
        assert np.allclose(cov[i] / expect_var, rho, atol=0.025), (d, cov[i] / expect_var, rho)
    # the closed form is not flat: the lags probe distinct values
    assert rho[0] - rho[-1] > 0.3


def test_isotropic_2d_correlation_is_isotropic_j0():
    lags = np.array([0.4, 0.9, 1.5])
    dirs = [(1, 0), (0, 1), (1, 1), (1, -2)]
    p, k, s, cov, var = _mc_cov(2, O.FIXED, dirs, lags, 6000, 6)
    w = s * k
    expect_var = p.ampMult ** 2 * w.sum() / 2.0          # E[gamma^2] = 1, E[cos^2] = 1/2
    assert expect_var == pytest.approx(1.0, abs=0.03)
    rho = np.array([(w * j0(k * t)).sum() / w.sum() for t in lags])
    assert var == pytest.approx(expect_var, rel=0.05)
    for i, d in enumerate(dirs):
        assert np.allclose(cov[i] / expect_var, rho, atol=0.025), (d, cov[i] / expect_var, rho)
    assert rho[0] - rho[-1] > 0.3


def test_isotropic_2d_ref_compat_variance():
    p = O.make_plan(2, O.ISOTROPIC, O.REF_COMPAT, [10.0, 10.0])
    kr = O.kpoints(p)
    sk = O.skvec(p, kr)
    k = kr[:, 0]
    x = np.array([[0.0, 0.0], [3.0, 4.0], [-7.5, 2.0], [12.0, -9.0]])
    nmc = 20000
    U = np.random.default_rng(9).random((nmc, 3 * p.kNTotal))
    u = O.isotropic(p, x, kr, sk, U, threads=8)
    # E_U over 2^17 midpoints of [0,1): psi = 2 pi U, theta = 2 pi psi
    uq = (np.arange(1 << 17) + 0.5) / (1 << 17)
    psi = 2 * PI * uq
    th = 2 * PI * psi
    g2 = (2 * math.pi) ** 2 / 3.0
    for i, xi in enumerate(x):
        ecos2 = np.array([np.mean(np.cos(kj * (np.cos(th) * xi[0] + np.sin(th) * xi[1]) + psi) ** 2) for kj in k])
        expect = p.ampMult ** 2 * g2 * (sk * k * ecos2).sum()
        # sampling error of a variance over 20000 near-Gaussian values ~ 1 %
        assert u[:, i].var() == pytest.approx(expect, rel=0.05), (xi, u[:, i].var(), expect)
    # at the origin the closed form is exact: E[cos^2 psi] = 1/2
    assert p.ampMult ** 2 * g2 * (sk * k).sum() / 2 == pytest.approx(u[:, 0].var(), rel=0.05)


def test_fft_plan_hand_derived():
    # method 4 lattice (spectra_RF.f90:135-139): kDelta = 2 pi / (1.1 xRange),
    # kMax = (kNStep - 1) kDelta, kNStep = grid points per axis.
    # 128 x 64 x 32 grid of normalised extent (25.4, 12.6, 6.2), chosen so that
    # (n - 1) / xRange = 5 on every axis:
    #   kDelta = 2 pi / 27.94, 2 pi / 13.86, 2 pi / 6.82 = 0.2248814, 0.4533323, 0.9212882
    #   kMax   = 5 * 2 pi / 1.1 = 28.559933 on every axis
    p = O.fft_plan(3, [25.4, 12.6, 6.2], [128, 64, 32])
    assert list(p.kNStep) == [128, 64, 32]
    assert p.kNTotal == 128 * 64 * 32
    assert list(p.kDelta) == pytest.approx([0.2248814, 0.4533323, 0.9212882], abs=1e-7)
    assert list(p.kMax) == pytest.approx([28.559933] * 3, abs=1e-6)
    # 2-D, 10 x 7 grid over (4.5, 3.0): (n - 1) / xRange = 2 ->
    #   kDelta = 2 pi / 4.95, 2 pi / 3.3 = 1.2693304, 1.9039955 ; kMax = 4 pi / 1.1 = 11.423973
    p2 = O.fft_plan(2, [4.5, 3.0], [10, 7])
    assert p2.kNTotal == 70
    assert list(p2.kDelta)[:2] == pytest.approx([1.2693304, 1.9039955], abs=1e-7)
    assert list(p2.kMax)[:2] == pytest.approx([11.423973, 11.423973], abs=1e-6)
    k = O.kpoints(p2)
    # dimension 1 fastest from 0: k[j] = (j mod 10, j div 10) * kDelta
    assert list(k[13]) == pytest.approx([3 * 1.2693304, 1 * 1.9039955], abs=1e-6)

"""Pins the oracle's sums, marginal map, partition of unity, blend and moments
with checks a plausible mistake would fail: single-term closed forms, the
translation identity u(x + delta; phi) = u(x; phi + k.delta) (catches a wrong
sign or transposed index in k.x), the reference's sqrt schedule on a single
wavevector, Monte-Carlo variance and correlation against the target
R(tau) = exp(-pi tau^2), and the partition-of-unity identities."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _plan_with_table(nDim, k, mode=O.FIXED):
    """A Shinozuka plan whose table is replaced by k (N_k x nDim)."""
    p = O.make_plan(nDim, O.SHINOZUKA, mode, [10.0] * nDim)
    p.kNTotal = k.shape[0]
    return p


def test_single_wavevector_closed_form():
    k = np.array([[0.7, -0.2, 0.3]])
    p = _plan_with_table(3, k)
    x = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [-4.0, 0.5, 2.0]])
    U = np.array([[0.125]])
    out = O.shinozuka(p, x, k, np.array([0.36]), U)
    phi = 2 * O.PI_REF * 0.125
    expect = p.ampMult * 0.6 * np.cos(x @ k[0] + phi)
    assert np.allclose(out[0], expect, rtol=1e-14, atol=1e-16)


def test_translation_identity():
    rng = np.random.default_rng(1)
    k = rng.normal(size=(37, 3))
    p = _plan_with_table(3, k)
    x = rng.random((50, 3)) * 5
    delta = np.array([0.3, -1.1, 0.7])
    U = rng.random((2, 37))
    sk = rng.random(37)
    a = O.shinozuka(p, x + delta, k, sk, U)
    U2 = U + (k @ delta) / (2 * O.PI_REF)
    b = O.shinozuka(p, x, k, sk, U2)
    assert np.allclose(a, b, rtol=0, atol=1e-11)


def test_ref_compat_sqrt_schedule():
    # randomFieldND.f90:312: realization n on a single slab uses S^(2^-n);
    # with xSlab = 10 on 30 points (3 slabs) pass c = 3(n-1) + m
    k = np.array([[0.4, 0.1, -0.3]])
    p = _plan_with_table(3, k, O.REF_COMPAT)
    x = np.random.default_rng(2).random((30, 3))
    U = np.array([[0.2], [0.7], [0.1]])
    S = 0.01
    out = O.shinozuka(p, x, k, np.array([S]), U)
    for n in range(3):
        c = np.cos(x @ k[0] + 2 * O.PI_REF * U[n, 0])
        assert np.allclose(out[n], p.ampMult * S ** (0.5 ** (n + 1)) * c, rtol=1e-13)
    out3 = O.shinozuka(p, x, k, np.array([S]), U, xSlab=10)
    for n in range(3):
        for m in range(3):
            sl = slice(10 * m, 10 * m + 10)
            c = np.cos(x[sl] @ k[0] + 2 * O.PI_REF * U[n, 0])
            e = 0.5 ** (3 * n + m + 1)
            assert np.allclose(out3[n, sl], p.ampMult * S ** e * c, rtol=1e-12)


def test_calculate_mask_and_draw_independence():
    k = np.array([[0.4, 0.1, -0.3], [0.2, 0.2, 0.2]])
    p = _plan_with_table(3, k, O.REF_COMPAT)
    x = np.random.default_rng(2).random((8, 3))
    U = np.random.default_rng(3).random((3, 2))
    full = O.shinozuka(p, x, k, np.array([0.3, 0.5]), U)
    masked = O.shinozuka(p, x, k, np.array([0.3, 0.5]), U, calc=np.array([1, 0, 1]))
    assert not masked[1].any()
    assert np.array_equal(masked[0], full[0])
    # realization 3 is the second pass: same exponent as full realization 2
    assert np.allclose(masked[2], full[1] * 0 + O.shinozuka(p, x, k, np.array([0.3, 0.5]), U[[0, 2]])[1])


def test_fixed_variance_closed_form():
    # Var u = ampMult^2 / 2 * sum S over phases, for every point
    p = O.make_plan(3, O.SHINOZUKA, O.FIXED, [6.0] * 3)
    k = O.kpoints(p)
    s = O.skvec(p, k)
    rng = np.random.default_rng(5)
    x = rng.random((40, 3)) * 6
    U = rng.random((400, p.kNTotal))
    u = O.shinozuka(p, x, k, s, U)
    expect = p.ampMult ** 2 / 2 * s.sum()
    assert u.var() == pytest.approx(expect, rel=0.05)
    assert abs(u.mean()) < 4 * math.sqrt(expect / u.size) * 10


def test_density_correlation_target():
    # k ~ N(0, 2 pi I), A = sqrt(2/N_k): E_phi[u(x)u(x+tau)] = mean_j cos(k_j.tau)
    # -> exp(-pi |tau|^2) up to O(1/sqrt(N_k))
    p = O.make_plan(3, O.RANDOMIZATION, O.FIXED, [1.0] * 3, ksamp=O.KDENSITY, nKDensity=4096)
    rng = np.random.default_rng(7)
    k = O.kpoints(p, rng.random(p.nKUniform))
    s = O.skvec(p, k)
    base = rng.random((60, 3)) * 20
    for tau in (0.0, 0.3, 0.6):
        x = np.concatenate([base, base + [tau, 0, 0]])
        U = rng.random((300, p.kNTotal))
        u = O.shinozuka(p, x, k, s, U)
        c = (u[:, :60] * u[:, 60:]).mean()
        assert c == pytest.approx(math.exp(-math.pi * tau * tau), abs=0.05)


def test_isotropic_variance_closed_form():
    # FIXED scale: Var = ampMult^2 * sum_j S_j k_j^2 E[sin phi] E[gamma^2] / 2
    #            = ampMult^2 * sum S k^2 / pi  (~1 on a fine lattice)
    p = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    k = O.kpoints(p)
    s = O.skvec(p, k)
    rng = np.random.default_rng(11)
    x = rng.random((30, 3)) * 25
    U = rng.random((300, 4 * p.kNTotal))
    u = O.isotropic(p, x, k, s, U)
    expect = p.ampMult ** 2 * (s * k[:, 0] ** 2).sum() / math.pi
    assert expect == pytest.approx(0.967, abs=0.01)
    assert u.var() == pytest.approx(expect, rel=0.06)
    # REF_COMPAT scale is 1/(4 pi^2) of that (randomFieldND.f90:500, A6)
    pr = O.make_plan(3, O.ISOTROPIC, O.REF_COMPAT, [25.4] * 3)
    assert pr.ampMult ** 2 * (2 * math.pi) ** 2 == pytest.approx(p.ampMult ** 2, rel=1e-12)


def test_isotropic_ref_compat_azimuth():
    # A4: theta = 2 pi * psi with psi already scaled -> the theta draw is unused
    p = O.make_plan(3, O.ISOTROPIC, O.REF_COMPAT, [25.4] * 3)
    k = O.kpoints(p)
    s = O.skvec(p, k)
    x = np.random.default_rng(1).random((10, 3))
    U = np.random.default_rng(2).random((1, 4 * p.kNTotal))
    U2 = U.copy()
    U2[0, p.kNTotal: 2 * p.kNTotal] = 0.5                 # theta draws
    assert np.array_equal(O.isotropic(p, x, k, s, U), O.isotropic(p, x, k, s, U2))
    pf = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    assert not np.array_equal(O.isotropic(pf, x, k, s, U), O.isotropic(pf, x, k, s, U2))


def test_marginal_closed_forms():
    g = np.random.default_rng(4).standard_normal(400000)
    y = O.marginal(2, 1.0, 0.25, g)
    assert y.mean() == pytest.approx(1.0, abs=0.005)
    assert y.var() == pytest.approx(0.25, rel=0.03)
    s2 = math.log(1 + 0.25)
    assert np.allclose(y, np.exp(g * math.sqrt(s2) + math.log(1.0) - s2 / 2), rtol=1e-15)
    assert np.allclose(O.marginal(1, 3.0, 4.0, g[:10]), g[:10] * 2 + 3, rtol=1e-15)
    with pytest.raises(ValueError):
        O.marginal(2, 0.0, 1.0, g[:4])


def test_partition_of_unity_identities():
    n = 7
    s = np.arange(1, n + 1)
    # REF_COMPAT faces of one overlap: w_A + w_B = 1 + [cos(pi s/(n+2)) - cos(pi (s+1)/(n+2))]/2
    up = O.unity_partition([30], [n], 0b10, O.REF_COMPAT)[30 - n:]
    low = O.unity_partition([30], [n], 0b01, O.REF_COMPAT)[:n]
    expect = 1 + (np.cos(O.PI_REF * s / (n + 2)) - np.cos(O.PI_REF * (s + 1) / (n + 2))) / 2
    assert np.allclose(up + low, expect, rtol=1e-15)
    # FIXED: squared weights of the two faces sum to one
    upf = O.unity_partition([30], [n], 0b10, O.FIXED)[30 - n:]
    lowf = O.unity_partition([30], [n], 0b01, O.FIXED)[:n]
    assert np.allclose(upf ** 2 + lowf ** 2, 1.0, rtol=1e-15)
    # separable product in 3-D
    pu = O.unity_partition([9, 10, 11], [3, 3, 3], 0b111111, O.FIXED).reshape(11, 10, 9)
    f = [O.unity_partition([m], [3], 0b11, O.FIXED) for m in (9, 10, 11)]
    assert np.allclose(pu, f[2][:, None, None] * f[1][None, :, None] * f[0][None, None, :], rtol=1e-15)


def test_blend_fixed_preserves_variance():
    # two 1-D subdomains overlapping by 5 points, independent unit fields
    rng = np.random.default_rng(8)
    Nmc = 4000
    fa = rng.standard_normal((Nmc, 20))
    fb = rng.standard_normal((Nmc, 20))
    g = O.blend([35], Nmc, [[0], [15]], [[20], [20]], [5], [0b10, 0b01], O.FIXED, [fa, fb])
    assert np.allclose(g.var(axis=0), 1.0, atol=0.08)
    # outside the overlap the field is copied
    assert np.array_equal(g[:, :15], fa[:, :15])
    gr = O.blend([35], 1, [[0], [15]], [[20], [20]], [5], [0b11, 0b11], O.REF_COMPAT, [fa[:1], fb[:1]])
    # REF_COMPAT: outer faces restored by the reverse taper, interior linear
    assert np.allclose(gr[0, :10], fa[0, :10], rtol=1e-12)


def test_normalize_and_stats():
    f = np.random.default_rng(9).normal(3.0, 2.0, size=(3, 5000))
    n = O.normalize(f)
    avg, sd, ga, gs = O.stats(n)
    assert np.allclose(avg, 0, atol=1e-12) and np.allclose(sd, 1, rtol=1e-12)
    a2, s2, _, _ = O.stats(f)
    assert np.allclose(a2, f.mean(1), rtol=1e-12) and np.allclose(s2, f.std(1), rtol=1e-10)


# ---------------------------------------------------------- directional correlation
def _ref_fft_steps(u, nstep, avg, std):
    """statistics_RF.f90:262-409 step by step with numpy's FFT: line FFTs
    along each axis, divide by n, times the conjugate, mean over the other
    axes, unnormalised inverse FFT, real part, first n/2 lags."""
    v = ((u - avg) / std).reshape(tuple(reversed(nstep)))       # [z][y][x]
    out = []
    nd = len(nstep)
    for a in range(nd):
        ax = nd - 1 - a
        F = np.fft.fft(v, axis=ax) / nstep[a]
        P = (F * np.conj(F))
        other = tuple(i for i in range(nd) if i != ax)
        Pm = P.mean(axis=other) if other else P
        r = np.real(np.fft.ifft(Pm) * nstep[a])
        out.append(r[: nstep[a] // 2])
    return out


@pytest.mark.parametrize("nstep", [(16,), (12, 9), (10, 7, 6), (5, 8, 11)])
def test_corr_matches_fft_statement(nstep):
    rng = np.random.default_rng(len(nstep) + nstep[0])
    u = rng.normal(0.3, 1.7, int(np.prod(nstep)))
    avg, std = float(u.mean()), float(u.std())
    got = O.corr(u, nstep, avg, std)
    ref = _ref_fft_steps(u, nstep, avg, std)
    for a in range(len(nstep)):
        assert np.allclose(got[a], ref[a], rtol=1e-11, atol=1e-12), a
    # lag 0 of a normalised field is its variance, 1
    assert all(abs(r[0] - 1.0) < 1e-12 for r in got)


def test_corrL_trapezoid_of_gaussian_correlation():
    """2 * trapezoid of R(tau) = exp(-pi tau^2 / L^2) over [0, (M-1) dx]
    approaches the integral L; a plain sum would be off by R(0) dx."""
    L, dx, M = 1.3, 0.01, 600
    r = np.exp(-np.pi * (np.arange(M) * dx) ** 2 / L ** 2)
    got = O.corrL([r], [2 * M], [dx])[0]
    assert abs(got - L) < 1e-6
    assert abs(2 * r.sum() * dx - L) > 5e-3


# ---------------------------------------------------------- interpolation to mesh nodes
@pytest.mark.parametrize("nDim", [2, 3])
def test_interp_reproduces_multilinear_fields(nDim):
    """Multilinear interpolation is exact on fields a + sum b_d x_d + c prod x_d
    (sample_RF.f90:1018-1086), at grid nodes, on the upper border (the
    floor(c) == n rule) and flags nodes outside the grid."""
    rng = np.random.default_rng(nDim)
    nstep = (7, 5, 4)[:nDim]
    xMin = (-1.0, 2.0, 0.5)[:nDim]
    xStep = (0.5, 0.25, 1.0)[:nDim]
    axes = [xMin[a] + xStep[a] * np.arange(nstep[a]) for a in range(nDim)]
    mesh = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([m.transpose().reshape(-1) for m in mesh], axis=1)   # x fastest

    def f(p, k):
        return 0.3 * k + p @ np.arange(1, nDim + 1) + 0.7 * k * np.prod(p, axis=1)
    grid = np.stack([f(pts, k) for k in range(3)])
    hi = np.array([xMin[a] + xStep[a] * (nstep[a] - 1) for a in range(nDim)])
    x = rng.uniform(np.array(xMin), hi, (500, nDim))
    x[0] = hi                      # upper corner
    x[1] = np.array(xMin)          # lower corner
    x[2] = pts[7]                  # a grid node
    out, bad = O.interp(grid, nstep, xMin, xStep, x)
    assert bad == 0
    ref = np.stack([f(x, k) for k in range(3)])
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)
    assert np.array_equal(out[:, 2], grid[:, 7])
    # below the grid, and more than one cell above it, are outside; the
    # border rule keeps [n, n+1) in the last cell (the reference's weights
    # 1 - |c - corner| there are not an extrapolation; only the count is
    # checked)
    xo = np.array([hi + 1.5 * np.array(xStep), np.array(xMin) - 0.1])
    out2, bad2 = O.interp(grid, nstep, xMin, xStep, xo)
    assert bad2 == 2 and np.isnan(out2).all()
    xe = hi + 0.5 * np.array(xStep)
    out3, bad3 = O.interp(grid, nstep, xMin, xStep, xe[None, :])
    assert bad3 == 0 and np.isfinite(out3).all()

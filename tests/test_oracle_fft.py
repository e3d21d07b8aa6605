"""Method 4 (gen_Std_Gauss_FFT, randomFieldND.f90:670-735) in the oracle,
pinned against scipy's DCT-III (type 3, unnormalised: y_i = x_0 + 2 sum_{j>=1}
x_j cos(pi j (2i+1) / (2n)), the transform FFTW calls REDFT01) and against
draw-order known answers: the kN - 1 uniforms between gammaK and phiK are
never read, gammaK = 1/2 gives a zero field, phiK = 0 gives cos = 1."""
import numpy as np
import pytest
import scipy.fft

from oracle import oracle as O


def _dct3(a, shape):
    # x fastest in memory: the array is (n_d, ..., n_1) in C order
    return scipy.fft.dctn(a.reshape(shape[::-1]), type=3).reshape(-1)


def _scale(xs):
    return (2.0 if len(xs) == 3 else 1.0) * np.sqrt(np.prod(xs))


@pytest.mark.parametrize("gn,xs", [((9, 6), (0.3, 0.5)), ((5, 7, 4), (0.25, 0.2, 0.4)), ((16, 16, 16), (0.1,) * 3)])
def test_fft_oracle_matches_scipy(gn, xs):
    nDim = len(gn)
    xr = [(n - 1) * s for n, s in zip(gn, xs)]
    po = O.fft_plan(nDim, xr, gn)
    kN = po.kNTotal
    S = O.skvec(po, O.kpoints(po))
    rng = np.random.default_rng(kN)
    Nmc = 3
    u = rng.random((Nmc, 3 * kN - 1))
    got = O.fft(po, S, u, xs, Nmc)
    for r in range(Nmc):
        g, ph = u[r, :kN], u[r, 2 * kN - 1:]
        coef = (g - 0.5) * np.sqrt(S) * np.cos(2.0 * O.PI_REF * ph)
        ref = _dct3(coef, gn) * _scale(xs)
        assert np.allclose(got[r], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_fft_oracle_draw_order_known_answers():
    gn, xs = (6, 5, 4), (0.5, 0.4, 0.3)
    po = O.fft_plan(3, [(n - 1) * s for n, s in zip(gn, xs)], gn)
    kN = po.kNTotal
    S = O.skvec(po, O.kpoints(po))
    u = np.full((4, 3 * kN - 1), np.nan)
    u[0, :kN], u[0, 2 * kN - 1:] = 1.0, 0.0          # coefficient 0.5 sqrt(S)
    u[1, :kN], u[1, 2 * kN - 1:] = 0.5, 0.3          # gammaK - 0.5 = 0
    u[2, :kN], u[2, 2 * kN - 1:] = 0.0, 0.5          # cos(2 PI 0.5) = -1 (PI = 3.1415926535898)
    u[3, :kN], u[3, 2 * kN - 1:] = 1.0, 0.0
    got = O.fft(po, S, u, xs, 4, calc=[1, 1, 1, 0])
    ref = _dct3(0.5 * np.sqrt(S), gn) * _scale(xs)
    assert np.isfinite(got).all()                     # the middle draws are never read
    assert np.allclose(got[0], ref, rtol=1e-12, atol=1e-14)
    assert not got[1].any()
    assert np.allclose(got[2], ref, rtol=1e-9, atol=1e-12)
    assert not got[3].any()                           # not calculated


def test_fft_oracle_single_mode_closed_form():
    # only coefficient (j1, j2) = (2, 1) nonzero: the field is
    # 4 c cos(pi 2 (2i1+1) / (2 n1)) cos(pi (2i2+1) / (2 n2)) times the scale
    gn, xs = (8, 6), (0.5, 0.25)
    po = O.fft_plan(2, [3.5, 1.25], gn)
    kN = po.kNTotal
    S = np.zeros(kN)
    j = 2 + gn[0] * 1
    S[j] = 4.0
    u = np.zeros((1, 3 * kN - 1))
    u[0, :kN] = 1.0                                  # gamma - 0.5 = 0.5, c = 0.5 * 2 = 1
    got = O.fft(po, S, u, xs, 1)[0].reshape(gn[1], gn[0])
    i1, i2 = np.arange(gn[0]), np.arange(gn[1])
    ref = 4.0 * np.outer(np.cos(np.pi * (2 * i2 + 1) / (2 * gn[1])), np.cos(np.pi * 2 * (2 * i1 + 1) / (2 * gn[0])))
    assert np.allclose(got, ref * np.sqrt(xs[0] * xs[1]), rtol=1e-13, atol=1e-14)

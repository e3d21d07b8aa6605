"""C-ABI checks that need no GPU: every symbol of include/rf_capi.h is
exported, host-side spectrum tables equal the oracle's bit for bit, argument
errors map to the documented status codes, and the random streams match their
known answers (the compiler's random_number for the reference seeds, and the
published Philox4x32-10 vectors)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def test_header_symbols_exported():
    hdr = (ROOT / "include" / "rf_capi.h").read_text()
    decl = set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", hdr))
    assert decl == set(R.EXPORTS), decl ^ set(R.EXPORTS)
    L = R.lib()
    for name in decl:
        assert hasattr(L, name), name
    assert L.rf_version() == 1


def test_fortran_facade_loads():
    F = R.flib()
    assert F.rf_fortran_seed_size() == 1      # amdflang: seed size 1


@pytest.mark.parametrize("nDim,method,mode,xr", [
    (3, R.SHINOZUKA, R.REF_COMPAT, (10, 10, 10)),
    (3, R.SHINOZUKA, R.FIXED, (7.5, 12.0, 3.3)),
    (2, R.SHINOZUKA, R.FIXED, (5.0, 9.0)),
    (1, R.SHINOZUKA, R.REF_COMPAT, (4.0,)),
    (3, R.ISOTROPIC, R.FIXED, (12.7, 12.7, 12.7)),
    (2, R.ISOTROPIC, R.REF_COMPAT, (3.0, 8.0)),
])
def test_spectrum_tables_match_oracle(nDim, method, mode, xr):
    corrL = (0.5, 1.0, 2.0)[:nDim]
    p = R.make_params(nDim=nDim, method=method, mode=mode, Nmc=2, corrL=corrL, xRange=xr)
    info = R.plan_spectrum(p)
    po = O.make_plan(nDim, method, mode, [xr[i] / corrL[i] for i in range(nDim)])
    assert info.kNTotal == po.kNTotal and info.ampMult == po.ampMult
    assert list(info.kDelta)[:nDim] == list(po.kDelta)[:nDim]
    k, s = R.spectrum_tables(p)
    ko = O.kpoints(po)
    assert np.array_equal(k, ko) and np.array_equal(s, O.skvec(po, ko))


@pytest.mark.parametrize("nDim,xr,gn", [(2, (5.0, 9.0), (11, 19)), (3, (2.0, 3.0, 1.5), (5, 7, 4))])
def test_fft_tables_match_oracle(nDim, xr, gn):
    corrL = (0.5, 1.5, 2.0)[:nDim]
    p = R.make_params(nDim=nDim, method=R.FFT, mode=R.FIXED, Nmc=2, corrL=corrL, xRange=xr, gridN=gn)
    info = R.plan_spectrum(p)
    po = O.fft_plan(nDim, [xr[i] / corrL[i] for i in range(nDim)], gn)
    assert info.kNTotal == po.kNTotal == int(np.prod(gn)) and info.nSignBlocks == 1
    assert info.nPhaseUniform == 3 * info.kNTotal - 1
    assert list(info.kDelta)[:nDim] == list(po.kDelta)[:nDim]
    assert list(info.kNStep)[:nDim] == list(gn)
    k, s = R.spectrum_tables(p)
    ko = O.kpoints(po)
    assert np.array_equal(k, ko) and np.array_equal(s, O.skvec(po, ko))
    # lattice from 0 in steps kDelta = 2 pi / (1.1 xRange), x fastest
    assert np.allclose(ko[1, 0], 2 * np.pi / (1.1 * xr[0] / corrL[0]), rtol=1e-12) and ko[0].tolist() == [0.0] * nDim
    assert ko[gn[0], 1] > 0 and ko[gn[0], 0] == 0.0


def test_fft_draws_continue_one_stream():
    # gen_Std_Gauss_FFT draws 3 kN - 1 uniforms per realization from one stream
    p = R.make_params(nDim=2, method=R.FFT, Nmc=2, xRange=(3, 3), gridN=(4, 5))
    _, ph = R.reference_draws(p, [90])
    p2 = R.make_params(nDim=2, method=R.SHINOZUKA, Nmc=1, xRange=(40, 40))
    _, flat = R.reference_draws(p2, [90])
    assert ph.shape == (2, 59)
    assert np.array_equal(ph.reshape(-1), flat[0, :118])


def test_randomization_tables_match_oracle():
    rng = np.random.default_rng(0)
    for ks, nk in ((R.KSAMP_BOX, 0), (R.KSAMP_DENSITY, 333)):
        for mode in (R.REF_COMPAT, R.FIXED):
            p = R.make_params(nDim=3, method=R.RANDOMIZATION, mode=mode, Nmc=1, kSampling=ks,
                              nKDensity=nk, corrL=(2, 2, 0.5), xRange=(6, 6, 4))
            info = R.plan_spectrum(p)
            kU = rng.random(info.nKUniform)
            po = O.make_plan(3, R.RANDOMIZATION, mode, [3, 3, 8], ksamp=ks, nKDensity=nk)
            assert info.ampMult == po.ampMult and info.kNTotal == po.kNTotal
            k, s = R.spectrum_tables(p, kU)
            ko = O.kpoints(po, kU)
            assert np.array_equal(k, ko) and np.array_equal(s, O.skvec(po, ko))


@pytest.mark.parametrize("kw,code", [
    (dict(Nmc=0), R.RF_E_ARG),
    (dict(nDim=4), R.RF_E_ARG),
    (dict(method=4), R.RF_E_ARG),                       # gridN missing
    (dict(method=4, nDim=1, xRange=(1.0,), gridN=(8,)), R.RF_E_UNSUPPORTED),
    (dict(method=5), R.RF_E_ARG),
    (dict(corrL=(0.0, 1.0, 1.0)), R.RF_E_ARG),
    (dict(margiFirst=R.FOM_LOGNORMAL, fieldAvg=0.0), R.RF_E_LOGNORMAL),
    (dict(method=R.ISOTROPIC, nDim=1, xRange=(1.0,)), R.RF_E_UNSUPPORTED),
    (dict(method=R.ISOTROPIC, form=R.FORM_MFMA_F16), R.RF_E_UNSUPPORTED),
    (dict(method=R.RANDOMIZATION, kSampling=R.KSAMP_DENSITY, nKDensity=0), R.RF_E_ARG),
    (dict(isoShared=1), R.RF_E_ARG),
])
def test_error_codes(kw, code):
    base = dict(nDim=3, method=R.SHINOZUKA, Nmc=2, xRange=(1.0, 1.0, 1.0))
    base.update(kw)
    p = R.make_params(**base)
    with pytest.raises(R.RFError) as e:
        R.plan_spectrum(p)
    assert e.value.code == code


def test_reference_rng_known_answers():
    # random_seed(PUT=[90]) (seedStart 1: 72 + 18) then random_number: the
    # stream of the compiler runtime (amdflang, ROCm 7.2)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, Nmc=2, xRange=(10, 10, 10))
    assert R.reference_seed(1).tolist() == [90]
    _, ph = R.reference_draws(p, R.reference_seed(1))
    kat = [4.1911399906169322e-08, 6.5292042272326944e-01, 2.4501493677125608e-01, 7.2079464651480463e-02]
    assert ph[0, :4].tolist() == kat
    # realization 2 continues the same stream (randomFieldND.f90:278)
    _, ph1 = R.reference_draws(R.make_params(nDim=3, method=R.SHINOZUKA, Nmc=1, xRange=(20, 20, 20)),
                               R.reference_seed(1))
    assert ph1[0, :4].tolist() == kat


def test_reference_draw_order_methods():
    # method 3 draws its wavevector uniforms before the phases (spectra_RF.f90:120)
    p3 = R.make_params(nDim=2, method=R.RANDOMIZATION, Nmc=1, xRange=(3, 3))
    kU, ph = R.reference_draws(p3, [90])
    info = R.plan_spectrum(p3)
    p2 = R.make_params(nDim=2, method=R.SHINOZUKA, Nmc=3, xRange=(3, 3))
    _, ph2 = R.reference_draws(p2, [90])
    flat = ph2.reshape(-1)
    n = info.nKUniform
    assert np.array_equal(kU, flat[:n]) and np.array_equal(ph[0], flat[n:n + info.nPhaseUniform])
    # isotropic: no draws for realizations that are not calculated (:421, :457)
    pi = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, Nmc=3, xRange=(2, 2, 2))
    _, a = R.reference_draws(pi, [90], calc=[1, 0, 1])
    _, b = R.reference_draws(R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, Nmc=2, xRange=(2, 2, 2)), [90])
    assert not a[1].any() and np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[1])


def test_philox_known_answers():
    L = R.lib()

    def run(ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        L.rf_philox4x32_10_host(c, k, o)
        return [hex(v) for v in o]

    assert run([0, 0, 0, 0], [0, 0]) == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    f = 0xFFFFFFFF
    assert run([f, f, f, f], [f, f]) == ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert run([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]
    u = np.array([L.rf_uniform_host(7, 2, e) for e in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01

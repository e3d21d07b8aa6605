// rf_spectrum.cpp -- wavevector tables, spectral density and amplitude
// multipliers of the spectral path, built on the host in FP64 (O(N_k) work,
// uploaded once per plan and kept resident in HBM).
//
// Rules follow spectra_RF.f90 (set_kMaxND :15-44, set_kPoints :50-168,
// set_SkVec :174-223, set_kSign :267-286) and the multipliers of
// randomFieldND.f90 (:262, :265, :409, :500-501), with the coordinates already
// divided by corrL (randomFieldND.f90:58-73).
#include <cmath>
#include <cstring>

#include "rf_internal.h"

namespace {

// set_kMaxND: kMax covering 99 % of the gaussian spectrum, per dimension count.
double kmax_for(int nDim) {
    switch (nDim) {
        case 1: return 6.457;
        case 2: return 7.035;
        default: return 7.355;
    }
}

// set_kSign row b (0-based) of the 2^(nDim-1) sign patterns.  Column 0 is
// always +1; column c >= 1 flips on the bit (nDim-1-c) of b, which is what
// cyclicMod(int((j-0.9)/2^(nDim-i)), 2) evaluates to for rows j = b+1.
double ksign(int nDim, int b, int c) {
    if (c == 0) return 1.0;
    int bit = (b >> (nDim - 1 - c)) & 1;
    return bit ? -1.0 : 1.0;
}

}  // namespace

int rf_validate_params(const rf_params* p, std::string* err) {
    if (!p) { *err = "params is NULL"; return RF_E_ARG; }
    if (p->nDim < 1 || p->nDim > 3) { *err = "nDim must be 1, 2 or 3"; return RF_E_ARG; }
    if (p->Nmc < 1) { *err = "Nmc must be >= 1"; return RF_E_ARG; }
    if (p->method < RF_ISOTROPIC || p->method > RF_FFT) { *err = "unknown method"; return RF_E_ARG; }
    if (p->method == RF_FFT) {
        // gen_Std_Gauss_FFT exists for 2-D and 3-D (randomFieldND.f90:565-597)
        if (p->nDim < 2) { *err = "method 4 (FFT) needs nDim 2 or 3"; return RF_E_UNSUPPORTED; }
        for (int i = 0; i < p->nDim; i++)
            if (p->gridN[i] < 2) { *err = "method 4 needs gridN >= 2 points per axis"; return RF_E_ARG; }
    }
    if (p->corrMod != RF_CM_GAUSSIAN && p->corrMod != RF_CM_COS) { *err = "unknown corrMod"; return RF_E_ARG; }
    if (p->mode != RF_REF_COMPAT && p->mode != RF_FIXED) { *err = "unknown mode"; return RF_E_ARG; }
    if (p->form < RF_FORM_FP64 || p->form > RF_FORM_MFMA_BF16) { *err = "unknown form"; return RF_E_ARG; }
    if (p->margiFirst < RF_FOM_NONE || p->margiFirst > RF_FOM_LOGNORMAL) { *err = "unknown margiFirst"; return RF_E_ARG; }
    if (p->margiFirst == RF_FOM_LOGNORMAL && !(p->fieldAvg > 0.0)) {
        *err = "ERROR - when using lognormal fieldAvg should be a positive number";
        return RF_E_LOGNORMAL;
    }
    for (int i = 0; i < p->nDim; i++) {
        if (!(p->corrL[i] > 0.0)) { *err = "corrL must be > 0"; return RF_E_ARG; }
        if (!(p->xRange[i] > 0.0) && p->method != RF_RANDOMIZATION) {
            *err = "xRange must be > 0"; return RF_E_ARG;
        }
    }
    if (p->method == RF_ISOTROPIC && p->nDim < 2) {
        *err = "isotropic method needs nDim 2 or 3"; return RF_E_UNSUPPORTED;
    }
    if (p->method == RF_RANDOMIZATION) {
        if (p->kSampling != RF_KSAMP_BOX && p->kSampling != RF_KSAMP_DENSITY) { *err = "unknown kSampling"; return RF_E_ARG; }
        if (p->kSampling == RF_KSAMP_BOX) {
            for (int i = 0; i < p->nDim; i++)
                if (!(p->xRange[i] > 0.0)) { *err = "xRange must be > 0"; return RF_E_ARG; }
        } else if (p->nKDensity < 1) {
            *err = "nKDensity must be >= 1 for density sampling"; return RF_E_ARG;
        }
    }
    if (p->isoShared && p->method != RF_ISOTROPIC) { *err = "isoShared needs the isotropic method"; return RF_E_ARG; }
    if (p->isoShared && p->mode != RF_FIXED) { *err = "isoShared is a FIXED-mode variant"; return RF_E_ARG; }
    if (p->method == RF_ISOTROPIC && !p->isoShared &&
        (p->form == RF_FORM_MFMA_F16 || p->form == RF_FORM_MFMA_BF16)) {
        *err = "the MFMA form needs directions shared across realizations (isoShared)";
        return RF_E_UNSUPPORTED;
    }
    return RF_OK;
}

int rf_spectrum_build_info(const rf_params* p, rf_spectrum_info* info, std::string* err) {
    int rc = rf_validate_params(p, err);
    if (rc) return rc;
    std::memset(info, 0, sizeof(*info));
    const int d = p->nDim;
    double xr[3] = {0, 0, 0};
    for (int i = 0; i < d; i++) xr[i] = p->xRange[i] / p->corrL[i];   // randomFieldND.f90:72
    const double km = kmax_for(d);
    for (int i = 0; i < d; i++) info->kMax[i] = km;

    if (p->method == RF_ISOTROPIC) {
        // spectra_RF.f90:76-79: one radial lattice spanning the box diagonal
        double diag2 = 0.0;
        for (int i = 0; i < d; i++) diag2 += xr[i] * xr[i];
        const double kd0 = 2.0 * RF_PI_REF / (1.1 * std::sqrt(diag2));
        const int nst = 1 + (int)(5.0 * (double)(int64_t)std::ceil(km / kd0));
        info->kNStep[0] = nst;
        info->kDelta[0] = km / (double)(nst - 1);
        info->kNTotal = nst;
        info->kN_unsigned = nst;
        info->nSignBlocks = 1;
        info->isoStep = info->kDelta[0];
        info->nVarIso = (d == 3) ? 4 : 3;
        info->nPhaseUniform = (int64_t)info->nVarIso * nst;
        // randomFieldND.f90:500-501; FIXED restores unit variance
        const double tp = 2.0 * RF_PI_REF;
        const double tpd = (d == 2) ? tp * tp : tp * tp * tp;
        info->ampMult = std::sqrt(info->isoStep / tpd);
        if (p->mode == RF_FIXED) info->ampMult *= (d == 3) ? 2.0 * RF_PI_REF : std::sqrt(4.0 * RF_PI_REF);
        return RF_OK;
    }
    if (p->method == RF_FFT) {
        // spectra_RF.f90:135-160: one wavevector per grid point, j * kDelta
        // with kDelta = 2 pi / (1.1 xRange), kMax = (kNStep - 1) kDelta
        int64_t kN = 1;
        for (int i = 0; i < d; i++) {
            info->kNStep[i] = p->gridN[i];
            info->kDelta[i] = 2.0 * RF_PI_REF / (1.1 * xr[i]);
            info->kMax[i] = ((double)(p->gridN[i] - 1) * info->kDelta[i]) / 1.0;
            kN *= p->gridN[i];
        }
        info->kNTotal = kN;
        info->kN_unsigned = kN;
        info->nSignBlocks = 1;
        // randomFieldND.f90:670-681: gammaK, kN - 1 discarded draws, phiK
        info->nPhaseUniform = 3 * kN - 1;
        info->ampMult = 1.0;            // the grid-step scale is applied at execution
        return RF_OK;
    }
    if (p->method == RF_RANDOMIZATION && p->kSampling == RF_KSAMP_DENSITY) {
        info->kNTotal = p->nKDensity;
        info->kN_unsigned = p->nKDensity;
        info->nSignBlocks = 1;
        info->nKUniform = 2 * (int64_t)d * p->nKDensity;
        info->nPhaseUniform = p->nKDensity;
        info->ampMult = std::sqrt(2.0 / (double)p->nKDensity);
        return RF_OK;
    }
    // spectra_RF.f90:88-101: half-space lattice, 2^(d-1) sign blocks
    int64_t kNu = 1;
    double dprod = 1.0;
    for (int i = 0; i < d; i++) {
        const double kd0 = 2.0 * RF_PI_REF / (1.1 * xr[i]);
        const int nst = 1 + (int)(1.0 * (double)(int64_t)std::ceil(km / kd0));
        info->kNStep[i] = nst;
        info->kDelta[i] = km / (double)(nst - 1);
        kNu *= nst;
        dprod *= info->kDelta[i];
    }
    info->nSignBlocks = 1 << (d - 1);
    info->kN_unsigned = kNu;
    info->kNTotal = kNu * info->nSignBlocks;
    info->nPhaseUniform = info->kNTotal;
    const double tpd = std::pow(2.0 * RF_PI_REF, (double)d);
    if (p->method == RF_RANDOMIZATION) {
        info->nKUniform = (int64_t)d * kNu;
        if (p->mode == RF_REF_COMPAT) {
            info->ampMult = 2.0 * std::sqrt(1.0 / ((double)info->kNTotal * tpd));   // :262
        } else {
            // the uniform box has volume 2^(d-1) prod(kMax); Monte-Carlo weight
            double vol = (double)info->nSignBlocks;
            for (int i = 0; i < d; i++) vol *= info->kMax[i];
            info->ampMult = 2.0 * std::sqrt(vol / ((double)info->kNTotal * tpd));
        }
    } else {
        info->ampMult = 2.0 * std::sqrt(dprod / tpd);   // :265
    }
    return RF_OK;
}

int rf_spectrum_build_tables(const rf_params* p, const rf_spectrum_info& info,
                             const double* kUniform, rf_spectrum* spec, std::string* err) {
    const int d = p->nDim;
    const int64_t kN = info.kNTotal;
    spec->info = info;
    if (p->method == RF_ISOTROPIC) {
        // intended radial lattice k_j = (j-1) kMax / (kNStep-1)
        // (get_Permutation with snapExtremes, math_RF.f90:137-145)
        spec->k.assign(kN, 0.0);
        spec->Sk.assign(kN, 0.0);
        for (int64_t j = 0; j < kN; j++)
            spec->k[j] = ((double)(j + 1) - 0.5 - 0.5) * info.kMax[0] / ((double)info.kNStep[0] - 1.0);
    } else {
        spec->k.assign((size_t)(d * kN), 0.0);
        spec->Sk.assign(kN, 0.0);
        if ((p->method == RF_RANDOMIZATION) && !kUniform) {
            *err = "method 3 needs its wavevector uniforms";
            return RF_E_ARG;
        }
        if (p->method == RF_RANDOMIZATION && p->kSampling == RF_KSAMP_DENSITY) {
            // k ~ N(0, 2 pi I): Box-Muller on consecutive uniform pairs
            const double sd = std::sqrt(2.0 * RF_PI_REF);
            for (int64_t c = 0; c < d * kN; c++) {
                const double r = std::sqrt(-2.0 * std::log(1.0 - kUniform[2 * c]));
                spec->k[c] = sd * (r * std::cos(2.0 * RF_PI_REF * kUniform[2 * c + 1]));
            }
        } else {
            const int64_t kNu = info.kN_unsigned;
            const int nb = info.nSignBlocks;
            for (int64_t j = 0; j < kNu; j++) {
                double base[3];
                if (p->method == RF_SHINOZUKA || p->method == RF_FFT) {
                    // setGrid(kPoints_base, 0, kDelta, kNStep), dim 1 fastest
                    int64_t rem = j;
                    for (int c = 0; c < d; c++) {
                        const int64_t idx = rem % info.kNStep[c];
                        rem /= info.kNStep[c];
                        base[c] = 0.0 + info.kDelta[c] * (double)idx;
                    }
                } else {
                    // random_number(kPoints_base): component fastest, * kMax(i)
                    for (int c = 0; c < d; c++) base[c] = kUniform[j * d + c] * info.kMax[c];
                }
                // block 0 is never written by the reference (:127); method 4
                // writes its one block directly (setGrid, :160)
                const bool blank0 = p->mode == RF_REF_COMPAT && p->method != RF_FFT;
                for (int b = 0; b < nb; b++) {
                    double* dst = &spec->k[(size_t)((b * kNu + j) * d)];
                    for (int c = 0; c < d; c++)
                        dst[c] = (b == 0 && blank0) ? 0.0 : base[c] * ksign(d, b, c);
                }
            }
        }
    }
    // set_SkVec (:200-221), corrL = 1 after normalisation
    if (p->method == RF_RANDOMIZATION && p->kSampling == RF_KSAMP_DENSITY) {
        for (int64_t j = 0; j < kN; j++) spec->Sk[j] = 1.0;
    } else if (p->corrMod == RF_CM_COS) {
        if (kN <= 6) {
            *err = "When using the cosinus correlation Model you should have more than 'freqK' kPoints";
            return RF_E_ARG;
        }
        spec->Sk[5] = 1.0;
    } else if (p->method == RF_ISOTROPIC) {
        for (int64_t j = 0; j < kN; j++) {
            const double k = spec->k[j];
            spec->Sk[j] = 1.0 * 1.0 * std::exp(-((k * k) * 1.0) / (4.0 * RF_PI_REF));
        }
    } else {
        for (int64_t j = 0; j < kN; j++) {
            double s = 1.0;
            int zeros = 0;
            for (int c = 0; c < d; c++) {
                const double k = spec->k[(size_t)(j * d + c)];
                s = s * 1.0 * std::exp(-((k * k) * 1.0) / (4.0 * RF_PI_REF));
                zeros += (k == 0.0);
            }
            // FIXED, Shinozuka lattice: trapezoid weight 1/2 per zero
            // component (the reference's k = 0 planes are counted twice by
            // the sign blocks, or whole at the end of the one-sided k_1 sum)
            if (p->mode == RF_FIXED && p->method == RF_SHINOZUKA) s = std::ldexp(s, -zeros);
            spec->Sk[j] = s;
        }
    }
    return RF_OK;
}

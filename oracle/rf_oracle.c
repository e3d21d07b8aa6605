/*
 * rf_oracle.c -- FP64 CPU restatement of randomField's spectral path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU product (see rf_oracle.h).
 * Compiled with -ffp-contract=off so every a*b+c rounds twice, as the
 * reference does at the -O0 its makefiles ship (makefileIGLOO:30).
 *
 * Citations are file:line in the reference (cottereau/randomField).
 */
#include "rf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- spectra */

/* spectra_RF.f90:15-44 -- 99 % spectral-area cut, gaussian model only */
void rfo_kmax(int corrMod, int nDim, double* kMax)
{
    double v = 0.0;
    if (corrMod == RFO_CM_GAUSSIAN || corrMod == RFO_CM_COS) {
        /* the reference only sets kMax for cm_GAUSSIAN; cm_COS reuses it */
        if (nDim == 1) v = 6.457;
        else if (nDim == 2) v = 7.035;
        else v = 7.355;
    }
    for (int i = 0; i < nDim; i++) kMax[i] = v;
}

/* math_RF.f90:380-391 */
static int cyclic_mod(int pos, int base)
{
    int r = pos % base;
    return r == 0 ? base : r;
}

/* spectra_RF.f90:267-286 -- kSign(j,i), j = 1..2^(nDim-1), i = 1..nDim */
void rfo_ksign(int nDim, double* kSign)
{
    int rows = 1 << (nDim - 1);
    for (int i = 0; i < nDim; i++)
        for (int j = 0; j < rows; j++) kSign[i * rows + j] = 1.0;
    for (int i = 2; i <= nDim; i++) {
        for (int j = 1; j <= rows; j++) {
            int seedStep = 1 << (nDim - i);
            /* int((j-0.9)/seedStep): Fortran int() truncates toward zero */
            int pos = cyclic_mod((int)(((double)j - 0.9) / (double)seedStep), 2);
            if (pos % 2 == 1) kSign[(i - 1) * rows + (j - 1)] = -1.0;
        }
    }
}

int rfo_make_plan(int nDim, int method, int corrMod, int mode, int ksamp,
                  int64_t nKDensity, const double* xRangeNorm, rfo_plan* p)
{
    memset(p, 0, sizeof(*p));
    if (nDim < 1 || nDim > 3) return -1;
    p->nDim = nDim; p->method = method; p->corrMod = corrMod;
    p->mode = mode; p->ksamp = ksamp;
    rfo_kmax(corrMod, nDim, p->kMax);
    if (method == RFO_ISOTROPIC) {
        /* spectra_RF.f90:76-79 */
        if (nDim < 2) return -1;
        double s2 = 0.0;
        for (int i = 0; i < nDim; i++) s2 += xRangeNorm[i] * xRangeNorm[i];
        double kd = 2.0 * RFO_PI / (1.1 * sqrt(s2));
        double kmax = p->kMax[0];
        for (int i = 1; i < nDim; i++) if (p->kMax[i] > kmax) kmax = p->kMax[i];
        int nst = 1 + (int)(5.0 * (double)(int64_t)ceil(kmax / kd));
        p->kNStep[0] = nst;
        p->kDelta[0] = kmax / (double)(nst - 1);
        p->kNTotal = nst;
        p->kN_unsigned = nst;
        p->nSignBlocks = 1;
        /* randomFieldND.f90:409 (intended table) */
        p->isoStep = p->kDelta[0];
        /* randomFieldND.f90:500-501; FIXED rescales to unit variance:
         * 3-D by 2*pi, 2-D by sqrt(4*pi) (closed form of the isotropic sum) */
        double tp = 2.0 * RFO_PI;
        double tpd = (nDim == 2) ? tp * tp : tp * tp * tp;
        p->ampMult = sqrt(p->isoStep / tpd);
        if (mode == RFO_FIXED) p->ampMult *= (nDim == 3) ? 2.0 * RFO_PI : sqrt(4.0 * RFO_PI);
        return 0;
    }
    if (method != RFO_SHINOZUKA && method != RFO_RANDOMIZATION) return -1;
    if (method == RFO_RANDOMIZATION && ksamp == RFO_KDENSITY) {
        if (nKDensity < 1) return -1;
        p->kNTotal = nKDensity;
        p->kN_unsigned = nKDensity;
        p->nSignBlocks = 1;
        p->nKUniform = 2 * (int64_t)nDim * nKDensity;
        p->ampMult = sqrt(2.0 / (double)nKDensity);
        return 0;
    }
    /* spectra_RF.f90:88-101 */
    int64_t kNu = 1;
    for (int i = 0; i < nDim; i++) {
        double kd = 2.0 * RFO_PI / (1.1 * xRangeNorm[i]);
        int nst = 1 + (int)(1.0 * (double)(int64_t)ceil(p->kMax[i] / kd));
        p->kNStep[i] = nst;
        p->kDelta[i] = p->kMax[i] / (double)(nst - 1);
        kNu *= nst;
    }
    p->nSignBlocks = 1 << (nDim - 1);
    p->kN_unsigned = kNu;
    p->kNTotal = kNu * p->nSignBlocks;
    double tpd = pow(2.0 * RFO_PI, (double)nDim);
    if (method == RFO_RANDOMIZATION) {
        p->nKUniform = (int64_t)nDim * kNu;
        /* randomFieldND.f90:262 */
        p->ampMult = 2.0 * sqrt(1.0 / ((double)p->kNTotal * tpd));
        if (mode == RFO_FIXED) {
            /* uniform draws over the half-space box of volume
             * 2^(d-1) prod(kMax): the amplitude needs that volume */
            double vol = (double)p->nSignBlocks;
            for (int i = 0; i < nDim; i++) vol *= p->kMax[i];
            p->ampMult = 2.0 * sqrt(vol / ((double)p->kNTotal * tpd));
        }
    } else {
        /* randomFieldND.f90:265 */
        double prod = 1.0;
        for (int i = 0; i < nDim; i++) prod *= p->kDelta[i];
        p->ampMult = 2.0 * sqrt(prod / tpd);
    }
    return 0;
}

int rfo_fft_plan(int nDim, int corrMod, const double* xRangeNorm, const int32_t* gridN, rfo_plan* p)
{
    memset(p, 0, sizeof(*p));
    if (nDim < 2 || nDim > 3) return -1;
    p->nDim = nDim; p->method = RFO_FFT; p->corrMod = corrMod; p->mode = RFO_FIXED;
    int64_t kN = 1;
    for (int i = 0; i < nDim; i++) {
        if (gridN[i] < 2) return -1;
        p->kNStep[i] = gridN[i];
        p->kDelta[i] = 2.0 * RFO_PI / (1.1 * xRangeNorm[i]);
        p->kMax[i] = ((double)(gridN[i] - 1) * p->kDelta[i]) / 1.0;
        kN *= gridN[i];
    }
    p->kNTotal = kN;
    p->kN_unsigned = kN;
    p->nSignBlocks = 1;
    p->ampMult = 1.0;
    return 0;
}

void rfo_kpoints(const rfo_plan* p, const double* kUnif, double* kPts)
{
    const int d = p->nDim;
    if (p->method == RFO_FFT) {
        /* setGrid(kPoints, 0, kDelta, kNStep) (spectra_RF.f90:160), dim 1 fastest */
        for (int64_t j = 0; j < p->kNTotal; j++) {
            int64_t rem = j;
            for (int dd = 0; dd < d; dd++) {
                const int64_t idx = rem % p->kNStep[dd];
                rem /= p->kNStep[dd];
                kPts[j * d + dd] = 0.0 + p->kDelta[dd] * (double)idx;
            }
        }
        return;
    }
    if (p->method == RFO_ISOTROPIC) {
        /* intended get_Permutation(i, kMax, kNStep, snapExtremes) value
         * (math_RF.f90:137-145): ((i - 0.5) - 0.5) * kMax / (kNStep - 1) */
        double kmax = p->kMax[0];
        for (int i = 1; i < d; i++) if (p->kMax[i] > kmax) kmax = p->kMax[i];
        (void)kmax;
        for (int64_t j = 0; j < p->kNTotal; j++)
            kPts[j] = ((double)(j + 1) - 0.5 - 1.0 / 2.0) * p->kMax[0]
                      / ((double)p->kNStep[0] - 1.0);
        return;
    }
    if (p->method == RFO_RANDOMIZATION && p->ksamp == RFO_KDENSITY) {
        /* build-defined: Box-Muller on uniform pairs, std sqrt(2*pi) in
         * coordinates normalised by corrL (S(k) = exp(-|k|^2/(4 pi))) */
        const double sd = sqrt(2.0 * RFO_PI);
        for (int64_t c = 0; c < d * p->kNTotal; c++) {
            double u1 = kUnif[2 * c], u2 = kUnif[2 * c + 1];
            double z = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * RFO_PI * u2);
            kPts[c] = sd * z;
        }
        return;
    }
    const int64_t kNu = p->kN_unsigned;
    double* base = (double*)malloc(sizeof(double) * (size_t)(d * kNu));
    if (p->method == RFO_SHINOZUKA) {
        /* setGrid(kPoints_base, 0, kDelta, kNStep) math_RF.f90:70-88:
         * dimension 1 fastest, value 0 + kDelta(d)*(i-1) */
        for (int64_t j = 0; j < kNu; j++) {
            int64_t rem = j;
            for (int dd = 0; dd < d; dd++) {
                int64_t idx = rem % p->kNStep[dd];
                rem /= p->kNStep[dd];
                base[j * d + dd] = 0.0 + p->kDelta[dd] * (double)idx;
            }
        }
    } else {
        /* spectra_RF.f90:120-123: random_number(kPoints_base) fills the
         * d x kN_unsigned array column-major, then scales row i by kMax(i) */
        for (int64_t j = 0; j < kNu; j++)
            for (int dd = 0; dd < d; dd++)
                base[j * d + dd] = kUnif[j * d + dd] * p->kMax[dd];
    }
    double kSign[3 * 4];
    rfo_ksign(d, kSign);
    const int rows = p->nSignBlocks;
    /* spectra_RF.f90:127-133: blocks 2..rows are written; block 1 is not
     * (REF_COMPAT reads it as zero, FIXED writes the base lattice) */
    for (int b = 0; b < rows; b++) {
        for (int64_t j = 0; j < kNu; j++) {
            for (int dd = 0; dd < d; dd++) {
                double v;
                if (b == 0 && p->mode == RFO_REF_COMPAT) v = 0.0;
                else v = base[j * d + dd] * kSign[dd * rows + b];
                kPts[(b * kNu + j) * d + dd] = v;
            }
        }
    }
    free(base);
}

/* spectra_RF.f90:200-221 with corrL = 1 after normalisation */
void rfo_skvec(const rfo_plan* p, const double* kPts, double* Sk)
{
    const int d = p->nDim;
    if (p->method == RFO_RANDOMIZATION && p->ksamp == RFO_KDENSITY) {
        /* density sampling: the spectrum is in the sampling, amplitude flat */
        for (int64_t j = 0; j < p->kNTotal; j++) Sk[j] = 1.0;
        return;
    }
    if (p->corrMod == RFO_CM_COS) {
        for (int64_t j = 0; j < p->kNTotal; j++) Sk[j] = 0.0;
        if (p->kNTotal > 6) Sk[5] = 1.0;
        return;
    }
    if (p->method == RFO_ISOTROPIC) {
        /* intended: S(|k|) = exp(-|k|^2/(4 pi)) on the radial table (the
         * reference loops nDim rows of a 1-row table, spectra_RF.f90:206) */
        for (int64_t j = 0; j < p->kNTotal; j++) {
            double k = kPts[j];
            Sk[j] = 1.0 * 1.0 * exp(-((k * k) * 1.0) / (4.0 * RFO_PI));
        }
        return;
    }
    for (int64_t j = 0; j < p->kNTotal; j++) {
        double s = 1.0;
        int zeros = 0;
        for (int i = 0; i < d; i++) {
            double k = kPts[j * d + i];
            s = s * 1.0 * exp(-((k * k) * 1.0) / (4.0 * RFO_PI));
            zeros += (k == 0.0);
        }
        /* FIXED, Shinozuka lattice: the lattice starts at k = 0 with full
         * weight, so the sign blocks count each k_a = 0 plane (a >= 2) twice
         * and the one-sided k_1 sum takes the k_1 = 0 plane whole (a left
         * Riemann sum): weight 1/2 per zero component is the trapezoid rule
         * (discrete variance ~1 instead of ~1.28 at 10-20 correlation
         * lengths).  REF_COMPAT keeps the reference's weights. */
        if (p->mode == RFO_FIXED && p->method == RFO_SHINOZUKA) s = ldexp(s, -zeros);
        Sk[j] = s;
    }
}

/* ------------------------------------------------------- Shinozuka sums */

void rfo_shinozuka(const rfo_plan* p, const double* x, int64_t N,
                   const double* kPts, const double* Sk, int Nmc,
                   const double* phaseUnif, const uint8_t* calc,
                   int64_t xSlab, double* out, int threads)
{
    const int d = p->nDim;
    const int64_t kN = p->kNTotal;
    /* randomFieldND.f90:237-239 */
    if (xSlab <= 0) xSlab = (int64_t)((double)d * 1e8 / (double)kN);
    if (xSlab > N) xSlab = N;
    if (xSlab < 1) xSlab = 1;
    int64_t xNSlab = (N + xSlab - 1) / xSlab;

    /* SkMat = sqrt(SkMat) is applied once per processed (n, m) pair
     * (randomFieldND.f90:312); REF_COMPAT keeps the running power, FIXED
     * uses sqrt(Sk) throughout.  amp[c] holds the table for the c-th pass. */
    double* amp = (double*)malloc(sizeof(double) * (size_t)kN);
    double* phi = (double*)malloc(sizeof(double) * (size_t)kN);
    double* skm = (double*)malloc(sizeof(double) * (size_t)kN);
    for (int64_t j = 0; j < kN; j++) skm[j] = Sk[j];
    for (int64_t i = 0; i < N * Nmc; i++) out[i] = 0.0;    /* :274 */
#ifdef _OPENMP
    int nthr = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    for (int n = 0; n < Nmc; n++) {
        /* :278-279 phiK = (2*pi) * U */
        for (int64_t j = 0; j < kN; j++) phi[j] = 2.0 * RFO_PI * phaseUnif[(int64_t)n * kN + j];
        for (int64_t m = 0; m < xNSlab; m++) {
            int64_t xs = m * xSlab, xe = xs + xSlab;
            if (xe > N) xe = N;
            if (calc && !calc[n]) continue;           /* :295 */
            for (int64_t j = 0; j < kN; j++) skm[j] = sqrt(skm[j]);   /* :312 */
            const double* a = (p->mode == RFO_REF_COMPAT) ? skm : amp;
            if (p->mode != RFO_REF_COMPAT)
                for (int64_t j = 0; j < kN; j++) amp[j] = sqrt(Sk[j]);
            double* col = out + (int64_t)n * N;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthr)
#endif
            for (int64_t i = xs; i < xe; i++) {
                const double* xi = x + i * d;
                double acc = 0.0;                      /* dgemm.f90:219-225 */
                for (int64_t j = 0; j < kN; j++) {
                    /* dgemm.f90:246-253: TEMP = 0 + sum_l x(l)*k(l), then
                     * C = 1*TEMP + 1*phi */
                    const double* kj = kPts + j * d;
                    double t = 0.0;
                    for (int l = 0; l < d; l++) t = t + xi[l] * kj[l];
                    double arg = 1.0 * t + 1.0 * phi[j];
                    double c = cos(arg);                /* :310 */
                    /* dgemm.f90:226-234: skip B(l,j) == 0, C += TEMP*A */
                    if (a[j] != 0.0) acc = acc + (1.0 * a[j]) * c;
                }
                col[i] = acc;
            }
            if (xe == N) break;                        /* :322 */
        }
    }
    /* :330 */
    for (int64_t i = 0; i < N * Nmc; i++) out[i] = p->ampMult * out[i];
    free(amp); free(phi); free(skm);
}

/* ------------------------------------------------------------ isotropic */

int rfo_iso_nvar(int nDim) { return nDim == 3 ? 4 : 3; }

void rfo_isotropic(const rfo_plan* p, const double* x, int64_t N,
                   const double* kRad, const double* Sk, int Nmc,
                   const double* unif, const uint8_t* calc, double* out,
                   int threads)
{
    const int d = p->nDim;
    const int64_t kN = p->kNTotal;
    const int nv = rfo_iso_nvar(d);
    double* rv = (double*)malloc(sizeof(double) * (size_t)(3 * kN));
    double* am = (double*)malloc(sizeof(double) * (size_t)kN);
    double* ps = (double*)malloc(sizeof(double) * (size_t)kN);
#ifdef _OPENMP
    int nthr = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    /* sqrt(12.0) is a default-real (single precision) literal in the
     * reference (randomFieldND.f90:468) */
    const double sq12 = (p->mode == RFO_REF_COMPAT) ? (double)sqrtf(12.0f) : sqrt(12.0);
    for (int n = 0; n < Nmc; n++) {
        double* col = out + (int64_t)n * N;
        for (int64_t i = 0; i < N; i++) col[i] = 0.0;
        if (calc && !calc[n]) continue;                /* :444-446, 487-489 */
        const double* U = unif + (int64_t)n * nv * kN;
        for (int64_t j = 0; j < kN; j++) {
            double k = kRad[j];
            if (d == 3) {
                /* :460-468: draws phi, theta, gamma, psi */
                double uphi = U[0 * kN + j], uth = U[1 * kN + j];
                double ugam = U[2 * kN + j], upsi = U[3 * kN + j];
                double psi = 2.0 * RFO_PI * upsi;
                double th = (p->mode == RFO_REF_COMPAT) ? 2.0 * RFO_PI * psi
                                                         : 2.0 * RFO_PI * uth;
                double ph = RFO_PI * uphi;
                double gam = sq12 * (ugam - 0.5);
                /* :474-476 */
                rv[3 * j + 0] = cos(th) * sin(ph) * k;
                rv[3 * j + 1] = sin(th) * sin(ph) * k;
                rv[3 * j + 2] = cos(ph) * k;
                /* :480 */
                am[j] = sqrt(Sk[j] * sin(ph) * (k * k)) * gam;
                ps[j] = psi;
            } else {
                /* :424-433: draws psi, theta, gamma */
                double upsi = U[0 * kN + j], uth = U[1 * kN + j];
                double ugam = U[2 * kN + j];
                double psi = 2.0 * RFO_PI * upsi;
                double th = (p->mode == RFO_REF_COMPAT) ? 2.0 * RFO_PI * psi
                                                         : 2.0 * RFO_PI * uth;
                double gam = (p->mode == RFO_REF_COMPAT) ? 2.0 * RFO_PI * ugam
                                                          : sq12 * (ugam - 0.5);
                rv[3 * j + 0] = cos(th) * k;
                rv[3 * j + 1] = sin(th) * k;
                rv[3 * j + 2] = 0.0;
                am[j] = sqrt(Sk[j] * k) * gam;
                ps[j] = psi;
            }
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthr)
#endif
        for (int64_t i = 0; i < N; i++) {
            const double* xi = x + i * d;
            double acc = 0.0;
            for (int64_t j = 0; j < kN; j++) {
                /* DGEMM T,N with beta = 0: C = 1*(0 + sum x*r) */
                double t = 0.0;
                for (int l = 0; l < d; l++) t = t + xi[l] * rv[3 * j + l];
                acc = am[j] * cos(1.0 * t + ps[j]) + acc;
            }
            col[i] = acc;
        }
    }
    /* :500-501 */
    for (int64_t i = 0; i < N * Nmc; i++) out[i] = p->ampMult * out[i];
    free(rv); free(am); free(ps);
}

/* --------------------------------------------------------- marginal map */

int rfo_marginal(int margiFirst, double fieldAvg, double fieldVar,
                 double* field, int64_t count)
{
    double nVar = 1.0, nAvg = 0.0;
    if (margiFirst == RFO_FOM_GAUSSIAN) {
        nVar = fieldVar; nAvg = fieldAvg;
    } else if (margiFirst == RFO_FOM_LOGNORMAL) {
        if (fieldAvg <= 0.0) return -2;                 /* :162-166 */
        nVar = log(1.0 + fieldVar / (fieldAvg * fieldAvg));
        nAvg = log(fieldAvg) - nVar / 2.0;
    }
    double s = sqrt(nVar);
    for (int64_t i = 0; i < count; i++) field[i] = field[i] * s + nAvg;   /* :171 */
    if (margiFirst == RFO_FOM_LOGNORMAL)
        for (int64_t i = 0; i < count; i++) field[i] = exp(field[i]);      /* :175 */
    return 0;
}

/* --------------------------------------------------- partition of unity */

/* --------------------------------------------------- localization geometry */

/* areEqual (math_RF.f90:397-413), TOLERANCE = 1e-9 (constants_RF.f90:8) */
static int are_equal(double a, double b) { return fabs(a - b) < 0.000000001; }

/* roundToMultiple (math_RF.f90:419-450) */
static double round_to_multiple(double nmb, double divisor, int up)
{
    if (divisor != 0.0) {
        nmb = nmb / fabs(divisor);
        if (are_equal(nmb, (double)lround(nmb))) nmb = (double)lround(nmb);
        else nmb = up ? ceil(nmb) : floor(nmb);
        nmb = nmb * fabs(divisor);
    }
    return nmb;
}

int rfo_localization_layout(int nDim, const int32_t* nFields, const int32_t* pointsPerCorrL,
                            const double* xMinGlobIn, const double* xMaxGlobIn, const double* corrL,
                            const double* overlapIn, double* g, int32_t* origins, double* xMinSub)
{
    double xStep[3], overlap[3], xMin[3], xMax[3], procExtent[3], stepProc[3];
    int32_t groupN[3], subN[3];
    for (int i = 0; i < nDim; i++) {
        const double P = (double)nFields[i];
        /* round_basic_inputs, topography_RF.f90:34-53 */
        xStep[i] = corrL[i] / (double)pointsPerCorrL[i];
        overlap[i] = overlapIn[i];
        if (nFields[i] == 1) overlap[i] = 0.0;
        overlap[i] = (double)lround(overlap[i] * corrL[i] / xStep[i]) * xStep[i] / corrL[i];
        if (nFields[i] != 1 && overlap[i] < 2.0 * xStep[i] / corrL[i]) overlap[i] = 2.0 * xStep[i] / corrL[i];
        /* set_global_extremes, topography_RF.f90:85-140 */
        xMin[i] = round_to_multiple(xMinGlobIn[i], xStep[i], 0);
        xMax[i] = round_to_multiple(xMaxGlobIn[i], xStep[i], 1);
        double ovlp = overlap[i] * corrL[i];
        double ovlpTot = ovlp * (P - 1.0) + 2.0 * (ovlp - xStep[i]);
        if (nFields[i] == 1) { ovlp = 0.0; ovlpTot = 0.0; }
        if (ovlpTot <= 0.0) ovlpTot = 0.0;
        double novlpTot = (xMax[i] - xMin[i]) - ovlpTot;
        if (novlpTot < 0.0) novlpTot = 0.0;
        novlpTot = round_to_multiple(novlpTot, xStep[i] * P, 1);
        const double novlp = novlpTot / P;
        xMax[i] = xMin[i] + (ovlpTot + novlpTot);
        double deltaProc = novlpTot / P + 2.0 * ovlp - 2.0 * xStep[i];
        if (nFields[i] == 1) deltaProc = novlpTot / P;
        stepProc[i] = novlp + ovlp;
        procExtent[i] = deltaProc;
        /* set_local_bounding_box, one rank: [start, start + procExtent];
         * find_xNStep = 1 + nint(range/xStep) (math_RF.f90:156-192) */
        subN[i] = 1 + (int32_t)lround(procExtent[i] / xStep[i]);
    }
    int nSub = 1;
    for (int i = 0; i < nDim; i++) nSub *= nFields[i];
    /* prepare_Localization, localization_RF.f90:150-176: starts on the
     * stride grid, setGrid inverse (last axis fastest); the group box */
    double gMin[3], gMax[3];
    for (int i = 0; i < nDim; i++) { gMin[i] = 1e308; gMax[i] = -1e308; }
    for (int s = 0; s < nSub; s++) {
        int rem = s, c[3] = {0, 0, 0};
        for (int i = nDim - 1; i >= 0; i--) { c[i] = rem % nFields[i]; rem /= nFields[i]; }
        for (int i = 0; i < nDim; i++) {
            const double start = xMin[i] + stepProc[i] * (double)c[i];
            if (xMinSub) xMinSub[s * nDim + i] = start;
            if (start < gMin[i]) gMin[i] = start;
            if (start + procExtent[i] > gMax[i]) gMax[i] = start + procExtent[i];
        }
    }
    for (int i = 0; i < nDim; i++) groupN[i] = 1 + (int32_t)lround((gMax[i] - gMin[i]) / xStep[i]);
    for (int s = 0; s < nSub && origins; s++) {
        int rem = s, c[3] = {0, 0, 0};
        for (int i = nDim - 1; i >= 0; i--) { c[i] = rem % nFields[i]; rem /= nFields[i]; }
        for (int i = 0; i < nDim; i++) {
            const double start = xMin[i] + stepProc[i] * (double)c[i];
            /* origin_Group-relative index: find_xNStep(xMin_Group, start) - 1 */
            origins[s * nDim + i] = (int32_t)lround((start - gMin[i]) / xStep[i]);
        }
    }
    int32_t nOvl[3];
    rfo_overlap_npoints(nDim, overlap, corrL, xStep, nOvl);
    for (int i = 0; i < nDim; i++) {
        double* r = g + 13 * i;
        r[0] = xStep[i]; r[1] = overlap[i]; r[2] = xMin[i]; r[3] = xMax[i];
        r[4] = procExtent[i]; r[5] = stepProc[i]; r[6] = groupN[i]; r[7] = subN[i]; r[8] = nOvl[i];
        r[9] = r[10] = r[11] = r[12] = 0.0;
    }
    return nSub;
}

/* localization_RF.f90:746 */
void rfo_overlap_npoints(int nDim, const double* overlap, const double* corrL,
                         const double* xStep, int32_t* nOvl)
{
    for (int i = 0; i < nDim; i++)
        nOvl[i] = (int32_t)lround((overlap[i] * corrL[i] - 2.0 * xStep[i]) / xStep[i]) + 1;
}

/* weight of tapered face position s = 1..n (s = 1 nearest the interior edge
 * of the overlap on an upper face; mirrored on a lower face).
 * REF_COMPAT: localization_RF.f90:800-813, v = s/(n+2), w = (1+cos(pi v))/2.
 * FIXED: v = s/(n+1) so that the two faces of an overlap sum to one, then
 * the square root so that the squared weights sum to one. */
static double face_weight(int s, int n, int upper, int mode)
{
    double v;
    if (mode == RFO_REF_COMPAT) {
        int q = upper ? s : (n + 1 - s);
        v = fabs((double)q / (double)(n + 2));
        return (1.0 + cos(RFO_PI * v)) / 2.0;
    }
    int q = upper ? s : (n + 1 - s);
    v = (double)q / (double)(n + 1);
    return sqrt((1.0 + cos(RFO_PI * v)) / 2.0);
}

void rfo_unity_partition(int nDim, const int32_t* xNStep, const int32_t* nOvl,
                         int faceMask, int mode, double* pu)
{
    int64_t tot = 1;
    for (int i = 0; i < nDim; i++) tot *= xNStep[i];
    for (int64_t p = 0; p < tot; p++) pu[p] = 1.0;     /* :748 */
    /* loop order j (face) outer, i (axis) inner: :755, :777 */
    for (int f = 0; f < 2; f++) {
        for (int i = 0; i < nDim; i++) {
            int n = nOvl[i];
            if (n < 1) continue;                          /* :779 */
            if (!(faceMask & (1 << (2 * i + f)))) continue;
            int dlim = (f == 0) ? 1 : xNStep[i] - n + 1;  /* :760-768 */
            for (int64_t p = 0; p < tot; p++) {
                int64_t rem = p, idx = 0;
                for (int a = 0; a <= i; a++) { idx = rem % xNStep[a]; rem /= xNStep[a]; }
                int s = (int)(idx + 1) - dlim + 1;
                if (s >= 1 && s <= n) pu[p] = pu[p] * face_weight(s, n, f == 1, mode);
            }
        }
    }
}

void rfo_blend(int nDim, const int32_t* groupN, int Nmc, int nSubdomains,
               const int32_t* origins, const int32_t* sizes,
               const int32_t* nOvl, const int32_t* faceMasks, int mode,
               const double* const* fields, double* group)
{
    int64_t gTot = 1;
    for (int i = 0; i < nDim; i++) gTot *= groupN[i];
    for (int64_t p = 0; p < gTot * Nmc; p++) group[p] = 0.0;   /* calls_RF.f90:202 */
    for (int s = 0; s < nSubdomains; s++) {
        const int32_t* sz = sizes + s * nDim;
        const int32_t* org = origins + s * nDim;
        int64_t sTot = 1;
        for (int i = 0; i < nDim; i++) sTot *= sz[i];
        double* pu = (double*)malloc(sizeof(double) * (size_t)sTot);
        rfo_unity_partition(nDim, sz, nOvl, faceMasks[s], mode, pu);
        for (int n = 0; n < Nmc; n++) {
            const double* F = fields[s] + (int64_t)n * sTot;
            double* G = group + (int64_t)n * gTot;
            for (int64_t p = 0; p < sTot; p++) {
                int64_t rem = p, g = 0, stride = 1;
                for (int i = 0; i < nDim; i++) {
                    int64_t idx = rem % sz[i]; rem /= sz[i];
                    g += (idx + org[i]) * stride;
                    stride *= groupN[i];
                }
                /* localization_RF.f90:228 then :253 */
                double w = F[p] * pu[p];
                G[g] = w + G[g];
            }
        }
        free(pu);
    }
    if (mode == RFO_REF_COMPAT) {
        /* calls_RF.f90:338-355: reverse taper on every face (no MPI
         * neighbours on one rank), column 1 only */
        double* pu1 = (double*)malloc(sizeof(double) * (size_t)gTot);
        rfo_unity_partition(nDim, groupN, nOvl, (1 << (2 * nDim)) - 1, mode, pu1);
        for (int64_t p = 0; p < gTot; p++) group[p] = group[p] / pu1[p];
        free(pu1);
    }
}

/* ---------------------------------------------------------- statistics */

void rfo_fft(const rfo_plan* p, const double* Sk, int Nmc, const double* unif, const uint8_t* calc,
             const double* xStepNorm, double* out)
{
    const int d = p->nDim;
    const int64_t kN = p->kNTotal, nu = 3 * kN - 1;
    const double pi = 3.14159265358979323846;   /* FFTW's transform is exact-math */
    double* cosT[3] = {NULL, NULL, NULL};
    int maxn = 1;
    for (int a = 0; a < d; a++) {
        const int n = p->kNStep[a];
        if (n > maxn) maxn = n;
        cosT[a] = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                cosT[a][(size_t)j * n + i] = cos(pi * (double)j * (2.0 * i + 1.0) / (2.0 * n));
    }
    double vol = 1.0;
    for (int a = 0; a < d; a++) vol *= xStepNorm[a];
    const double scale = (d == 3 ? 2.0 : 1.0) * sqrt(vol);
    double* line = (double*)malloc(sizeof(double) * (size_t)maxn);
    for (int r = 0; r < Nmc; r++) {
        double* f = out + (size_t)r * kN;
        if (calc && !calc[r]) {
            for (int64_t j = 0; j < kN; j++) f[j] = 0.0;
            continue;
        }
        const double* U = unif + (size_t)r * nu;
        for (int64_t j = 0; j < kN; j++) {
            const double g = U[j] - 0.5;
            f[j] = g * sqrt(Sk[j]) * cos(2.0 * RFO_PI * U[2 * kN - 1 + j]);
        }
        /* REDFT01 along axis a: lines of n[a] values at stride prod(n[<a]) */
        int64_t stride = 1;
        for (int a = 0; a < d; a++) {
            const int n = p->kNStep[a];
            const int64_t outer = kN / ((int64_t)n * stride);
            for (int64_t o = 0; o < outer; o++)
                for (int64_t in = 0; in < stride; in++) {
                    double* base = f + o * n * stride + in;
                    for (int i = 0; i < n; i++) {
                        double y = base[0];
                        for (int j = 1; j < n; j++) y += 2.0 * base[(int64_t)j * stride] * cosT[a][(size_t)j * n + i];
                        line[i] = y;
                    }
                    for (int i = 0; i < n; i++) base[(int64_t)i * stride] = line[i];
                }
            stride *= n;
        }
        for (int64_t j = 0; j < kN; j++) f[j] *= scale;
    }
    free(line);
    for (int a = 0; a < d; a++) free(cosT[a]);
}

void rfo_normalize(double* field, int64_t N, int Nmc)
{
    for (int n = 0; n < Nmc; n++) {
        double* c = field + (int64_t)n * N;
        double s = 0.0;
        for (int64_t i = 0; i < N; i++) s += c[i];
        double avg = s / (double)N;                      /* :869 */
        for (int64_t i = 0; i < N; i++) c[i] = c[i] - avg;   /* :874 */
        double s2 = 0.0;
        for (int64_t i = 0; i < N; i++) s2 += c[i] * c[i];
        double sd = sqrt(s2 / (double)N);               /* :904 */
        for (int64_t i = 0; i < N; i++) c[i] = c[i] / sd;    /* :908 */
    }
}

void rfo_stats(const double* field, int64_t N, int Nmc, double* avg,
               double* std, double* gAvg, double* gStd)
{
    double ts = 0.0, ts2 = 0.0;
    for (int n = 0; n < Nmc; n++) {
        const double* c = field + (int64_t)n * N;
        double s = 0.0, s2 = 0.0;
        for (int64_t i = 0; i < N; i++) { s += c[i]; s2 += c[i] * c[i]; }
        avg[n] = s / (double)N;
        std[n] = sqrt(s2 / (double)N - avg[n] * avg[n]);
        ts += s; ts2 += s2;
    }
    *gAvg = ts / ((double)N * Nmc);
    *gStd = sqrt(ts2 / ((double)N * Nmc) - (*gAvg) * (*gAvg));
}

/* rebuild_Sk_FFT + rebuild_Sk + rebuild_corrL (statistics_RF.f90:180-443) on
 * one realization u of a structured grid (x fastest), restated without the
 * FFT: per direction a, the line-wise FFT, division by n_a, |.|^2, average
 * over lines and unnormalised inverse FFT equal the circular
 * autocorrelation averaged over lines,
 *     R_a[m] = (1/N) sum_x v(x) v(x + m e_a mod n_a),  v = (u - avg)/std,
 * of which lags m = 0 .. n_a/2 - 1 are kept (:215-233).  R holds them for
 * a = 0 .. nDim-1 back to back.  corrL_a = 2 (sum R_a - (R_a[0] +
 * R_a[M-1])/2) xStep_a (:438-445).  Sums in FP64, loop order x fastest. */
void rfo_corr(const double* u, int nDim, const int32_t* n, double avg, double std, double* R)
{
    int64_t N = 1;
    for (int a = 0; a < nDim; a++) N *= n[a];
    int64_t stride[3] = {1, 1, 1};
    for (int a = 1; a < nDim; a++) stride[a] = stride[a - 1] * n[a - 1];
    double* out = R;
    for (int a = 0; a < nDim; a++) {
        const int M = n[a] / 2;
        for (int m = 0; m < M; m++) {
            double s = 0.0;
            for (int64_t i = 0; i < N; i++) {
                const int64_t ia = (i / stride[a]) % n[a];
                const int64_t ja = (ia + m) % n[a];
                const int64_t j = i + (ja - ia) * stride[a];
                s += ((u[i] - avg) / std) * ((u[j] - avg) / std);
            }
            out[m] = s / (double)N;
        }
        out += M;
    }
}

void rfo_corrL(const double* R, int nDim, const int32_t* n, const double* xStep, double* corrL)
{
    const double* r = R;
    for (int a = 0; a < nDim; a++) {
        const int M = n[a] / 2;
        double s = 0.0;
        for (int m = 0; m < M; m++) s += r[m];
        corrL[a] = M > 0 ? 2.0 * (s - (r[0] + r[M - 1]) / 2.0) * xStep[a] : 0.0;
        r += M;
    }
}

/* interpolateToMesh (sample_RF.f90:872-1106): multilinear interpolation of
 * a structured-grid field (x fastest, nstep points from xMin at xStep) to
 * nodes x[i*nDim + d], every realization.  1-based cell position
 * c = (x - xMin)/xStep + 1, cell floor(c) moved down one at the upper border
 * (:1024); the 2^d corners in the reference's order (:996-1013), weight
 * prod(1 - |c - corner|), accumulated in that order (:1072-1086).  Returns
 * the number of nodes outside the grid (their output is NaN). */
int64_t rfo_interp(int nDim, const int32_t* nstep, const double* xMin, const double* xStep, int Nmc,
                   const double* grid, const double* x, int64_t nNodes, double* out)
{
    static const int nb2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    static const int nb3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    int64_t Ng = 1, stride[3] = {1, 1, 1};
    for (int a = 0; a < nDim; a++) Ng *= nstep[a];
    for (int a = 1; a < nDim; a++) stride[a] = stride[a - 1] * nstep[a - 1];
    const int nc = 1 << nDim;
    int64_t bad = 0;
    for (int64_t i = 0; i < nNodes; i++) {
        double c[3];
        int64_t ci[3];
        int inside = 1;
        for (int a = 0; a < nDim; a++) {
            c[a] = (x[i * nDim + a] - xMin[a]) / xStep[a] + 1.0;
            ci[a] = (int64_t)floor(c[a]);
            if (ci[a] == nstep[a]) ci[a] -= 1;
            if (ci[a] < 1 || ci[a] + 1 > nstep[a]) inside = 0;
        }
        if (nDim == 1 && nstep[0] == 1) inside = 0;
        if (!inside) {
            bad++;
            for (int n = 0; n < Nmc; n++) out[(int64_t)n * nNodes + i] = NAN;
            continue;
        }
        for (int n = 0; n < Nmc; n++) {
            const double* g = grid + (int64_t)n * Ng;
            double acc = 0.0;
            for (int j = 0; j < nc; j++) {
                double w = 1.0;
                int64_t idx = 0;
                for (int a = 0; a < nDim; a++) {
                    const int off = nDim == 3 ? nb3[j][a] : (nDim == 2 ? nb2[j][a] : j);
                    const double dist = fabs(c[a] - (double)(ci[a] + off));
                    w = w * (1.0 - dist);
                    idx += (ci[a] + off - 1) * stride[a];
                }
                acc = acc + g[idx] * w;
            }
            out[(int64_t)n * nNodes + i] = acc;
        }
    }
    return bad;
}

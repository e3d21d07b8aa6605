/*
 * rf_oracle.h -- FP64 CPU restatement of randomField's spectral generation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (librf_hip.so) never links,
 * loads or falls back to it.
 *
 * Every function restates one routine of the reference (Fortran 90) and cites
 * the file:line it follows.  Two arithmetic modes:
 *   RFO_REF_COMPAT  bug-for-bug wherever the reference's behaviour is defined
 *                   (unwritten sign block 1 reads as zero, sqrt(Sk) re-applied
 *                   per realization and point slab, dependent isotropic azimuth,
 *                   off-by-one linear partition-of-unity weights);
 *   RFO_FIXED       the intended mathematics.
 * The Shinozuka/randomization/isotropic sums follow the reference's operation
 * order exactly (netlib DGEMM sequential contraction, no FMA contraction, libm
 * cos), so REF_COMPAT output is what the reference computes at -O0.
 *
 * Layouts (all column-major, as in the Fortran):
 *   x      nDim x N        x[i*nDim + d]
 *   kPts   nDim x kNTotal  k[j*nDim + d]
 *   out    N x Nmc         out[n*N + i]
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* constants_RF.f90:5 -- not the full-precision pi; parity needs this literal */
#define RFO_PI 3.1415926535898

enum { RFO_ISOTROPIC = 1, RFO_SHINOZUKA = 2, RFO_RANDOMIZATION = 3, RFO_FFT = 4 };
enum { RFO_CM_GAUSSIAN = 1, RFO_CM_COS = 2 };
enum { RFO_FOM_GAUSSIAN = 1, RFO_FOM_LOGNORMAL = 2 };
enum { RFO_REF_COMPAT = 0, RFO_FIXED = 1 };
/* wavevector sampling for method 3: the reference's uniform box with the
 * lattice-rule count and sign blocks, or i.i.d. draws from the normalised
 * spectral density (k ~ N(0, 2*pi*I), amplitude sqrt(2/N_k)). */
enum { RFO_KBOX = 0, RFO_KDENSITY = 1 };

typedef struct {
    int32_t nDim, method, corrMod, mode, ksamp;
    int32_t kNStep[3];
    int64_t kN_unsigned;   /* points of the base (unsigned) block        */
    int64_t kNTotal;       /* wavevectors in the table                    */
    int64_t nKUniform;     /* uniforms consumed to place k (method 3)     */
    int32_t nSignBlocks;   /* 2^(nDim-1) for methods 2/3 box, else 1      */
    double kMax[3], kDelta[3];
    double ampMult;        /* methods 2/3 multiplier; isotropic scale     */
    double isoStep;        /* isotropic radial step                       */
} rfo_plan;

/* spectra_RF.f90:15-44 */
void rfo_kmax(int corrMod, int nDim, double* kMax);
/* spectra_RF.f90:50-133 (+ randomFieldND.f90:260-266, 409, 500) ; xRange is
 * already divided by corrL (randomFieldND.f90:72).  Returns 0 or -1. */
int rfo_make_plan(int nDim, int method, int corrMod, int mode, int ksamp,
                  int64_t nKDensity, const double* xRangeNorm, rfo_plan* p);
/* spectra_RF.f90:135-139 (case FFT) on one rank: kNStep = gridN (the grid
 * points, find_xNStep(xRange, xStep) = 1 + nint(xRange/xStep)), kDelta =
 * 2 pi / (1.1 xRange), lattice from 0, one block.  nDim 2 or 3; returns 0 or -1. */
int rfo_fft_plan(int nDim, int corrMod, const double* xRangeNorm, const int32_t* gridN, rfo_plan* p);
/* spectra_RF.f90:267-286 ; kSign is (2^(nDim-1)) x nDim column-major */
void rfo_ksign(int nDim, double* kSign);
/* spectra_RF.f90:75-133 + math_RF.f90:16-92,98-148.  kUnif holds the
 * nKUniform draws of method 3 (NULL otherwise). */
void rfo_kpoints(const rfo_plan* p, const double* kUnif, double* kPts);
/* spectra_RF.f90:174-223 (intended form for the isotropic table) */
void rfo_skvec(const rfo_plan* p, const double* kPts, double* Sk);

/* randomFieldND.f90:184-337 (methods 2/3).  phaseUnif: Nmc x kNTotal raw
 * uniforms in draw order, phaseUnif[n*kNTotal + j].  calc: Nmc flags or NULL.
 * xSlab <= 0 selects the reference slab rule int(nDim*1e8/kNTotal).
 * threads: OpenMP threads over points (0 = library default). */
void rfo_shinozuka(const rfo_plan* p, const double* x, int64_t N,
                   const double* kPts, const double* Sk, int Nmc,
                   const double* phaseUnif, const uint8_t* calc,
                   int64_t xSlab, double* out, int threads);

/* randomFieldND.f90:368-514.  unif: per realization nVar x kNTotal uniforms
 * in the reference's draw order (3-D: phi, theta, gamma, psi; 2-D: psi, theta,
 * gamma) at unif[(n*nVar + v)*kNTotal + j]; only calculated realizations are
 * read. */
int  rfo_iso_nvar(int nDim);
void rfo_isotropic(const rfo_plan* p, const double* x, int64_t N,
                   const double* kRad, const double* Sk, int Nmc,
                   const double* unif, const uint8_t* calc, double* out,
                   int threads);

/* gen_Std_Gauss_FFT, randomFieldND.f90:670-735, on one rank, per
 * realization: unif[n*(3 kN - 1) + .] = gammaK (kN), kN - 1 discarded draws,
 * phiK (kN); coefficient (gammaK - 0.5) sqrt(Sk) cos(2 PI phiK); the FFTW
 * REDFT01 transform (DCT-III, y_i = x_0 + 2 sum_{j>=1} x_j cos(pi j (2i+1)/(2n)))
 * along every axis as direct FP64 sums; times sqrt(prod xStepNorm), and 2 more
 * in 3-D.  out[n*kN + i], x fastest; realizations with calc[n] == 0 are zero. */
void rfo_fft(const rfo_plan* p, const double* Sk, int Nmc, const double* unif, const uint8_t* calc,
             const double* xStepNorm, double* out);

/* randomFieldND.f90:139-178.  Returns 0, or -2 for lognormal with avg <= 0. */
int rfo_marginal(int margiFirst, double fieldAvg, double fieldVar,
                 double* field, int64_t count);

/* localization_RF.f90:705-857 (partitionType 1).  faceMask bit (2*i + f)
 * selects face f (0 lower, 1 upper) of axis i for tapering; REF_COMPAT uses
 * the reference pattern s/(n+2), FIXED the exact partition s/(n+1) with
 * square-root (variance-preserving) weights.  pu has prod(xNStep) entries,
 * x fastest. */
/* Localization geometry of one localization level, in the reference's own
 * floating-point steps (topography_RF.f90:19-144, localization_RF.f90:127-192,
 * topography_RF.f90:150-255 with one rank per subdomain).  g[] receives per
 * axis a (stride 13): xStep, overlap, xMinGlob, xMaxGlob, procExtent,
 * stepProc, groupN, subN, nOvl (localization_RF.f90:746), and 4 unused; per
 * subdomain s (field s+1, last axis fastest): origins (0-based group index),
 * xMinSub.  Returns the number of subdomains. */
int rfo_localization_layout(int nDim, const int32_t* nFields, const int32_t* pointsPerCorrL,
                            const double* xMinGlob, const double* xMaxGlob, const double* corrL,
                            const double* overlap, double* g, int32_t* origins, double* xMinSub);

void rfo_overlap_npoints(int nDim, const double* overlap, const double* corrL,
                         const double* xStep, int32_t* nOvl);
void rfo_unity_partition(int nDim, const int32_t* xNStep, const int32_t* nOvl,
                         int faceMask, int mode, double* pu);

/* localization_RF.f90:199-263 + calls_RF.f90:324-369: G[window_s] += F_s*PU_s
 * for every subdomain s, then (REF_COMPAT) column 1 divided by the reverse
 * taper of the group grid.  F_s is nSub[s] x Nmc, origin is the subdomain's
 * first grid index in the group grid, sizes its xNStep. */
void rfo_blend(int nDim, const int32_t* groupN, int Nmc, int nSubdomains,
               const int32_t* origins, const int32_t* sizes,
               const int32_t* nOvl, const int32_t* faceMasks, int mode,
               const double* const* fields, double* group);

/* randomFieldND.f90:820-926 over a whole field of N points: subtract the
 * per-realization mean then divide by the root mean square. */
void rfo_normalize(double* field, int64_t N, int Nmc);
/* statistics_RF.f90:62-87: per-realization mean and (biased) std, then the
 * global mean and std over all realizations. */
void rfo_corr(const double* u, int nDim, const int32_t* n, double avg, double std, double* R);
int64_t rfo_interp(int nDim, const int32_t* nstep, const double* xMin, const double* xStep, int Nmc,
                   const double* grid, const double* x, int64_t nNodes, double* out);
void rfo_corrL(const double* R, int nDim, const int32_t* n, const double* xStep, double* corrL);
void rfo_stats(const double* field, int64_t N, int Nmc, double* avg,
               double* std, double* gAvg, double* gStd);

#ifdef __cplusplus
}
#endif
#endif

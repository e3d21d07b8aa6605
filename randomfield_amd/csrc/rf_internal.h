// rf_internal.h -- shared declarations of librf_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/rf_capi.h"

// constants_RF.f90:5 -- the reference's pi literal (phases, spectra, tapers).
#define RF_PI_REF 3.1415926535898
// exact 2*pi and its reciprocal, for reductions to revolutions
#define RF_TWO_PI_EXACT 6.283185307179586476925286766559
#define RF_INV_TWO_PI 0.15915494309189533576888376337251

// A plan kept by the host-buffer entry points (rf_generate, rf_generate_grid)
// for calls with the same parameters, wavevector draws and mask.
struct rf_cached_plan {
    rf_params p;
    bool hasKU = false;            // caller wavevector uniforms (method 3)
    std::vector<double> kU;
    bool hasCalc = false;          // caller calculate mask
    std::vector<uint8_t> calc;
    rf_field_plan* plan = nullptr;
    uint64_t lastUse = 0;
};

struct rf_context {
    int device = -1;
    hipStream_t stream = nullptr;
    std::string last_error;
    int num_cu = 0;
    void* d_scratch = nullptr;   // reduction partials (grown on demand)
    size_t scratch_bytes = 0;
    std::vector<rf_cached_plan> planCache;   // at most kPlanCache, least recently used evicted
    uint64_t useClock = 0;
    void* d_io = nullptr;        // host-buffer entry points' points and outputs (grown on demand)
    size_t ioBytes = 0;
};
constexpr size_t kPlanCache = 4;

// Host-side spectrum: the wavevector table in coordinates normalised by corrL.
struct rf_spectrum {
    rf_spectrum_info info{};
    std::vector<double> k;    // nDim x kNTotal (isotropic: radii)
    std::vector<double> Sk;   // kNTotal
};

// Fused epilogue: optional marginal map (randomFieldND.f90:171-176).
struct rf_epilogue {
    int margi;      // RF_FOM_NONE / GAUSSIAN / LOGNORMAL
    double scale;   // sqrt(normalVar)
    double shift;   // normalAvg
};

struct rf_field_plan {
    rf_context* ctx = nullptr;
    rf_params p{};
    rf_spectrum spec;
    std::vector<uint8_t> calc;        // Nmc flags
    int nCalc = 0;
    // device buffers (owned)
    double* d_k = nullptr;       // shared k: kN x nDim (normalised units)
    double* d_kRad = nullptr;    // isotropic: radii kN
    double* d_kRev = nullptr;    // shared k / (2 pi)
    double* d_kIso = nullptr;    // per-realization k: Nmc x kN x nDim
    double* d_kIsoT = nullptr;   // FP32 isotropic form: kN x Rpad x nDim, revolutions
    double* d_Sk = nullptr;      // isotropic spectral density
    uint8_t* d_calc = nullptr;   // Nmc flags
    double* d_phase = nullptr;   // Nmc x kN radians (2*pi_ref*U or psi)
    double* d_amp = nullptr;     // Nmc x kN amplitudes (before ampMult)
    double* d_uni = nullptr;     // device copy of the phase uniforms
    float* d_phT = nullptr;      // kN x Rpad   FP32 phase in revolutions
    float* d_aT = nullptr;       // kN x Rpad   FP32 amplitude * ampMult
    void* d_B = nullptr;         // MFMA coefficient panels (hi, lo)
    int Rpad = 0;
    int64_t kNpad = 0;
    double bScale = 1.0;         // power-of-two scale folded into B
    double ampMax = 1.0;         // bound of |amp| over all realizations
    // REF_COMPAT: sqrt chain of Sk, row c = Sk^(2^-(c+1)), until it saturates
    std::vector<std::vector<double>> sqrtChain;
    bool coeffsReady = false;    // FIXED: coefficients independent of N
    rf_epilogue epi{};
    void* d_gridScratch = nullptr;   // structured-grid form (grown on demand)
    size_t gridScratchBytes = 0;
    float2* d_gridCoef = nullptr;    // folded grid coefficients (first grid execution)
    double gridCoefKey = NAN;
    int mfmaVariant = 0;         // MFMA launch shape (rf_plan_set_variant / rf_plan_autotune)
};

// ---- host spectrum (rf_spectrum.cpp)
int rf_spectrum_build_info(const rf_params* p, rf_spectrum_info* info, std::string* err);
int rf_spectrum_build_tables(const rf_params* p, const rf_spectrum_info& info,
                             const double* kUniform, rf_spectrum* spec, std::string* err);
int rf_validate_params(const rf_params* p, std::string* err);

// ---- kernel launchers (rf_sum.hip, rf_prep.hip, rf_mfma.hip, rf_post.hip)
hipError_t rf_launch_uniforms(uint64_t seed, uint32_t tag, int64_t count, double* d_out,
                              hipStream_t s);
// elements first .. first + count - 1 of stream (seed, tag)
hipError_t rf_launch_uniforms_at(uint64_t seed, uint32_t tag, uint64_t first, int64_t count,
                                 double* d_out, hipStream_t s);
hipError_t rf_launch_phase_from_uniform(const double* d_u, int64_t n, double* d_phase,
                                        hipStream_t s);
hipError_t rf_launch_iso_prep(int nDim, int mode, int shared, int64_t kN, int Nmc,
                              const double* d_kRad, const double* d_Sk, const double* d_u,
                              const uint8_t* d_calc, double* d_kIso, double* d_amp,
                              double* d_phase, hipStream_t s);
hipError_t rf_launch_prep_fp32(int64_t kN, int Nmc, int Rpad, const double* d_phase,
                               const double* d_amp, double ampMult, float* d_phT, float* d_aT,
                               hipStream_t s);
hipError_t rf_launch_krev(int64_t count, const double* d_k, double* d_kRev, hipStream_t s);
hipError_t rf_launch_fft_coeffs(int64_t kN, int Nmc, const double* d_Sk, const double* d_u,
                                const uint8_t* d_calc, double* d_amp, double* d_phase, hipStream_t s);
// k_fft_coeffs's amplitudes with the uniforms drawn in place from element
// `first` of stream (seed, tag 2); the phases are left as they are (zero)
hipError_t rf_launch_fft_coeffs_draw(int64_t kN, int Nmc, const double* d_Sk, const uint8_t* d_calc,
                                     uint64_t seed, uint64_t first, double* d_amp, hipStream_t s);

struct rf_sum_args {
    const double* x;      // nDim x N caller units
    int64_t N;
    int64_t i0, i1;       // point range of this launch
    double invL[3];       // 1/corrL ... FP64 kernel divides instead (see kernel)
    double corrL[3];
    int nDim;
    int Nmc;
    int64_t kN;
    double ampMult;
    rf_epilogue epi;
    void* out;
    int outDtype;
};
hipError_t rf_launch_sum_fp64(const rf_sum_args& a, const double* d_k, int64_t kStrideR,
                              const double* d_amp, const double* d_phase, hipStream_t s);
hipError_t rf_launch_sum_fp32(const rf_sum_args& a, const double* d_kRev, const float* d_phT,
                              const float* d_aT, int Rpad, hipStream_t s);
hipError_t rf_launch_sum_fp32_iso_t(const rf_sum_args& a, const double* d_kT, const float* d_pT,
                                    const float* d_aT, int Rpad, hipStream_t s);
hipError_t rf_launch_prep_fp32_iso(int nDim, int64_t kN, int Nmc, int Rpad, const double* d_kIso,
                                   const double* d_phase, const double* d_amp, double ampMult,
                                   double* d_kT, float* d_pT, float* d_aT, hipStream_t s);
int rf_fp32_rchunk();

// MFMA form (rf_mfma.hip)
int rf_mfma_kpad(int form);
int rf_mfma_num_variants();
hipError_t rf_launch_mfma_prep(int form, int nDim, int64_t kN, int64_t kNpad, int Nmc, int Npad,
                               const double* d_kRev, const double* d_phase, const double* d_amp,
                               int64_t ampStrideR, double scale, void* d_B, hipStream_t s);
hipError_t rf_launch_sum_mfma(int form, int variant, const rf_sum_args& a, int64_t kNpad, int Npad,
                              const void* d_B, double outScale, hipStream_t s);
int rf_mfma_npad(int Nmc);
// Power-of-two scale that brings the largest |amp * ampMult| near 2^14, so the
// FP16 hi/lo split keeps small amplitudes out of the subnormal range.
inline double rf_mfma_bscale(double ampMult, double maxAmp) {
    const double m = maxAmp * ampMult;
    if (!(m > 0.0)) return 1.0;
    int e = 0;
    std::frexp(m, &e);   // m = f * 2^e, f in [0.5, 1)
    return std::ldexp(1.0, 14 - e);
}
size_t rf_mfma_bbytes(int form, int64_t kNpad, int Npad);

// grouped launches: a batch of independent fields (serving requests) in one
// launch.  Field q owns tiles tile0 .. tile0 + nTiles(q) - 1 (ascending in q);
// block b finds its field by binary search over tile0.
struct rf_group_req {
    rf_sum_args a;        // points, output, epilogue of the field (i0 = 0, i1 = N)
    const void* ops0;     // MFMA: coefficient panels; direct isotropic: kT
    const void* ops1;     // MFMA: wavevector stream;  direct isotropic: pT
    const void* ops2;     //                           direct isotropic: aT
    int64_t tile0;        // first tile of the field
    int64_t nSteps;       // MFMA K-steps (multiple of 4)
    int32_t mtTotal;      // MFMA M-tiles of the panels
    int32_t Rpad;         // direct isotropic: padded realizations (multiple of 16)
    float outScale;       // MFMA: 1 / panel scale
    int32_t pad;
};
constexpr int kGroupMT = 4;       // M-tiles per tile, grouped MFMA kernel
constexpr int kGroupW = 8;        // waves per block, grouped MFMA kernel (256 points)
constexpr int kGroupIsoPts = 256; // points per tile, grouped direct isotropic kernel
constexpr int kGroupIsoR = 16;    // realizations per tile, grouped direct isotropic kernel
// tiles of one field
inline int64_t rf_group_tiles_mfma(int64_t N, int mtTotal) {
    return (N + 32 * kGroupW - 1) / (32 * kGroupW) * ((mtTotal + kGroupMT - 1) / kGroupMT);
}
inline int64_t rf_group_tiles_iso(int64_t N, int Rpad) {
    return (N + kGroupIsoPts - 1) / kGroupIsoPts * (Rpad / kGroupIsoR);
}
hipError_t rf_launch_group_mfma(int form, const rf_group_req* d_reqs, int nReq, int64_t nTiles,
                                hipStream_t s);
hipError_t rf_launch_group_iso(const rf_group_req* d_reqs, int nReq, int64_t nTiles, hipStream_t s);

// epilogue constants of multiVariateTransformation (randomFieldND.f90:154-169)
inline rf_epilogue rf_make_epilogue(int margi, double avg, double var) {
    rf_epilogue e{};
    e.margi = margi;
    e.scale = 1.0;
    e.shift = 0.0;
    if (margi == RF_FOM_GAUSSIAN) {
        e.scale = std::sqrt(var);
        e.shift = avg;
    } else if (margi == RF_FOM_LOGNORMAL) {
        const double nv = std::log(1.0 + var / (avg * avg));
        e.scale = std::sqrt(nv);
        e.shift = std::log(avg) - nv / 2.0;
    }
    return e;
}

// structured-grid form (rf_grid.hip), nDim 2 or 3, Shinozuka lattice:
// batched complex FP32 GEMMs D[b] = A[(b / aDiv) % aMod] B[b]
// The coefficient fold of one grid execution: Cf[k][m] for rows m (r, j3, q,
// j2) and K index k (h, j1) from amp / phase [r][b][j] (rf_grid.hip).
struct rf_fold_desc {
    int64_t rows; int K, nDim, fold, nk1, nk2, nk3;
    int64_t kNu, kNr;
    const double* amp; int64_t ampStrideR;  // amplitudes [r * ampStrideR + b kNu + j]
    const double* uni;                      // phase uniforms [r][kNr]; NULL: real coefficients (method 4)
    double ampMult;
    float2* C;
};
struct rf_gemm_desc {
    const float2* A; int64_t aSm, aSk, aStrideB; int aDiv, aMod;
    const float2* B; int64_t bSk, bSn, bStrideB;
    void* D; int64_t dSm, dSn, dStrideB;
    int M, N, K, batches;
    int realOut;          // 0: complex D, 1: Re(D) with the epilogue, outDtype
    int outDtype;
    rf_epilogue epi;
    int64_t tile0;        // grouped launches: first tile of this GEMM
    float outScale;       // D = outScale * (A B) (power of two)
    int split16;          // 1: FP16 hi/lo operands (3 products), 0: FP32 operands
    // two-level columns (nLo > 0): column n = (n / nLo, n % nLo) sits at
    // (n / nLo) * xSnHi + (n % nLo) * xSn in B and in D, so one GEMM spans
    // several batch blocks along N (a short axis is then padded once, not
    // once per block)
    int nLo;
    int64_t bSnHi, dSnHi;
    int rw;               // tile kind, FP16 operands only: 2 128x64 (complex), 3 64x128 (real),
                          // 4 64x128 (complex); 0, 1: 64x64
    // blend fused into a real last stage (grp != NULL): output row m (the
    // last axis) and column n = x + grpNx y (the plane) whose indices all lie
    // in the part of the field no other field covers (grpZlo <= m < grpZhi,
    // grpXlo <= x < grpXhi, grpYlo <= y < grpYhi) -- weight 1, the blended
    // value -- is stored to the group grid at grp + batch * grpStrideB +
    // grpRow0 + m grpPlaneG + grpCol0 + x + grpG0 y instead of to D
    void* grp;
    int64_t grpStrideB, grpRow0, grpPlaneG, grpCol0, grpG0;
    int grpNx, grpXlo, grpXhi, grpYlo, grpYhi, grpZlo, grpZhi;
    // fused: the points of the first-axis overlap bands (x < grpXlo, the low
    // band, x >= grpXhi, the high band) go to band arrays [batch][m][y][x']
    // (x' = x, resp. x - grpXhi; grpNy plane rows) instead of to D, so the
    // blend reads them as whole cache lines
    void* bandLo; void* bandHi;
    int64_t bandLoStrideB, bandHiStrideB;
    int grpNy;
    // fresh realizations formed by the first contraction of a batch: the
    // field's coefficient fold, seed and Nmc, and the batch's draw state
    // [0] next draw, [1] draw of the current realizations, [2] 1 while the
    // first contraction forms the coefficients of draw [1] itself instead of
    // reading them (genState NULL: always read)
    rf_fold_desc gen;
    uint64_t genSeed;
    int genNmc;
    const int64_t* genState;
};
struct rf_grid_stages {
    rf_gemm_desc g[3];    // stage 1, stage 2 (3-D only), last stage (real)
    int n;
    rf_fold_desc fold;    // the coefficient fold this execution formed (or reused)
};
// A field of a batch redraw: its fold, the seed of its phase stream and its
// first block in the grouped launch.
struct rf_fold_item {
    rf_fold_desc f;
    uint64_t seed;
    int Nmc;
    int nbx;          // row blocks of the fold
    int64_t block0;
};
int rf_fold_item_blocks(const rf_fold_desc& f);
void rf_fold_item_init(rf_fold_item* it, const rf_fold_desc& f, int Nmc, uint64_t seed, int64_t block0);
// Re-forms the folded coefficients of the batch's Shinozuka fields (n items,
// `blocks` = the last item's block0 + its blocks) from new phases 2 pi U, U
// element (*d_draw * Nmc + r) * kNr + b kNu + j of each field's device stream
// (seed, tag 2): realizations *d_draw * Nmc .. of the stream whose first Nmc
// are the plan's own device draws (draw index 0).  Reads the draw index on the
// device, so a captured graph draws what the counter holds at replay.
hipError_t rf_launch_refold_draw(const rf_fold_item* d_items, int n, int64_t blocks, const int64_t* d_draw,
                                 hipStream_t s);
// The draw state [0] next draw, [1] current draw, [2] coefficient mode of a
// batch, one thread: value >= 0 sets [0] = [1] = value; RF_DRAW_NEXT (-1)
// sets [1] = [0] and advances [0]; -2 initialises ([0] 1, [1] 0); -3 leaves
// the draws; then [2] = gen
hipError_t rf_launch_draw_counter(int64_t* d_state, int64_t value, int gen, hipStream_t s);
size_t rf_grid_scratch_elems(int nDim, const int32_t* kNStep, int nb, int Nmc, const int32_t* n);
// kind 0: Shinozuka lattice, tables exp(i k_j x); kind 1: method 4, tables of
// the DCT-III (FFTW REDFT01) of each axis, w_j cos(pi j (2i+1) / (2n))
// Folded coefficients kept across executions: buf holds Nmc * nb * prod(kNStep)
// elements, valid for the scale key (NaN: not yet computed).
struct rf_grid_coef {
    float2* buf;
    double key;
};
// coefMax bounds |ampMult * amp| (it sets the power-of-two scales that keep
// the FP16 operands of the split-precision contractions in range).  coef may
// be NULL (coefficients formed in scratch on every call).
hipError_t rf_grid_prepare(int kind, int nDim, int Nmc, const int32_t* kNStep, const double* kDelta, int nb,
                           const double* origin, const double* step, const int32_t* n,
                           const double* corrL, const double* d_amp, int ampShared, const double* d_unif,
                           double ampMult, double coefMax, const rf_epilogue& epi, void* d_out, int outDtype,
                           float2* scratch, hipStream_t s, rf_grid_stages* st, rf_grid_coef* coef = nullptr);
hipError_t rf_grid_run(const rf_grid_stages& st, hipStream_t s);
// Isotropic sum (per-realization directions) on a structured grid x = o + i
// step per axis (x fastest): per-axis phase tables, then per realization one
// real-part MFMA GEMM whose B operand is formed while staged (rf_grid.hip).
struct rf_iso_grid_args {
    int nDim, Nmc;
    int64_t kN;
    int n[3];                     // grid points per axis (1 beyond nDim)
    double origin[3], step[3], invL[3];
    const double* kIso;           // [Nmc][kN][nDim] wavevectors (radians per correlation length)
    const double* amp;            // [Nmc][kN]
    const double* phase;          // [Nmc][kN] radians
    double coefMax;               // bound of ampMult * |amp|
    double ampMult;
    rf_epilogue epi;
    void* d_out;                  // [Nmc][prod n]
    int outDtype;
    float2* scratch;              // rf_iso_grid_scratch_elems float2
};
size_t rf_iso_grid_scratch_elems(int nDim, int64_t kN, int Nmc, const int32_t* n);
hipError_t rf_launch_iso_grid(const rf_iso_grid_args& a, hipStream_t s);
int64_t rf_gemm_tiles(const rf_gemm_desc& g);
hipError_t rf_launch_gemm(const rf_gemm_desc& g, hipStream_t s);
hipError_t rf_launch_gemm_group(int realOut, const rf_gemm_desc* d_gs, int nG, int64_t nTiles, int rwMax,
                                hipStream_t s);

// post-processing (rf_post.hip)
hipError_t rf_launch_marginal(int margi, double scale, double shift, void* d, int64_t count,
                              int dtype, hipStream_t s);

// utilities
int rf_set_error(rf_context* ctx, int code, const std::string& msg);
// device address of the allocation-free entry's static scratch, looked up
// once per device with that device current (rf_noalloc.hip)
hipError_t rf_noalloc_setup(int device);
#define RF_HIP_TRY(ctx, expr)                                                        \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            return rf_set_error((ctx), RF_E_HIP,                                     \
                                std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

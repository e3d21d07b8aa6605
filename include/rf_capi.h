/*
 * rf_capi.h -- C ABI of librf_hip.so, the gfx950 spectral random-field path.
 *
 * This is the boundary the Fortran facade (randomfield_amd/fortran/
 * randomField_gpu.f90) binds to with iso_c_binding, and that C, C++ and
 * Python (ctypes) callers use directly.  Plain pointers, sizes and PODs only.
 *
 * What each entry point replaces in the reference (cottereau/randomField):
 *   rf_plan_spectrum / rf_spectrum_tables
 *       set_kMaxND, set_kPoints, set_kSign, set_SkVec and the amplitude
 *       multipliers (spectra_RF.f90:15-223, randomFieldND.f90:260-266,409,
 *       500-501).
 *   rf_plan_create + rf_plan_execute (rf_generate on host buffers)
 *       create_RF_Unstruct_Init -> gen_Std_Gauss -> gen_Std_Gauss_{Shinozuka,
 *       Randomization,Isotropic} (randomFieldND.f90:27-132,184-514), the body
 *       behind the generic createRandomField (calls_RF.f90:20-22).
 *   rf_marginal_device
 *       multiVariateTransformation (randomFieldND.f90:139-178).
 *   rf_blend_device
 *       generateUnityPartition_Matrix + add_RF_to_Group + the boundary
 *       correction of build_random_field (localization_RF.f90:199-263,
 *       705-857; calls_RF.f90:209-369).
 *   rf_stats_device / rf_normalize_device
 *       calculate_average_and_stdVar_MPI (statistics_RF.f90:18-174) and
 *       normalize_randField (randomFieldND.f90:820-926).
 *
 * Conventions
 *   - Layouts are the reference's column-major ones: points x[i*nDim + d]
 *     (nDim x N), fields out[n*N + i] (N x Nmc, point fastest).
 *   - Every call returns RF_OK (0) or a negative RF_E_* status; the message
 *     of the last failure on a context is rf_last_error(ctx).  The reference
 *     stops or calls MPI_ABORT instead (randomFieldND.f90:162-166,
 *     math_RF.f90:318-331).
 *   - "Uniforms" are raw U[0,1) draws in the reference's draw order
 *     (random_number after random_seed(PUT=seed)); the library turns them into
 *     phases 2*pi*U with the reference's pi literal.  A NULL draw pointer
 *     selects the library's device generator (Philox4x32-10 keyed by
 *     rf_params.seed), a documented stream that differs from the compiler's.
 *   - One context per host thread; calls on a context are ordered on its
 *     stream.  Device-pointer entry points take a hipStream_t as void*
 *     (NULL = the context's stream).
 */
#ifndef RF_CAPI_H
#define RF_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_CAPI_VERSION 1

/* status codes */
#define RF_OK               0
#define RF_E_ARG          (-1)   /* invalid argument / shape mismatch        */
#define RF_E_LOGNORMAL    (-2)   /* lognormal marginal with fieldAvg <= 0    */
#define RF_E_NOMEM        (-3)   /* device or host allocation failed         */
#define RF_E_HIP          (-4)   /* HIP runtime error                        */
#define RF_E_UNSUPPORTED  (-5)   /* combination not implemented              */
#define RF_E_NODEVICE     (-6)   /* no usable gfx950 device                  */

/* methods: constants_RF.f90:12-15 */
#define RF_ISOTROPIC      1
#define RF_SHINOZUKA      2
#define RF_RANDOMIZATION  3
#define RF_FFT            4   /* structured grids only (rf_plan_execute_grid) */
/* correlation models: constants_RF.f90:17-18 */
#define RF_CM_GAUSSIAN    1
#define RF_CM_COS         2
/* first-order marginals: constants_RF.f90:20-21 (0 = leave standard) */
#define RF_FOM_NONE       0
#define RF_FOM_GAUSSIAN   1
#define RF_FOM_LOGNORMAL  2
/* arithmetic modes */
#define RF_REF_COMPAT     0   /* reference behaviour wherever it is defined */
#define RF_FIXED          1   /* intended mathematics                        */
/* precision forms of the spectral sum */
#define RF_FORM_FP64        0 /* FP64 direct sum, reference operation order  */
#define RF_FORM_FP32        1 /* FP64 k.x mod 2pi, FP32 cos, FP32 accumulate */
#define RF_FORM_MFMA_F16    2 /* [cos Kx|sin Kx].[A cos;-A sin] FP16 hi/lo   */
#define RF_FORM_MFMA_BF16   3 /* same with BF16 hi/lo                         */
/* wavevector sampling for RF_RANDOMIZATION */
#define RF_KSAMP_BOX      0   /* reference: U[0,kMax]^d, lattice-rule count  */
#define RF_KSAMP_DENSITY  1   /* k ~ N(0, 2 pi I) (normalised), N_k given    */
/* output element types */
#define RF_DTYPE_F64      0
#define RF_DTYPE_F32      1
/* launch shapes of the MFMA forms (rf_plan_set_variant) */
#define RF_MFMA_VARIANTS  9

typedef struct rf_params {
    int32_t nDim;        /* 1..3 (isotropic: 2..3)                            */
    int32_t method;      /* RF_ISOTROPIC / RF_SHINOZUKA / RF_RANDOMIZATION     */
    int32_t corrMod;     /* RF_CM_*                                            */
    int32_t mode;        /* RF_REF_COMPAT / RF_FIXED                           */
    int32_t form;        /* RF_FORM_*                                          */
    int32_t kSampling;   /* RF_KSAMP_* (method 3 only)                         */
    int32_t Nmc;         /* realizations                                       */
    int32_t margiFirst;  /* RF_FOM_*: marginal map fused into generation       */
    int64_t nKDensity;   /* N_k for RF_KSAMP_DENSITY                            */
    int64_t xSlab;       /* REF_COMPAT point-slab size; 0 = reference rule      */
    double  corrL[3];    /* correlation lengths (> 0)                          */
    double  xRange[3];   /* extent of the generation box, caller units         */
    double  fieldAvg;    /* marginal mean                                      */
    double  fieldVar;    /* marginal variance                                  */
    uint64_t seed;       /* device generator key when no uniforms are given    */
    int32_t isoShared;   /* isotropic: one direction set for all realizations  */
    int32_t gridN[3];    /* method 4: grid points per axis (the DCT sizes)     */
    int32_t reserved[4]; /* zero                                               */
} rf_params;

typedef struct rf_spectrum_info {
    int64_t kNTotal;        /* wavevectors in the table                        */
    int64_t kN_unsigned;    /* points of one sign block                        */
    int64_t nKUniform;      /* uniforms placing k (method 3), drawn first      */
    int64_t nPhaseUniform;  /* uniforms per realization                        */
    int32_t kNStep[3];
    int32_t nSignBlocks;
    int32_t nVarIso;        /* isotropic draw arrays per realization           */
    int32_t pad;
    double  kMax[3];
    double  kDelta[3];
    double  ampMult;        /* sum multiplier (isotropic: overall scale)       */
    double  isoStep;        /* isotropic radial step                           */
} rf_spectrum_info;

typedef struct rf_context rf_context;
typedef struct rf_field_plan rf_field_plan;

/* ---- context ------------------------------------------------------------ */
int         rf_context_create(rf_context** ctx, int device);
int         rf_context_destroy(rf_context* ctx);
const char* rf_last_error(const rf_context* ctx);   /* ctx may be NULL       */
int         rf_version(void);
int         rf_context_sync(rf_context* ctx);

/* ---- host-side spectrum (no GPU) ------------------------------------------ */
int rf_plan_spectrum(const rf_params* p, rf_spectrum_info* info);
/* kPoints: nDim x kNTotal (isotropic: 1 x kNTotal radii); Sk: kNTotal.
 * kUniform: info.nKUniform draws (method 3) or NULL otherwise. */
int rf_spectrum_tables(const rf_params* p, const double* kUniform,
                       double* kPoints, double* Sk);

/* ---- generation ----------------------------------------------------------- */
/* Builds the device-resident wavevector table and per-realization
 * coefficients.  All pointers are HOST pointers:
 *   kUniform     info.nKUniform draws (method 3), NULL = device generator
 *   phaseUniform Nmc x info.nPhaseUniform draws, realization-major, NULL =
 *                device generator
 *   calculate    Nmc flags (0 = leave the realization zero), NULL = all.  */
int rf_plan_create(rf_context* ctx, const rf_params* p,
                   const double* kUniform, const double* phaseUniform,
                   const uint8_t* calculate, rf_field_plan** plan);
int rf_plan_destroy(rf_field_plan* plan);
int rf_plan_info(const rf_field_plan* plan, rf_spectrum_info* info);
/* The hot path.  d_x: device nDim x N FP64 points in caller units;
 * d_out: device N x Nmc in outDtype.  Stream-ordered, asynchronous. */
int rf_plan_execute(rf_field_plan* plan, const double* d_x, int64_t N,
                    void* d_out, int32_t outDtype, void* stream);
/* Launch shape of the MFMA forms (no counterpart in the reference, whose
 * sum has one code path): variant 0 .. RF_MFMA_VARIANTS-1 = {waves per
 * block, K-steps per LDS stage, MFMA issue order} = {4,1,pairs}, {4,2,pairs},
 * {8,1,pairs}, {8,2,pairs}, {8,4,pairs}, {8,2,product-major},
 * {8,4,product-major}, {8,2,pairs, wavevectors staged in LDS},
 * {8,4,pairs, wavevectors in LDS}.  Results are identical across variants (same
 * per-point operation order); only the speed differs.  rf_plan_autotune
 * times every variant on the caller's points (each runs three times, the
 * last run's result stays in d_out), keeps the fastest in the plan and
 * returns it in *chosen; for the FP64/FP32 forms it does nothing. */
int rf_plan_set_variant(rf_field_plan* plan, int32_t variant);
int rf_plan_autotune(rf_field_plan* plan, const double* d_x, int64_t N,
                     void* d_out, int32_t outDtype, void* stream,
                     int32_t* chosen);
/* The same sum on a structured grid x = origin + (i1,i2,i3)*step (caller
 * units, x fastest; the point order of set_XPoints/setGrid, mesh_RF.f90:44,
 * math_RF.f90:70-88), evaluated axis by axis: the Shinozuka lattice is a
 * tensor product per sign block, so sum_j C_j exp(i k.x) factorises into
 * three small contractions.  Method 2, FIXED mode, nDim 2 or 3, or method 4
 * (gen_Std_Gauss_FFT: nstep = params.gridN, DCT-III tables), or method 1
 * (randomFieldND.f90:368-514), nDim 2 or 3, either mode: each term
 * exp(i(k.x + phi)) factorises over the axes, so per realization the field
 * is the real part of one GEMM over the shells whose operands are per-axis
 * phase tables; otherwise RF_E_UNSUPPORTED.  d_out: prod(nstep) x Nmc.  The
 * plan keeps a scratch
 * buffer, grown on the first call with a larger grid, and the folded
 * coefficients, formed on the first call and reused while the scale is the
 * same: calls on one plan are ordered by their stream, so calls on different
 * streams must not overlap. */
int rf_plan_execute_grid(rf_field_plan* plan, const double origin[3],
                         const double step[3], const int32_t nstep[3],
                         void* d_out, int32_t outDtype, void* stream);
/* Fresh realizations for a plan (createRandomField draws new phases on each
 * call, randomFieldND.f90:276-279 and :683-689 for method 4): the phase
 * uniforms become realizations d*Nmc .. d*Nmc + Nmc - 1 of the device
 * stream (rf_params.seed, tag 2), d = drawIndex >= 0 (0 repeats the draws of
 * a plan created without caller draws; a plan created with them switches to
 * the stream), and everything built from them is re-formed: phases (methods
 * 2/3), directions, amplitudes and phases (method 1), DCT coefficients
 * (method 4), the form's operand tables, the grid form's folded coefficients.
 * Method 3's wavevector draws (tag 1) stay.  Enqueued on stream (NULL: the
 * context's); allocates nothing; later executes on another stream must be
 * ordered after it by the caller.  No separate entry in the reference. */
int rf_plan_redraw(rf_field_plan* plan, int64_t drawIndex, void* stream);
/* Host-buffer form of rf_plan_execute_grid (the Fortran facade's structured
 * path; method 4 runs here, as gen_Std_Gauss_FFT does on the mesh's grid,
 * randomFieldND.f90:520-763): plan, grid of nstep points from origin at step
 * (caller units, x fastest), out N x Nmc; synchronous.
 *
 * rf_generate and rf_generate_grid keep the plans they build on the context
 * (at most 4, the least recently used dropped; freed by rf_context_destroy)
 * with a staging buffer that grows to the largest call.  A call whose
 * rf_params bytes, wavevector uniforms and calculate mask equal a kept
 * plan's refreshes that plan's draws instead of building a new one: the
 * caller's phaseUniform re-forms it, a NULL phaseUniform redraws draw 0 of
 * the device stream -- the values a new plan would hold, bit for bit.  The
 * reference rebuilds its k table and spectrum on every call (gen_Std_Gauss,
 * randomFieldND.f90:44); here only the draws are renewed. */
int rf_generate_grid(rf_context* ctx, const rf_params* p, const double* origin,
                     const double* step, const int32_t* nstep, const double* phaseUniform,
                     const uint8_t* calculate, void* out, int32_t outDtype);
/* Host-buffer convenience: plan (kept, as above), upload, execute, download.
 * x: nDim x N, out: N x Nmc (outDtype). */
int rf_generate(rf_context* ctx, const rf_params* p, const double* x,
                int64_t N, const double* kUniform, const double* phaseUniform,
                const uint8_t* calculate, void* out, int32_t outDtype);

/* ---- post-processing on device buffers ------------------------------------ */
/* multiVariateTransformation in place over count values of dtype. */
int rf_marginal_device(rf_context* ctx, int32_t margiFirst, double fieldAvg,
                       double fieldVar, void* d_field, int64_t count,
                       int32_t dtype, void* stream);

/* multiVariateTransformation of count FP64 values in a HOST buffer (uploaded,
 * transformed on the device, downloaded; synchronous). */
int rf_marginal_host(rf_context* ctx, int32_t margiFirst, double fieldAvg,
                     double fieldVar, double* field, int64_t count);

/* Overlap blend of nSub structured subdomain fields into a group grid.
 *   groupN[nDim]            group grid points per axis (x fastest)
 *   origins[nSub*nDim]      first group index of each subdomain
 *   sizes[nSub*nDim]        points per axis of each subdomain
 *   nOvl[nDim]              overlap points (localization_RF.f90:746)
 *   faceMasks[nSub]         bit 2*axis+face: taper that face (0 lower, 1 up)
 *   d_fields[nSub]          device pointers, sizes[s] product x Nmc
 *   d_group                 device, prod(groupN) x Nmc, overwritten
 * REF_COMPAT: linear reference weights then column 1 divided by the reverse
 * taper of the group grid; FIXED: square-root exact partition of unity. */
int rf_blend_device(rf_context* ctx, int32_t nDim, const int32_t* groupN,
                    int32_t Nmc, int32_t nSub, const int32_t* origins,
                    const int32_t* sizes, const int32_t* nOvl,
                    const int32_t* faceMasks, int32_t mode,
                    const void* const* d_fields, void* d_group,
                    int32_t dtype, void* stream);

/* generateUnityPartition_Matrix (localization_RF.f90:705-857) on the host:
 * pu (prod(xNStep), x fastest) = product of the face factors of the faces
 * faceMask selects (bit 2a lower, 2a+1 upper face of axis a), nOvl[a] points
 * per face (< 1: axis untapered).  REF_COMPAT: the reference's cosine taper
 * (partitionType 1, :800-813); FIXED: its square-root exact form. */
int rf_unity_partition(int32_t nDim, const int32_t* xNStep, const int32_t* nOvl, int32_t faceMask,
                       int32_t mode, double* pu);
/* rf_blend_device on host buffers (the Fortran facade's localization step,
 * calls_RF.f90:209-369 with every field added at once): fields[s] FP64,
 * group FP64 prod(groupN) x Nmc, overwritten.  Synchronous. */
int rf_blend_host(rf_context* ctx, int32_t nDim, const int32_t* groupN, int32_t Nmc, int32_t nSub,
                  const int32_t* origins, const int32_t* sizes, const int32_t* nOvl,
                  const int32_t* faceMasks, int32_t mode, const double* const* fields, double* group);

/* The blend split into set-up and launch (same arguments and result as
 * rf_blend_device): create analyses the boxes and uploads the tables once
 * (synchronous); execute enqueues one kernel and allocates nothing, so it
 * can be captured into a graph; destroy frees the tables (no launch of the
 * plan may be pending). */
typedef struct rf_blend_plan rf_blend_plan;
int rf_blend_plan_create(rf_context* ctx, int32_t nDim, const int32_t* groupN,
                         int32_t Nmc, int32_t nSub, const int32_t* origins,
                         const int32_t* sizes, const int32_t* nOvl,
                         const int32_t* faceMasks, int32_t mode,
                         const void* const* d_fields, int32_t dtype,
                         rf_blend_plan** plan);
int rf_blend_plan_execute(rf_blend_plan* plan, void* d_group, void* stream);
int rf_blend_plan_destroy(rf_blend_plan* plan);

/* Localization geometry (one localization level; the reference's input
 * file gives xMinGlob/xMaxGlob, corrL, pointsPerCorrL, overlap and nFields):
 *   round_basic_inputs     topography_RF.f90:19-58
 *   set_global_extremes    topography_RF.f90:63-144
 *   prepare_Localization   localization_RF.f90:127-192
 *   set_local_bounding_box topography_RF.f90:150-255 (one rank per subdomain)
 *   overlap points         localization_RF.f90:746
 * Host only, no GPU.  Outputs per subdomain s = 0 .. info.nSub-1 (the
 * reference's field number s+1, last axis fastest; its seed is
 * seedStart + s + 1, sample_RF.f90:309-312), each nullable:
 *   origins[s*nDim+a]  first point of the subdomain in the group grid (0-based)
 *   sizes[s*nDim+a]    points per axis (all equal in the reference)
 *   faceMasks[s]       bit 2a / 2a+1: the lower / upper face along axis a has a
 *                      neighbour (the faces rf_blend_device tapers in FIXED
 *                      mode; REF_COMPAT tapers every face, as the reference)
 *   xMinSub[s*nDim+a]  coordinate of the subdomain's first point
 * origins/sizes/faceMasks/info.nOvl are the arguments rf_blend_device takes. */
typedef struct rf_layout_params {
    int32_t nDim;
    int32_t nFields[3];         /* subdomains per axis (IPT%nFields)            */
    int32_t pointsPerCorrL[3];
    int32_t pad;
    double  xMinGlob[3], xMaxGlob[3];
    double  corrL[3];
    double  overlap[3];         /* in correlation lengths (IPT%overlap)         */
} rf_layout_params;

typedef struct rf_layout_info {
    int32_t nDim, nSub;
    int32_t groupN[3];          /* group grid points per axis (xNStep_Group)    */
    int32_t subN[3];            /* subdomain points per axis (xNStep)           */
    int32_t nOvl[3];            /* overlap points, >= 0 (blend argument)        */
    int32_t nOvlRaw[3];         /* the reference's formula (< 1: no taper)      */
    double  xStep[3];
    double  overlap[3];         /* rounded overlap, correlation lengths         */
    double  xMinGlob[3], xMaxGlob[3];  /* rounded and grown global box          */
    double  procExtent[3];      /* subdomain extent                             */
    double  stepProc[3];        /* stride between subdomain starts             */
    double  xRangeSub[3];       /* RDF%xRange of a subdomain (its lattice box)  */
} rf_layout_info;

int rf_localization_layout(const rf_layout_params* in, rf_layout_info* info, int32_t* origins,
                           int32_t* sizes, int32_t* faceMasks, double* xMinSub);

/* Batched structured-grid generation: the subdomains of a localization
 * group (one MPI rank each in the reference, localization_RF.f90:127-192,
 * calls_RF.f90:250-296) as one batch.  Field i is rf_plan_execute_grid of
 * plans[i] on origins/steps/nsteps[3i..3i+2] into d_outs[i]; create builds
 * every field's phase tables and coefficients and the GEMM descriptors
 * (allocates, synchronous); execute enqueues three grouped launches (all
 * fields' first, middle and last contractions) and allocates nothing, so it
 * can be captured into a graph.  Results equal rf_plan_execute_grid's. */
typedef struct rf_grid_batch rf_grid_batch;
int rf_grid_batch_create(rf_context* ctx, int32_t n, rf_field_plan* const* plans,
                         const double* origins, const double* steps,
                         const int32_t* nsteps, void* const* d_outs,
                         int32_t outDtype, rf_grid_batch** batch);
int rf_grid_batch_execute(rf_grid_batch* batch, void* stream);
/* Fresh realizations for the next executes: every field's coefficients are
 * re-formed from new phases 2 pi U, U from the device stream of its plan
 * (rf_params.seed, tag 2), realizations d*Nmc .. d*Nmc + Nmc - 1 of that
 * stream, where d = drawIndex (>= 0; 0 gives the plan's own realizations when
 * it was created without caller draws) or, for RF_DRAW_NEXT, a counter the
 * batch keeps on the device (1 after create), used and then advanced -- so a
 * graph captured with RF_DRAW_NEXT draws new realizations on every replay.
 * Each realization is generated as in the reference (a new random_number
 * draw per createRandomField call, randomFieldND.f90:276-279).  Allocation-
 * free, capturable; method 4 batches: RF_E_UNSUPPORTED. */
#define RF_DRAW_NEXT (-1)
/* Times each contraction stage of the batch with the wide tile kinds the
 * set-up chose, with 128 x 64 tiles where it chose a 64 x 160 tile spanning a
 * whole 129..160-point axis, and with 64 x 64 tiles on every GEMM of the
 * stage, keeps the fastest (chosen[3]: the kind now used per stage, 1 =
 * 64 x 64, 2 128x64, 3 64x128 real part, 4 64x128, 5 64x160) and
 * leaves one execute's result in the outputs.  The field values do not depend
 * on the tile kind (the same K order per element).  Synchronous; no
 * counterpart in the reference.  This and rf_grid_batch_fuse_blend rewrite
 * the batch's device tables: a graph captured before either keeps its old
 * launch shapes and must be captured again. */
int rf_grid_batch_autotune(rf_grid_batch* batch, void* stream, int32_t* chosen);
/* The blend fused into the batch's last contraction (FIXED mode, a
 * tensor-product layout whose fields are the batch's outputs; same element
 * type): points that one field alone covers -- weight 1 -- are stored by the
 * last contraction straight into d_group, the others into the field as
 * before -- except the overlap bands along the first axis, which go to
 * arrays the blend plan allocates here, row by row, so that the blend reads
 * them as whole cache lines -- and rf_blend_plan_execute then forms only the
 * points several fields cover.  The group is the same as without fusion,
 * bit for bit; the field buffers then hold only the points of the overlaps
 * along the other axes (the single-cover interior and the first-axis bands
 * are neither written nor read there).  execute of the blend
 * plan must then be given d_group; d_group NULL undoes the fusion, which a
 * caller that destroys the batch and keeps using the blend plan must do
 * first (a fused blend plan forms the overlaps only).  Synchronous set-up;
 * the batch's tables are replaced. */
int rf_grid_batch_fuse_blend(rf_grid_batch* batch, rf_blend_plan* blend, void* d_group);
int rf_grid_batch_redraw(rf_grid_batch* batch, int64_t drawIndex, void* stream);
int rf_grid_batch_destroy(rf_grid_batch* batch);

/* Graph capture of the launches a thread enqueues on an explicit stream
 * between begin and end (only allocation-free entry points: the execute
 * functions of a batch or blend plan, rf_plan_execute of a FIXED-mode plan);
 * launch replays them on a stream.  No counterpart in the reference. */
typedef struct rf_graph rf_graph;
int rf_graph_capture_begin(rf_context* ctx, void* stream);
int rf_graph_capture_end(rf_context* ctx, void* stream, rf_graph** graph);
int rf_graph_launch(rf_graph* graph, void* stream);
int rf_graph_destroy(rf_graph* graph);

/* Per-realization FP64 mean/std (biased) and global mean/std; host outputs
 * avg[Nmc], std[Nmc], g[2] (synchronous). */
int rf_stats_device(rf_context* ctx, const void* d_field, int64_t N,
                    int32_t Nmc, int32_t dtype, double* avg, double* std,
                    double* g);
/* rebuild_Sk_FFT + rebuild_Sk + rebuild_corrL (statistics_RF.f90:180-443) of
 * one realization of a structured-grid field (x fastest; d_field holds
 * realizations of prod(nstep) points back to back).  For each axis a, R
 * receives lags 0 .. nstep[a]/2-1 of the line-averaged circular
 * autocorrelation of (u - avg)/std (what the reference's FFT, |.|^2, line
 * average and inverse FFT compute), axes back to back, and corrL[a] (may be
 * NULL) the reference's estimate 2 (sum R - (R[0] + R[last])/2) xStep[a].
 * FP64 sums; host outputs; synchronous. */
int rf_corr_device(rf_context* ctx, const void* d_field, int32_t dtype,
                   int32_t nDim, const int32_t* nstep, const double* xStep,
                   int32_t realization, double avg, double std, double* R,
                   double* corrL);
/* interpolateToMesh (sample_RF.f90:872-1106): multilinear interpolation of
 * a structured-grid field (nstep points per axis from xMin at xStep, x
 * fastest, Nmc realizations back to back) to nNodes mesh nodes d_nodes
 * (nDim x nNodes) for every realization (the reference interpolates the
 * first only); d_out: nNodes x Nmc.  The reference's cell rule and corner
 * order (border nodes use the last cell); nodes outside the grid get NaN
 * and set *outside (may be NULL).  Synchronous. */
int rf_interpolate_device(rf_context* ctx, int32_t nDim, const int32_t* nstep,
                          const double* xMin, const double* xStep, int32_t Nmc,
                          const void* d_grid, int32_t gridDtype,
                          const double* d_nodes, int64_t nNodes, void* d_out,
                          int32_t outDtype, void* stream, int32_t* outside);
/* interpolateToMesh on the bounding box it reads (sample_RF.f90:936-975):
 * d_box holds global grid points minPos..maxPos (1-based, inclusive, x
 * fastest, Nmc realizations back to back) of the grid that starts at xMinGlob;
 * positions are measured from xMinGlob and the cell rule's border is maxPos,
 * as in the reference.  Otherwise as rf_interpolate_device. */
int rf_interpolate_box_device(rf_context* ctx, int32_t nDim, const int32_t* minPos,
                              const int32_t* maxPos, const double* xMinGlob,
                              const double* xStep, int32_t Nmc, const void* d_box,
                              int32_t gridDtype, const double* d_nodes, int64_t nNodes,
                              void* d_out, int32_t outDtype, void* stream, int32_t* outside);
/* readUNV (readUNV_RF.f90:14-215, one rank): node coordinates of datasets
 * 2411/781 and connectivity of 2412/780 of a universal file.  Host only.
 * Files with element types other than FE descriptor 115 (8-node brick) are
 * refused (RF_E_UNSUPPORTED): the reference misreads their records. */
typedef struct rf_unv_mesh rf_unv_mesh;
int rf_unv_read(const char* path, int32_t nDim, rf_unv_mesh** out);
/* nNodes, nElem, maxConnect (= size(connectList,1)), monoType (0/1) */
int rf_unv_sizes(const rf_unv_mesh* m, int64_t* nNodes, int64_t* nElem,
                 int32_t* maxConnect, int32_t* monoType);
/* coordList (nDim x nNodes), connectList (maxConnect x nElem, node labels as
 * in the file, -1 padded), sizeList (nElem); any may be NULL. */
int rf_unv_copy(const rf_unv_mesh* m, double* coords, int32_t* connect, int32_t* sizeList);
int rf_unv_destroy(rf_unv_mesh* m);
/* Subtract each realization's mean and divide by its RMS (FP64 sums). */
int rf_normalize_device(rf_context* ctx, void* d_field, int64_t N,
                        int32_t Nmc, int32_t dtype, void* stream);

/* normalize_randField (randomFieldND.f90:820-926) as the reference computes
 * it: the mean and then the root mean square of realization 1 over the box
 * minPos..maxPos (1-based, inclusive; the points that are not repeated
 * between ranks, sample_RF.f90:829-842) of its nDim grid gridN (x fastest),
 * subtracted from / divided into realizations 1..Nmc of d_field (N x Nmc,
 * N = prod(gridN)).  Synchronous.  _host: on a host FP64 buffer. */
int rf_normalize_window_device(rf_context* ctx, void* d_field, int32_t nDim, const int32_t* gridN,
                               const int32_t* minPos, const int32_t* maxPos, int32_t Nmc, int32_t dtype,
                               void* stream);
int rf_normalize_window_host(rf_context* ctx, double* field, int32_t nDim, const int32_t* gridN,
                             const int32_t* minPos, const int32_t* maxPos, int32_t Nmc);

/* rf_normalize_device without allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that must not call hipMalloc).  Same arguments and
 * result as rf_normalize_device (same kernels and sum partition), but the
 * reduction partials, sums and coefficients live in a static device array of
 * the library (about 4.2 MB per device, held until the process ends), the
 * coefficients are formed on the device, and the call is stream-ordered and
 * asynchronous.  Nmc <= RF_NOALLOC_MAX_NMC, else RF_E_ARG.  Every call on a
 * device shares that device's array: calls that overlap on one device (other
 * streams or host threads) race. */
#define RF_NOALLOC_MAX_NMC 1024
int rf_normalize_device_noalloc(rf_context* ctx, void* d_field, int64_t N, int32_t Nmc,
                                int32_t dtype, void* stream);

/* Device generator: count uniforms U[0,1) of stream (seed, tag) into d_out
 * (FP64), element e = Philox4x32-10(counter=(e lo, e hi, tag, 0),
 * key=(seed lo, seed hi)), 53-bit mantissa from words 0..1. */
int rf_device_uniforms(rf_context* ctx, uint64_t seed, uint32_t tag,
                       int64_t count, double* d_out, void* stream);
/* Host mirror of that stream (element e), and the raw generator. */
double rf_uniform_host(uint64_t seed, uint32_t tag, uint64_t e);
void   rf_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2],
                             uint32_t out[4]);
/* Streams of the device generator: wavevector draws use tag 1, phase draws
 * tag 2 (element = realization * nPhaseUniform + index). */

/* ---- serving: field requests with continuous batching ----------------------
 * No counterpart in the reference, whose unit of work is one
 * createRandomField call per MPI rank; a server answers many such calls
 * (single_realization, sample_RF.f90:309-372, as requests) on one GPU.
 *
 * Every request is a 3-D FIXED-mode field on the server's lattice: the
 * Shinozuka lattice (method 2, MFMA form of the server) or the isotropic
 * shells (method 1, FP32 direct form with per-realization directions) of the
 * normalised extent cfg.xRangeNorm (corrL 1).  A request's points are
 * divided by its own corrL, so a request whose box is xRangeNorm * corrL
 * gets exactly the field rf_plan_create would build for it.  Phases come
 * from the device generator (seed, tag 2), as in rf_plan_create with
 * phaseUniform NULL.
 *
 * A worker thread admits queued requests in arrival order while their
 * device memory fits the server's budget (one pre-allocated block, HBM
 * admission), groups them into batches (one grouped launch per method),
 * and keeps up to cfg.maxInflight batches on the device at once, each on
 * its own stream: requests that arrive while others run are admitted as
 * memory and a stream slot free up, and their work overlaps.  A request's
 * result depends only on the request (per-point operation order is fixed),
 * not on what it was batched with.  Outputs stay in the server's memory
 * until rf_server_release. */
typedef struct rf_server rf_server;

typedef struct rf_server_config {
    int64_t hbmBudget;     /* bytes the server may hold; 0 = half the free memory */
    int32_t maxInflight;   /* batches on the device at once, 1..8; 0 = 3 */
    int32_t maxBatch;      /* requests per batch; 0 = 64 */
    int32_t form;          /* method-2 form: RF_FORM_MFMA_F16 (0 = default) or _BF16 */
    int32_t maxNmc;        /* largest Nmc accepted; 0 = 1024 */
    double  xRangeNorm[3]; /* lattice extent in correlation lengths; 0 = 10 */
    double  batchTerms;    /* batch closes above this many N*Nmc*N_k terms; 0 = 2e11 */
    int32_t reserved[8];
} rf_server_config;

typedef struct rf_request {
    int32_t method;        /* RF_SHINOZUKA or RF_ISOTROPIC */
    int32_t Nmc;
    int32_t margiFirst;    /* RF_FOM_* (lognormal needs fieldAvg > 0) */
    int32_t xOnDevice;     /* 1: x is a device pointer, 0: host (copied at launch) */
    int64_t N;
    const double* x;       /* 3 x N (point-major), valid until the request completes */
    double corrL[3];
    double fieldAvg, fieldVar;
    uint64_t seed;
    int32_t reserved[6];
} rf_request;

typedef struct rf_request_result {
    int32_t status;        /* 0 done, 1 queued/running, < 0 failed (RF_E_*) */
    int32_t Nmc;
    int64_t N;
    const float* d_out;    /* device, N x Nmc FP32 (realization-major rows of N) */
    double tSubmit, tLaunch, tDone;   /* steady-clock seconds */
    int64_t batch;         /* batch the request ran in */
} rf_request_result;

typedef struct rf_server_stats {
    int64_t submitted, completed, failed, batches;
    int64_t arenaCapacity, arenaUsed, arenaPeak;
    int32_t maxBatchSeen, maxInflightSeen;
    double  pointRealizations;   /* sum of N * Nmc over completed requests */
} rf_server_stats;

int rf_server_create(rf_context* ctx, const rf_server_config* cfg, rf_server** server);
/* Fails queued requests, waits for the batches on the device, frees all. */
int rf_server_destroy(rf_server* server);
/* Thread-safe.  RF_E_NOMEM if the request can never fit the budget. */
int rf_server_submit(rf_server* server, const rf_request* req, int64_t* id);
/* Blocks until request id completes (timeoutSec < 0: no limit); status 1 in
 * res on time-out.  RF_E_ARG for an unknown or released id. */
int rf_server_wait(rf_server* server, int64_t id, double timeoutSec, rf_request_result* res);
/* Copies a completed request's field to a host buffer (N x Nmc FP32). */
int rf_server_read(rf_server* server, int64_t id, float* host_out);
/* Frees a completed request's output; its id becomes unknown. */
int rf_server_release(rf_server* server, int64_t id);
int rf_server_get_stats(rf_server* server, rf_server_stats* stats);

#ifdef __cplusplus
}
#endif
#endif

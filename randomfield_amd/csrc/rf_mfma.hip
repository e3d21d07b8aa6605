// rf_mfma.hip -- the spectral sum as a dense contraction on the matrix cores.
//
//   u[r][i] = sum_j  C[r][j] cos(2 pi t_ij) + S[r][j] sin(2 pi t_ij),
//   t_ij = (k_j / 2pi) . x_i  (revolutions),  C = A cos(phi), S = -A sin(phi)
//
// i.e. D (realizations x points) = P (realizations x 2N_k) . Phi (2N_k x points),
// the rewrite of randomFieldND.f90:276-330 with the realization loop batched.
//
// * Phi, the basis, is never stored: every wave produces the fragments it
//   consumes.  Lane l owns point l%32 and the wavevectors 4h..4h+3 (h = l/32)
//   of each 8-wavevector K-step; it forms k.x in FP64, reduces it to
//   revolutions (v_fract_f64), takes the hardware sin/cos (v_sin_f32 and
//   v_cos_f32 take revolutions) and splits each value into hi + lo halves.
// * P, the coefficient panels, is built once per plan in FP64 and stored
//   pre-split (hi, lo), pre-scaled by a power of two, in exactly the per-lane
//   order of the MFMA A operand.  Each K-step's block -- its 8 wavevectors
//   and the panels -- is contiguous in HBM and goes to LDS by direct
//   global->LDS loads (global_load_lds_dwordx4) one step ahead of use, into a
//   double buffer; one barrier per step.  K order inside a step is
//   (wavevector q, cos/sin), identical on both operands, so any hardware K
//   permutation cancels.
// * D += P_hi Phi_hi + P_hi Phi_lo + P_lo Phi_hi with FP32 accumulation
//   (v_mfma_f32_32x32x16_{f16,bf16}); the dropped lo.lo term is 2^-22 (FP16)
//   relative.  A wave holds 32 points x (32*MT) realizations: 16*MT
//   accumulators per lane.
#include "rf_internal.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16x __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int kMaxMT = 8;               // up to 256 realizations per block
constexpr int kStepK = 8;               // wavevectors per K-step (16 K entries)
constexpr int kKPad = 4 * kStepK;       // K padding: whole stages of up to 4 steps

// launch variants: waves per workgroup (points per block = 32 * waves),
// K-steps per LDS stage (one barrier per stage) and MFMA issue order (0:
// tiles in pairs, the three products of a pair interleaved; 1: one product
// over all tiles, then the next; 2: as 0 with the wavevectors read from LDS)
struct Variant { int waves, steps, order; };
constexpr Variant kVariants[] = {{4, 1, 0}, {4, 2, 0}, {8, 1, 0}, {8, 2, 0}, {8, 4, 0}, {8, 2, 1}, {8, 4, 1},
                                 {8, 2, 2}, {8, 4, 2}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
static_assert(kNumVariants == RF_MFMA_VARIANTS, "variant count of the C ABI");

template <typename V> struct Elem;
template <> struct Elem<h8> { typedef _Float16 T; };
template <> struct Elem<b8> { typedef __bf16 T; };

__device__ inline f16x mfma(const h8& a, const h8& b, const f16x& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ inline f16x mfma(const b8& a, const b8& b, const f16x& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

typedef __attribute__((address_space(3))) void lds_void;

// (c, s) -> hi and lo halves at positions 2q, 2q+1 of the basis fragments.
// FP16: one packed conversion for hi, the residuals v - hi by mixed-precision
// FMAs reading the f16 halves in place (v_fma_mix_f32), one packed conversion
// for lo.  BF16: plain conversions.
template <typename V>
__device__ inline void split_pair(float c, float s, int q, V* bh, V* bl) {
    typedef typename Elem<V>::T T;
    const T ch = (T)c, sh = (T)s;
    (*bh)[2 * q] = ch;
    (*bh)[2 * q + 1] = sh;
    (*bl)[2 * q] = (T)(c - (float)ch);
    (*bl)[2 * q + 1] = (T)(s - (float)sh);
}
template <>
__device__ inline void split_pair<h8>(float c, float s, int q, h8* bh, h8* bl) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 hi = {(_Float16)c, (_Float16)s};        // v_cvt_pk_f16_f32 (RNE)
    const unsigned hb = __builtin_bit_cast(unsigned, hi);
    float lc, ls;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[0,0,0] op_sel_hi:[1,0,0]" : "=v"(lc) : "v"(hb), "v"(c));
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(ls) : "v"(hb), "v"(s));
    (*bh)[2 * q] = hi.x;
    (*bh)[2 * q + 1] = hi.y;
    (*bl)[2 * q] = (_Float16)lc;
    (*bl)[2 * q + 1] = (_Float16)ls;
}

// Basis fragments of a lane: point x, the wavevectors 4h..4h+3 of the step
// whose 8 wavevectors (x,y,z in revolutions, 24 doubles) start at k.  The
// wavevectors are wave-uniform (scalar loads); both halves' dot products are
// formed and the lane's half selected, which is cheaper than selecting the
// 64-bit operands.
template <typename V, int D>
__device__ inline void basis(const double* __restrict__ k, int h, const double* xn, V* bh, V* bl) {
    typedef typename Elem<V>::T T;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double s0 = k[3 * q] * xn[0];
        double s1 = k[3 * (q + 4)] * xn[0];
        if (D > 1) {
            s0 = fma(k[3 * q + 1], xn[1], s0);
            s1 = fma(k[3 * (q + 4) + 1], xn[1], s1);
        }
        if (D > 2) {
            s0 = fma(k[3 * q + 2], xn[2], s0);
            s1 = fma(k[3 * (q + 4) + 2], xn[2], s1);
        }
        const double s = h ? s1 : s0;
        const float f = (float)__builtin_amdgcn_fract(s);
        const float c = __builtin_amdgcn_cosf(f);
        const float sn = __builtin_amdgcn_sinf(f);
        split_pair<V>(c, sn, q, bh, bl);
    }
}

// The same basis from the lane's own half of a step's wavevectors, read from
// LDS (kslot: the step's 24 doubles; a lane in half h takes doubles 12h ..
// 12h+11, a broadcast read): one dot product per wavevector and no select.
template <typename V, int D>
__device__ inline void basis_lds(const u32x4* kslot, int h, const double* xn, V* bh, V* bl) {
    d2 k2[6];
#pragma unroll
    for (int i = 0; i < 6; i++) k2[i] = __builtin_bit_cast(d2, kslot[6 * h + i]);
    const double* k = reinterpret_cast<const double*>(k2);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double s = k[3 * q] * xn[0];
        if (D > 1) s = fma(k[3 * q + 1], xn[1], s);
        if (D > 2) s = fma(k[3 * q + 2], xn[2], s);
        const float f = (float)__builtin_amdgcn_fract(s);
        const float c = __builtin_amdgcn_cosf(f);
        const float sn = __builtin_amdgcn_sinf(f);
        split_pair<V>(c, sn, q, bh, bl);
    }
}

// Panels in HBM (16-byte units): [t][mtile][hilo][lane]; LDS double buffer of
// S K-steps x MT x 2 x 64 units.  Wavevectors: separate stream [t][8][3]
// doubles read with scalar loads, so the LDS serves only the coefficient
// fragments and the only vector-memory traffic in the loop is the panel DMA.
// The basis of step t+1 is formed while the MFMAs of step t run.
// One block's work: points pt0 .. pt0 + 32W - 1 (clipped to a.i1) and the
// M-tiles mt0 .. mt0 + MT - 1.  ldsA and ldsB are two distinct __shared__
// objects of S * MT * 128 units or more (distinct, so the waitcnt pass can
// tell the DMA target of the next stage from the buffer being read).
// KS = 0: each step's wavevectors come through scalar loads; KS = 1: every
// wave copies a stage's wavevectors into its own LDS slot (kqA/kqB, W x 12S
// units each) with the stage's DMA and reads its lane half from there.
template <typename V, int D, int MT, int W, int S, int ORD, int KS>
__device__ __forceinline__ void sum_mfma_block(const rf_sum_args& a, int64_t nSteps, int mtTotal,
                                               const u32x4* __restrict__ panels,
                                               const u32x4* __restrict__ kq, float outScale,
                                               int64_t pt0, int mt0, u32x4* ldsA, u32x4* ldsB,
                                               u32x4* kqA, u32x4* kqB) {
    constexpr int kStepUnits = MT * 2 * 64;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int h = lane >> 5;
    const int n = lane & 31;
    const int64_t hbmStepUnits = (int64_t)mtTotal * 128;
    const int64_t pt = pt0 + wave * 32 + n;
    const int64_t ptc = pt < a.i1 ? pt : a.i1 - 1;
    const int64_t nStages = nSteps / S;             // nSteps is a multiple of S

    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = a.x[ptc * D + d] / a.corrL[d];

    f16x acc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;

    // direct global->LDS copy of stage u (S steps x 2*MT chunks of 64 units)
    // spread over the waves (static trip count, so waits on older loads need
    // not drain the DMA)
    constexpr int kChunks = 2 * MT * S;
    constexpr int kDmaPerWave = (kChunks + W - 1) / W;
    static_assert(kDmaPerWave < 16, "vmcnt field of the wavevector wait");
    auto dma_stage = [&](int64_t u, u32x4* dst) {
#pragma unroll
        for (int i = 0; i < kDmaPerWave; i++) {
            const int c = wave + i * W;
            if (kChunks >= W * (i + 1) || c < kChunks) {
                const int sI = c / (2 * MT), cI = c % (2 * MT);
                const u32x4* src = panels + (u * S + sI) * hbmStepUnits + (int64_t)mt0 * 128;
                __builtin_amdgcn_global_load_lds((const void*)(src + cI * 64 + lane),
                                                 (lds_void*)&dst[sI * kStepUnits + cI * 64], 16, 0, 0);
            }
        }
    };

    // this wave's copy of stage u's wavevectors (12 units per step)
    auto dma_kq = [&](int64_t u, u32x4* dst) {
        if (KS == 1 && lane < 12 * S)
            __builtin_amdgcn_global_load_lds((const void*)(kq + u * S * 12 + lane),
                                             (lds_void*)&dst[wave * 12 * S], 16, 0, 0);
    };

    V bh, bl;
    const double* kqd = reinterpret_cast<const double*>(kq);
    // one stage: prefetch stage u+1 into nxt; per step, form the basis of the
    // next step while the MFMAs of this step read cur
    auto stage = [&](int64_t u, const u32x4* cur, u32x4* nxt, const u32x4* kcur, u32x4* knxt) {
        // unconditional (the last stage re-reads itself into the unused
        // buffer), so every path leaves the same number of loads in flight;
        // the wavevector copy goes first so a wait can single it out
        const int64_t un = u + 1 < nStages ? u + 1 : u;
        dma_kq(un, knxt);
        dma_stage(un, nxt);
#pragma unroll
        for (int sI = 0; sI < S; sI++) {
            const int64_t t = u * S + sI;
            const u32x4* c = cur + sI * kStepUnits;
            V nbh, nbl;
            if (KS == 1) {
                if (sI + 1 < S) {
                    basis_lds<V, D>(kcur + wave * 12 * S + (sI + 1) * 12, h, xn, &nbh, &nbl);
                } else {
                    // next stage's first step: this wave's own copy, issued
                    // before the kDmaPerWave panel loads of this stage
                    __builtin_amdgcn_s_waitcnt(0x0F70 | kDmaPerWave);
                    basis_lds<V, D>(knxt + wave * 12 * S, h, xn, &nbh, &nbl);
                }
            } else {
                basis<V, D>(kqd + (t + 1 < nSteps ? t + 1 : t) * 24, h, xn, &nbh, &nbl);
            }
            if (ORD == 1) {
                V ah[MT], al[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    ah[m] = __builtin_bit_cast(V, c[(m * 2 + 0) * 64 + lane]);
                    al[m] = __builtin_bit_cast(V, c[(m * 2 + 1) * 64 + lane]);
                }
#pragma unroll
                for (int m = 0; m < MT; m++) acc[m] = mfma(ah[m], bh, acc[m]);
#pragma unroll
                for (int m = 0; m < MT; m++) acc[m] = mfma(ah[m], bl, acc[m]);
#pragma unroll
                for (int m = 0; m < MT; m++) acc[m] = mfma(al[m], bh, acc[m]);
            } else
#pragma unroll
            for (int m = 0; m < MT; m += 2) {
                const V ah0 = __builtin_bit_cast(V, c[(m * 2 + 0) * 64 + lane]);
                const V al0 = __builtin_bit_cast(V, c[(m * 2 + 1) * 64 + lane]);
                if (m + 1 < MT) {
                    const V ah1 = __builtin_bit_cast(V, c[(m * 2 + 2) * 64 + lane]);
                    const V al1 = __builtin_bit_cast(V, c[(m * 2 + 3) * 64 + lane]);
                    acc[m] = mfma(ah0, bh, acc[m]);
                    acc[m + 1] = mfma(ah1, bh, acc[m + 1]);
                    acc[m] = mfma(ah0, bl, acc[m]);
                    acc[m + 1] = mfma(ah1, bl, acc[m + 1]);
                    acc[m] = mfma(al0, bh, acc[m]);
                    acc[m + 1] = mfma(al1, bh, acc[m + 1]);
                } else {
                    acc[m] = mfma(ah0, bh, acc[m]);
                    acc[m] = mfma(ah0, bl, acc[m]);
                    acc[m] = mfma(al0, bh, acc[m]);
                }
            }
            bh = nbh;
            bl = nbl;
        }
        // this wave's global->LDS copies are complete (vmcnt(0); expcnt and
        // lgkmcnt untouched) before the barrier publishes them to the others
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();   // stage u+1 landed in LDS; stage u's buffer is free
    };

    dma_kq(0, kqA);
    dma_stage(0, ldsA);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (KS == 1) basis_lds<V, D>(kqA + wave * 12 * S, h, xn, &bh, &bl);
    else basis<V, D>(kqd, h, xn, &bh, &bl);
    __syncthreads();
    for (int64_t u = 0; u < nStages; u += 2) {
        stage(u, ldsA, ldsB, kqA, kqB);
        if (u + 1 < nStages) stage(u + 1, ldsB, ldsA, kqB, kqA);
    }
    if (pt >= a.i1) return;
    // D layout of a 32x32 tile: lane column n, rows 8*(r/4) + 4*h + r%4.
    // The marginal map and the output type are uniform: one store loop each,
    // so the common cases carry no exp evaluation or dtype select per value.
    auto store = [&](auto map, auto* out) {
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int real = (mt0 + m) * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                if (real < a.Nmc) out[(int64_t)real * a.N + pt] = map(acc[m][r] * outScale);
            }
        }
    };
    const float sc = (float)a.epi.scale, sh = (float)a.epi.shift;
    auto none = [](float u) { return u; };
    auto gauss = [=](float u) { return fmaf(u, sc, sh); };
    auto logn = [=](float u) { return expf(fmaf(u, sc, sh)); };
    auto none64 = [](float u) { return (double)u; };
    auto gauss64 = [=](float u) { return (double)fmaf(u, sc, sh); };
    auto logn64 = [=](float u) { return (double)expf(fmaf(u, sc, sh)); };
    float* o32 = static_cast<float*>(a.out);
    double* o64 = static_cast<double*>(a.out);
    if (a.outDtype == RF_DTYPE_F64) {
        if (a.epi.margi == RF_FOM_NONE) store(none64, o64);
        else if (a.epi.margi == RF_FOM_LOGNORMAL) store(logn64, o64);
        else store(gauss64, o64);
    } else {
        if (a.epi.margi == RF_FOM_NONE) store(none, o32);
        else if (a.epi.margi == RF_FOM_LOGNORMAL) store(logn, o32);
        else store(gauss, o32);
    }
}


template <typename V, int D, int MT, int W, int S, int ORD, int KS>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_sum_mfma(rf_sum_args a, int64_t nSteps, int mtTotal, const u32x4* __restrict__ panels,
                const u32x4* __restrict__ kq, float outScale) {
    __shared__ u32x4 ldsA[S * MT * 128];
    __shared__ u32x4 ldsB[S * MT * 128];
    __shared__ u32x4 kqA[KS ? W * 12 * S : 1];
    __shared__ u32x4 kqB[KS ? W * 12 * S : 1];
    sum_mfma_block<V, D, MT, W, S, ORD, KS>(a, nSteps, mtTotal, panels, kq, outScale,
                                            a.i0 + (int64_t)blockIdx.x * (32 * W), blockIdx.y * kMaxMT, ldsA,
                                            ldsB, kqA, kqB);
}

// Grouped form for a batch of independent fields (requests of a serving
// batch): block b runs one tile = (field, point block, block of up to
// kGroupMT M-tiles).  Each field's points, panels and output are its own;
// the per-point operation order is that of k_sum_mfma, so a field's result
// does not depend on what it is batched with.
__device__ inline int group_find(const rf_group_req* __restrict__ reqs, int nReq, int64_t b) {
    int lo = 0, hi = nReq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename V, int W, int S>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_group_mfma(const rf_group_req* __restrict__ reqs, int nReq) {
    __shared__ u32x4 ldsA[S * kGroupMT * 128];
    __shared__ u32x4 ldsB[S * kGroupMT * 128];
    __shared__ u32x4 kqA[W * 12 * S];
    __shared__ u32x4 kqB[W * 12 * S];
    const int64_t b = blockIdx.x;
    const rf_group_req& q = reqs[group_find(reqs, nReq, b)];
    const int nSub = (q.mtTotal + kGroupMT - 1) / kGroupMT;
    const int64_t local = b - q.tile0;
    const int64_t pt0 = (local / nSub) * (32 * W);
    const int mt0 = (int)(local % nSub) * kGroupMT;
    const int mt = q.mtTotal - mt0 < kGroupMT ? q.mtTotal - mt0 : kGroupMT;
    const rf_sum_args a = q.a;
    const u32x4* P = static_cast<const u32x4*>(q.ops0);
    const u32x4* K = static_cast<const u32x4*>(q.ops1);
    // wavevectors staged in LDS: with at most kGroupMT tiles per block the
    // basis, not the matrix cores, bounds a step, and this form spends the
    // least VALU on it
    switch (mt) {
        case 1: sum_mfma_block<V, 3, 1, W, S, 0, 1>(a, q.nSteps, q.mtTotal, P, K, q.outScale, pt0, mt0, ldsA, ldsB, kqA, kqB); break;
        case 2: sum_mfma_block<V, 3, 2, W, S, 0, 1>(a, q.nSteps, q.mtTotal, P, K, q.outScale, pt0, mt0, ldsA, ldsB, kqA, kqB); break;
        case 3: sum_mfma_block<V, 3, 3, W, S, 0, 1>(a, q.nSteps, q.mtTotal, P, K, q.outScale, pt0, mt0, ldsA, ldsB, kqA, kqB); break;
        default: sum_mfma_block<V, 3, kGroupMT, W, S, 0, 1>(a, q.nSteps, q.mtTotal, P, K, q.outScale, pt0, mt0, ldsA, ldsB, kqA, kqB); break;
    }
}

// Coefficient panels from FP64 phases and amplitudes, one thread per (step,
// mtile, lane, element) writing both halves; the first 24 threads of each
// step also write its 8 wavevectors (revolutions, 3 components, zero padded)
// to the wavevector stream kqd[j*3 + d].
template <typename T>
__global__ void k_mfma_prep(int nDim, int64_t kN, int64_t nSteps, int Nmc, int mtTotal,
                            const double* __restrict__ kRev, const double* __restrict__ phase,
                            const double* __restrict__ amp, int64_t ampStrideR, double scale,
                            T* __restrict__ panels, double* __restrict__ kqd) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t perStep = (int64_t)mtTotal * 64 * 8;
    if (e >= nSteps * perStep) return;
    const int64_t t = e / perStep;
    const int64_t rem = e - t * perStep;
    if (rem < kStepK * 3) {
        const int64_t j = t * kStepK + rem / 3;
        const int d = (int)(rem % 3);
        kqd[j * 3 + d] = (j < kN && d < nDim) ? kRev[j * nDim + d] : 0.0;
    }
    const int el = (int)(rem & 7);
    const int lane = (int)((rem >> 3) & 63);
    const int mt = (int)(rem >> 9);
    const int h = lane >> 5;
    const int r = mt * 32 + (lane & 31);
    const int q = el >> 1;
    const int64_t j = t * kStepK + 4 * h + q;
    double v = 0.0;
    if (r < Nmc && j < kN) {
        const double A = amp[(int64_t)r * ampStrideR + j] * scale;
        const double ph = phase[(int64_t)r * kN + j];
        v = (el & 1) ? -A * sin(ph) : A * cos(ph);
    }
    const T hi = (T)(float)v;
    const T lo = (T)(float)(v - (double)(float)hi);
    T* pan = panels + ((t * mtTotal + mt) * 2) * 64 * 8;
    pan[(int64_t)lane * 8 + el] = hi;
    pan[64 * 8 + (int64_t)lane * 8 + el] = lo;
}

template <typename V, int D, int W, int S, int ORD>
hipError_t launch_mt(int mt, const rf_sum_args& a, int64_t nSteps, int mtTotal, const void* B,
                     float outScale, hipStream_t s) {
    constexpr int O = ORD == 2 ? 0 : ORD, KS = ORD == 2 ? 1 : 0;
    const u32x4* P = static_cast<const u32x4*>(B);
    const u32x4* K = P + nSteps * mtTotal * 128;
    const int64_t n = a.i1 - a.i0;
    const dim3 grid((unsigned)((n + 32 * W - 1) / (32 * W)), (unsigned)((mtTotal + kMaxMT - 1) / kMaxMT));
    const dim3 block(64 * W);
    switch (mt) {
        case 1: hipLaunchKernelGGL((k_sum_mfma<V, D, 1, W, S, O, KS>), grid, block, 0, s, a, nSteps, mtTotal, P, K, outScale); break;
        case 2: hipLaunchKernelGGL((k_sum_mfma<V, D, 2, W, S, O, KS>), grid, block, 0, s, a, nSteps, mtTotal, P, K, outScale); break;
        case 4: hipLaunchKernelGGL((k_sum_mfma<V, D, 4, W, S, O, KS>), grid, block, 0, s, a, nSteps, mtTotal, P, K, outScale); break;
        default: hipLaunchKernelGGL((k_sum_mfma<V, D, 8, W, S, O, KS>), grid, block, 0, s, a, nSteps, mtTotal, P, K, outScale); break;
    }
    return hipGetLastError();
}

template <typename V, int D>
hipError_t launch_variant(int variant, int mt, const rf_sum_args& a, int64_t nSteps, int mtTotal,
                          const void* B, float outScale, hipStream_t s) {
    static_assert(kNumVariants == 9, "variant table and dispatch agree");
    switch (variant) {
        case 1: return launch_mt<V, D, 4, 2, 0>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 2: return launch_mt<V, D, 8, 1, 0>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 3: return launch_mt<V, D, 8, 2, 0>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 4: return launch_mt<V, D, 8, 4, 0>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 5: return launch_mt<V, D, 8, 2, 1>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 6: return launch_mt<V, D, 8, 4, 1>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 7: return launch_mt<V, D, 8, 2, 2>(mt, a, nSteps, mtTotal, B, outScale, s);
        case 8: return launch_mt<V, D, 8, 4, 2>(mt, a, nSteps, mtTotal, B, outScale, s);
        default: return launch_mt<V, D, 4, 1, 0>(mt, a, nSteps, mtTotal, B, outScale, s);
    }
}

}  // namespace

// The six (element type, dimension) instantiations of the point-wise kernel
// are separate translation units of this file (RF_MFMA_PART 1..6; part 0 holds
// everything else), so the build compiles them in parallel.
#ifndef RF_MFMA_PART
#error "rf_mfma.hip is compiled once per RF_MFMA_PART (0..6)"
#endif
#define RF_MFMA_LAUNCH(NAME)                                                                      \
    hipError_t NAME(int variant, int mt, const rf_sum_args& a, int64_t nSteps, int mtTotal,       \
                    const void* B, float outScale, hipStream_t s)
RF_MFMA_LAUNCH(rf_mfma_launch_f16_1);
RF_MFMA_LAUNCH(rf_mfma_launch_f16_2);
RF_MFMA_LAUNCH(rf_mfma_launch_f16_3);
RF_MFMA_LAUNCH(rf_mfma_launch_bf16_1);
RF_MFMA_LAUNCH(rf_mfma_launch_bf16_2);
RF_MFMA_LAUNCH(rf_mfma_launch_bf16_3);
#define RF_MFMA_DEFINE(PART, NAME, V, D) \
    RF_MFMA_LAUNCH(NAME) { return launch_variant<V, D>(variant, mt, a, nSteps, mtTotal, B, outScale, s); }
#if RF_MFMA_PART == 1
RF_MFMA_DEFINE(1, rf_mfma_launch_f16_1, h8, 1)
#elif RF_MFMA_PART == 2
RF_MFMA_DEFINE(2, rf_mfma_launch_f16_2, h8, 2)
#elif RF_MFMA_PART == 3
RF_MFMA_DEFINE(3, rf_mfma_launch_f16_3, h8, 3)
#elif RF_MFMA_PART == 4
RF_MFMA_DEFINE(4, rf_mfma_launch_bf16_1, b8, 1)
#elif RF_MFMA_PART == 5
RF_MFMA_DEFINE(5, rf_mfma_launch_bf16_2, b8, 2)
#elif RF_MFMA_PART == 6
RF_MFMA_DEFINE(6, rf_mfma_launch_bf16_3, b8, 3)
#endif

#if RF_MFMA_PART == 0
namespace {

hipError_t launch_form(bool bf16, int variant, const rf_sum_args& a, int64_t kNpad, int Npad, const void* B,
                       double outScale, hipStream_t s) {
    const int mtTotal = Npad / 32;
    const int mt = mtTotal >= kMaxMT ? kMaxMT : mtTotal;
    const int64_t nSteps = kNpad / kStepK;
    const float sc = (float)outScale;
    switch (a.nDim) {
        case 1: return (bf16 ? rf_mfma_launch_bf16_1 : rf_mfma_launch_f16_1)(variant, mt, a, nSteps, mtTotal, B, sc, s);
        case 2: return (bf16 ? rf_mfma_launch_bf16_2 : rf_mfma_launch_f16_2)(variant, mt, a, nSteps, mtTotal, B, sc, s);
        default: return (bf16 ? rf_mfma_launch_bf16_3 : rf_mfma_launch_f16_3)(variant, mt, a, nSteps, mtTotal, B, sc, s);
    }
}

}  // namespace

int rf_mfma_kpad(int) { return kKPad; }
int rf_mfma_num_variants() { return kNumVariants; }

// Realizations are padded to whole 32-row tiles; above 256 to whole blocks of
// 8 tiles, and below that to a power-of-two tile count (1, 2, 4, 8).
int rf_mfma_npad(int Nmc) {
    int mt = (Nmc + 31) / 32;
    if (mt > kMaxMT) return ((mt + kMaxMT - 1) / kMaxMT) * kMaxMT * 32;
    int p = 1;
    while (p < mt) p <<= 1;
    return p * 32;
}

// panels [steps][mtiles][hilo][64] x 16 B, then the wavevector stream
// (3 doubles per padded wavevector)
size_t rf_mfma_bbytes(int, int64_t kNpad, int Npad) {
    return (size_t)(kNpad / kStepK) * (size_t)(Npad / 32) * 128 * 16 + (size_t)kNpad * 3 * 8;
}

hipError_t rf_launch_mfma_prep(int form, int nDim, int64_t kN, int64_t kNpad, int Nmc, int Npad,
                               const double* d_kRev, const double* d_phase, const double* d_amp,
                               int64_t ampStrideR, double scale, void* d_B, hipStream_t s) {
    const int64_t nSteps = kNpad / kStepK;
    const int mtTotal = Npad / 32;
    const int64_t total = nSteps * mtTotal * 64 * 8;
    const unsigned grid = (unsigned)((total + 255) / 256);
    char* base = static_cast<char*>(d_B);
    double* kqd = reinterpret_cast<double*>(base + (size_t)nSteps * mtTotal * 128 * 16);
    if (form == RF_FORM_MFMA_BF16)
        hipLaunchKernelGGL(k_mfma_prep<__bf16>, dim3(grid), dim3(256), 0, s, nDim, kN, nSteps, Nmc, mtTotal,
                           d_kRev, d_phase, d_amp, ampStrideR, scale, reinterpret_cast<__bf16*>(base), kqd);
    else
        hipLaunchKernelGGL(k_mfma_prep<_Float16>, dim3(grid), dim3(256), 0, s, nDim, kN, nSteps, Nmc, mtTotal,
                           d_kRev, d_phase, d_amp, ampStrideR, scale, reinterpret_cast<_Float16*>(base), kqd);
    return hipGetLastError();
}

hipError_t rf_launch_group_mfma(int form, const rf_group_req* d_reqs, int nReq, int64_t nTiles,
                                hipStream_t s) {
    if (nTiles <= 0 || nReq <= 0) return hipSuccess;
    if (nTiles > 0x7fffffffLL) return hipErrorInvalidValue;
    if (form == RF_FORM_MFMA_BF16)
        hipLaunchKernelGGL((k_group_mfma<b8, kGroupW, 4>), dim3((unsigned)nTiles), dim3(64 * kGroupW), 0, s, d_reqs, nReq);
    else
        hipLaunchKernelGGL((k_group_mfma<h8, kGroupW, 4>), dim3((unsigned)nTiles), dim3(64 * kGroupW), 0, s, d_reqs, nReq);
    return hipGetLastError();
}

hipError_t rf_launch_sum_mfma(int form, int variant, const rf_sum_args& a, int64_t kNpad, int Npad,
                              const void* d_B, double outScale, hipStream_t s) {
    if (a.i1 <= a.i0) return hipSuccess;
    if (kNpad % kKPad != 0 || variant < 0 || variant >= kNumVariants) return hipErrorInvalidValue;
    return launch_form(form == RF_FORM_MFMA_BF16, variant, a, kNpad, Npad, d_B, outScale, s);
}
#endif  // RF_MFMA_PART == 0

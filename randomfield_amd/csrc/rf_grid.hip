// rf_grid.hip -- Shinozuka sums on structured grids by axis-wise contraction.
//
// On a grid x = o + (i1, i2, i3) * step (x fastest, setGrid order,
// math_RF.f90:70-88) and the Shinozuka lattice -- per sign block b the tensor
// product k = (s1 k1_j1, s2 k2_j2, s3 k3_j3) of 1-D lattices k_a,j = j dk_a
// (spectra_RF.f90:88-133) -- the sum of randomFieldND.f90:276-330 factorises:
//
//   u_r(x) = Re sum_b sum_j3 E3^{s3}[j3][x3] sum_j2 E2^{s2}[j2][x2]
//                        sum_j1 E1[j1][x1] C_r[b][j3][j2][j1],
//   E_a[j][x] = exp(i k_a,j x_a),  C = ampMult A exp(i phi),
//
// three small batched complex GEMMs instead of N * N_k cosines
// (O(n N_k + n^2 N_k^(2/3) + n^3 N_k^(1/3)) per realization and block).
// Every stage is one tiled complex GEMM on the matrix cores,
// D[m][n] = sum_k A[m][k] B[k][n] with per-stage strides: operands split
// into FP16 hi + lo halves (power-of-two scaled into range), three
// v_mfma_f32_32x32x16_f16 products with FP32 accumulation, on 64x64 tiles or
// the wide kinds along each stage's long axis; an FP32-operand MFMA tile
// (v_mfma_f32_32x32x2_f32) serves shapes whose K would leave the FP16 range.
// The negative sign blocks are folded into the first contraction
// (conjugated tables and coefficients).  The phase tables are formed in FP64
// and rounded to FP32; accumulation is FP32 over <= 4*N_k^(1/3) terms.
//
// The isotropic method (randomFieldND.f90:368-514) has no lattice, but on a
// grid each of its terms still factorises over the axes: per realization one
// real-part GEMM over the shells whose operands are per-axis phase tables
// (k_iso_tables, k_iso_gemm).
#include "rf_internal.h"
#include "rf_philox.h"

namespace {

constexpr int kTM = 64, kTN = 64, kTK = 16;   // block tile; 256 threads, 4x4 each

// element offsets of column n in B and in D (two-level when nLo > 0)
__device__ __forceinline__ int64_t col_b(const rf_gemm_desc& g, int n) {
    return g.nLo ? (int64_t)(n / g.nLo) * g.bSnHi + (int64_t)(n % g.nLo) * g.bSn : (int64_t)n * g.bSn;
}
__device__ __forceinline__ int64_t col_d(const rf_gemm_desc& g, int n) {
    return g.nLo ? (int64_t)(n / g.nLo) * g.dSnHi + (int64_t)(n % g.nLo) * g.dSn : (int64_t)n * g.dSn;
}

__device__ inline float epi32(float u, const rf_epilogue& e) {
    if (e.margi == RF_FOM_NONE) return u;
    const float y = fmaf(u, (float)e.scale, (float)e.shift);
    return (e.margi == RF_FOM_LOGNORMAL) ? expf(y) : y;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// One 64 x 64 output tile (complex rows m0.., columns n0.., batch) of
// D = A B on the matrix cores in FP32 (v_mfma_f32_32x32x2_f32): the complex
// product is the real GEMM
//   [Re D; Im D] = [[Re A, -Im A], [Im A, Re A]] [Re B; Im B]
// with the real rows interleaved (2m: Re, 2m+1: Im), one complex k per
// instruction (lanes 0-31 take the k' = 2k column, 32-63 the k' = 2k+1
// column).  RealOnly forms Re D only: [Re A, -Im A] [Re B; Im B].  Operand
// tiles of 16 complex k are staged in LDS as FP32 complex, as loaded.
// Waves: complex output 128 real rows x 64 columns, wave w owns real rows
// 32w.. and both 32-column halves; RealOnly 64 x 64, wave w one 32 x 32 tile.
template <bool RealOnly>
__device__ __forceinline__ void gemm_tile(const rf_gemm_desc& g, int m0, int n0, int batch,
                                          float2 (*sA)[kTM + 1], float2 (*sB)[kTN + 1]) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const float2* A = g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB;
    const float2* B = g.B + (int64_t)batch * g.bStrideB;
    f32x16 acc0, acc1;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        acc0[i] = 0.0f;
        acc1[i] = 0.0f;
    }
    // RealOnly tile of this wave
    const int rt = wave >> 1, ct = wave & 1;
    // operand chunks of kTK complex k: each thread fetches kPer elements of
    // A and of B into registers one chunk ahead (zero outside the matrices)
    constexpr int kPer = kTM * kTK / 256;
    static_assert(kTM * kTK == kTK * kTN && kPer * 256 == kTM * kTK, "staging split");
    float2 ra[kPer], rb[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int e = tid + i * 256;
            const int mm = e / kTK, ka = e % kTK;
            const int m = m0 + mm, k = k0 + ka;
            ra[i] = (m < g.M && k < g.K) ? A[m * g.aSm + k * g.aSk] : make_float2(0.f, 0.f);
            const int kb = e / kTN, nn = e % kTN;
            const int n = n0 + nn, k2 = k0 + kb;
            rb[i] = (n < g.N && k2 < g.K) ? B[k2 * g.bSk + col_b(g, n)] : make_float2(0.f, 0.f);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int e = tid + i * 256;
            sA[e % kTK][e / kTK] = ra[i];
            sB[e / kTN][e % kTN] = rb[i];
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        // 32 x 32 sub-tiles entirely outside D are skipped (wave-uniform), so
        // edge tiles of the ~130-point grid axes cost only their valid part
        if (!RealOnly) {
            if (m0 + wave * 16 < g.M) {
                if (n0 + 32 < g.N) {
#pragma unroll 4
                    for (int kk = 0; kk < kTK; kk++) {
                        const float2 a = sA[kk][wave * 16 + (l32 >> 1)];
                        // row part (l32 & 1): 0 -> Re row (Re A, -Im A), 1 -> Im row (Im A, Re A)
                        const float av = (l32 & 1) ? (half ? a.x : a.y) : (half ? -a.y : a.x);
                        const float2 b0 = sB[kk][l32], b1 = sB[kk][32 + l32];
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, half ? b0.y : b0.x, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, half ? b1.y : b1.x, acc1, 0, 0, 0);
                    }
                } else {
#pragma unroll 4
                    for (int kk = 0; kk < kTK; kk++) {
                        const float2 a = sA[kk][wave * 16 + (l32 >> 1)];
                        const float av = (l32 & 1) ? (half ? a.x : a.y) : (half ? -a.y : a.x);
                        const float2 b0 = sB[kk][l32];
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, half ? b0.y : b0.x, acc0, 0, 0, 0);
                    }
                }
            }
        } else if (m0 + rt * 32 < g.M && n0 + ct * 32 < g.N) {
#pragma unroll 4
            for (int kk = 0; kk < kTK; kk++) {
                const float2 a = sA[kk][rt * 32 + l32];
                const float2 b = sB[kk][ct * 32 + l32];
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(half ? -a.y : a.x, half ? b.y : b.x, acc0, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    if (!RealOnly) {
#pragma unroll
        for (int i = 0; i < 16; i += 2) {
            const int r = 8 * (i >> 2) + 4 * half + (i & 3);       // even: the Re row
            const int m = m0 + wave * 16 + (r >> 1);
            if (m >= g.M) continue;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int n = n0 + c * 32 + l32;
                if (n >= g.N) continue;
                const int64_t o = (int64_t)batch * g.dStrideB + m * g.dSm + col_d(g, n);
                static_cast<float2*>(g.D)[o] = c ? make_float2(acc1[i] * g.outScale, acc1[i + 1] * g.outScale)
                                                 : make_float2(acc0[i] * g.outScale, acc0[i + 1] * g.outScale);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = m0 + rt * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
            const int n = n0 + ct * 32 + l32;
            if (m >= g.M || n >= g.N) continue;
            const int64_t o = (int64_t)batch * g.dStrideB + m * g.dSm + col_d(g, n);
            const float u = epi32(acc0[i] * g.outScale, g.epi);
            if (g.outDtype == RF_DTYPE_F64) static_cast<double*>(g.D)[o] = (double)u;
            else static_cast<float*>(g.D)[o] = u;
        }
    }
}

// The same tile on the FP16 matrix cores (v_mfma_f32_32x32x16_f16), operands
// split into FP16 hi + lo and D += A_hi B_hi + A_hi B_lo + A_lo B_hi (FP32
// accumulate): ~22 significant bits per operand, 3x the instructions of one
// product at 10x the FP32 MFMA rate.  A chunk of 16 complex k is converted
// once while staged: A into planes Re, Im, -Im (hi and lo each) as [row][k],
// B into Re, Im (hi, lo) as [column][k], so every MFMA operand is one 16-byte
// LDS read.  In the complex product a 32-row MFMA tile holds 16 complex rows,
// Re D in rows 0-15 ([Re A | -Im A] against [Re B ; Im B]) and Im D in rows
// 16-31 ([Im A | Re A]); K' = 16 covers 8 complex k, lanes 0-31 the Re half.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef const f2v __attribute__((address_space(1))) gf2;      // global-memory views
typedef f2v __attribute__((address_space(1))) gf2o;
typedef float __attribute__((address_space(1))) gfo;
typedef double __attribute__((address_space(1))) gdo;
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const f4v __attribute__((address_space(1))) gf4;
// n consecutive float2 (n even, 16-byte aligned) by 16-byte loads
template <int n>
__device__ __forceinline__ void ld_run(const gf2* p, f2v* v) {
    const gf4* q = reinterpret_cast<const gf4*>(p);
#pragma unroll
    for (int i = 0; i < n / 2; i++) {
        const f4v t = q[i];
        v[2 * i] = f2v{t.x, t.y};
        v[2 * i + 1] = f2v{t.z, t.w};
    }
}
typedef _Float16 hv8 __attribute__((ext_vector_type(8)));
typedef _Float16 hv4 __attribute__((ext_vector_type(4)));
constexpr int kHP = 24;                       // LDS row pitch (halves): 16 k + pad
// Fused blend: the group-grid column offset of plane column n, or -1 where
// another field also covers it; the group row offset of row m, or -1.
__device__ __forceinline__ int64_t grp_col(const rf_gemm_desc& g, int n) {
    const int x = n % g.grpNx, y = n / g.grpNx;
    return (x >= g.grpXlo && x < g.grpXhi && y >= g.grpYlo && y < g.grpYhi)
               ? g.grpCol0 + x + g.grpG0 * (int64_t)y : (int64_t)-1;
}
__device__ __forceinline__ int64_t grp_row(const rf_gemm_desc& g, int m) {
    return (m >= g.grpZlo && m < g.grpZhi) ? g.grpRow0 + g.grpPlaneG * (int64_t)m : (int64_t)-1;
}
// Fused blend: plane column n inside a first-axis overlap band sits in the
// band array of its side (*side 0 low, 1 high) at element offset (return) +
// m * (*rowStride) for row m; -1 outside the bands.
// (selects, not early returns: the blend's lookup of the same bands was
// miscompiled in an if-return form)
__device__ __forceinline__ int64_t band_col(const rf_gemm_desc& g, int n, int batch, int* side, int* rowStride) {
    const int x = n % g.grpNx, y = n / g.grpNx;
    const int wLo = g.grpXlo, wHi = g.grpNx - g.grpXhi;
    const bool lo = g.bandLo != nullptr && x < g.grpXlo;
    const bool hi = !lo && g.bandHi != nullptr && x >= g.grpXhi;
    const int64_t offLo = (int64_t)batch * g.bandLoStrideB + (int64_t)y * wLo + x;
    const int64_t offHi = (int64_t)batch * g.bandHiStrideB + (int64_t)y * wHi + (x - g.grpXhi);
    *side = hi ? 1 : 0;
    *rowStride = lo ? g.grpNy * wLo : hi ? g.grpNy * wHi : 0;
    return lo ? offLo : hi ? offHi : (int64_t)-1;
}

// Tile kinds (rf_gemm_desc::rw): 1 64x64; 2 128x64 complex product (two
// 16-row groups per wave share each B fragment read); 3 64x128 real part
// (two 32-column tiles per wave share each A fragment read; A holds only the
// Re and -Im planes); 4 64x128 complex product (four 32-column tiles per
// wave share each A fragment read); 5 64x160 complex product (a column axis
// of 129..160 in one tile: each row slab of A is staged once, not once per
// 64 columns with the last tile nearly empty; five 32-column tiles per wave
// share each A fragment read)
constexpr int kTN5 = 160;
template <int RW>
struct HStage {
    _Float16 a[RW == 3 ? 4 : 6][kTM * (RW == 2 ? 2 : 1)][kHP];   // Re hi, Re lo, Im hi, Im lo, -Im hi, -Im lo
    _Float16 b[4][RW == 5 ? kTN5 : kTN * (RW >= 3 ? 2 : 1)][kHP];  // Re hi, Re lo, Im hi, Im lo
};
__host__ __device__ inline int tile_m(const rf_gemm_desc& g) { return g.rw == 2 ? 2 * kTM : kTM; }
__host__ __device__ inline int tile_n(const rf_gemm_desc& g) {
    return g.rw == 5 ? kTN5 : (g.rw == 3 || g.rw == 4) ? 2 * kTN : kTN;
}

// v -> FP16 hi (round to nearest) and lo = v - hi: packed conversions for
// hi and lo, the residuals by mixed-precision FMAs reading the f16 halves in
// place (v_fma_mix_f32), two instructions per value.
__device__ inline void split4(const float* v, hv4* hi, hv4* lo) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int i = 0; i < 4; i += 2) {
        const h2 h = {(_Float16)v[i], (_Float16)v[i + 1]};       // v_cvt_pk_f16_f32
        const unsigned hb = __builtin_bit_cast(unsigned, h);
        float l0, l1;
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[0,0,0] op_sel_hi:[1,0,0]" : "=v"(l0) : "v"(hb), "v"(v[i]));
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(l1) : "v"(hb), "v"(v[i + 1]));
        (*hi)[i] = h.x;
        (*hi)[i + 1] = h.y;
        (*lo)[i] = (_Float16)l0;
        (*lo)[i + 1] = (_Float16)l1;
    }
}

// split4 for two values
typedef _Float16 hv2 __attribute__((ext_vector_type(2)));
__device__ inline void split2(const float* v, hv2* hi, hv2* lo) {
    const hv2 h = {(_Float16)v[0], (_Float16)v[1]};
    const unsigned hb = __builtin_bit_cast(unsigned, h);
    float l0, l1;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[0,0,0] op_sel_hi:[1,0,0]" : "=v"(l0) : "v"(hb), "v"(v[0]));
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(l1) : "v"(hb), "v"(v[1]));
    *hi = h;
    *lo = hv2{(_Float16)l0, (_Float16)l1};
}

// -x of FP16 values: the sign bits flipped
__device__ inline hv4 neg4(hv4 x) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(hv4, __builtin_bit_cast(u2, x) ^ 0x80008000u);
}

template <bool RealOnly>
__device__ __forceinline__ void gemm_tile_h(const rf_gemm_desc& g, int m0, int n0, int batch, HStage<1>* st) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    // tile base pointers in 64-bit, offsets inside the tile in 32-bit (the
    // launcher keeps K * stride and 64 * stride below 2^31)
    // global address space: the descriptor of a grouped launch is read from
    // memory, and generic (flat) loads would also count against the LDS
    // wait counter, so every LDS read of the K-step would wait for the
    // prefetch of the next one
    const gf2* A = (const gf2*)(g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB + (int64_t)m0 * g.aSm);
    const int64_t bCol0 = col_b(g, n0);
    const gf2* B = (const gf2*)(g.B + (int64_t)batch * g.bStrideB + bCol0);
    const int aSm = (int)g.aSm, aSk = (int)g.aSk, bSk = (int)g.bSk;
    f32x16 acc0, acc1;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        acc0[i] = 0.0f;
        acc1[i] = 0.0f;
    }
    const int rt = wave >> 1, ct = wave & 1;
    // staging: A row t/4, k 4(t%4)..+3; B column t%64, k 4(t/64)..+3
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 63, bK = (tid >> 6) * 4;
    const bool aIn = m0 + aRow < g.M, bIn = n0 + bCol < g.N;
    // rows / columns / k outside the operand stage zeros; the prefetch
    // loads are waited for only when staged (a clamped, branch-free form
    // with the zeroing at staging time measured 5 % slower: more VALU)
    const int aOff = aRow * aSm, bOff = bIn ? (int)(col_b(g, n0 + bCol) - bCol0) : 0;
    const int kLast = g.K - 1;
    f2v ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ka = k0 + aK + i, kb = k0 + bK + i;
            ra[i] = (aIn && ka <= kLast) ? A[aOff + ka * aSk] : f2v{0.f, 0.f};
            rb[i] = (bIn && kb <= kLast) ? B[kb * bSk + bOff] : f2v{0.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = ra[i].x;
            im[i] = ra[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[0][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[1][aRow][aK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[2][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[3][aRow][aK]) = lo;
        *reinterpret_cast<hv4*>(&st->a[4][aRow][aK]) = neg4(hi);   // split(-x) = -split(x)
        *reinterpret_cast<hv4*>(&st->a[5][aRow][aK]) = neg4(lo);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = rb[i].x;
            im[i] = rb[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->b[0][bCol][bK]) = hi;
        *reinterpret_cast<hv4*>(&st->b[1][bCol][bK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->b[2][bCol][bK]) = hi;
        *reinterpret_cast<hv4*>(&st->b[3][bCol][bK]) = lo;
    };
    // this lane's operand planes: complex product, tile row l32 = Re (< 16)
    // or Im row of complex row (l32 & 15); K' half 0 reads Re B, 1 reads Im B
    const int isIm = l32 >> 4;
    const int aPlane = RealOnly ? (half ? 4 : 0) : (isIm ? (half ? 0 : 2) : (half ? 4 : 0));
    const int aRowL = RealOnly ? rt * 32 + l32 : wave * 16 + (l32 & 15);
    const int bPlane = half ? 2 : 0;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 8) {
            if (!RealOnly) {
                if (m0 + wave * 16 < g.M) {
                    const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                    const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
                    const hv8 bh0 = *reinterpret_cast<const hv8*>(&st->b[bPlane][l32][kk]);
                    const hv8 bl0 = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][l32][kk]);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh0, acc0, 0, 0, 0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl0, acc0, 0, 0, 0);
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh0, acc0, 0, 0, 0);
                    if (n0 + 32 < g.N) {
                        const hv8 bh1 = *reinterpret_cast<const hv8*>(&st->b[bPlane][32 + l32][kk]);
                        const hv8 bl1 = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][32 + l32][kk]);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh1, acc1, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl1, acc1, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh1, acc1, 0, 0, 0);
                    }
                }
            } else if (m0 + rt * 32 < g.M && n0 + ct * 32 < g.N) {
                const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
                const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][ct * 32 + l32][kk]);
                const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][ct * 32 + l32][kk]);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc0, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc0, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc0, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dCol0 = col_d(g, n0);
    const int64_t dBase = (int64_t)batch * g.dStrideB + (int64_t)m0 * g.dSm + dCol0;
    const int dSm = (int)g.dSm;
    // this lane's column offsets (32-bit inside the tile)
    auto dcol = [&](int nn) { return n0 + nn < g.N ? (int)(col_d(g, n0 + nn) - dCol0) : 0; };
    if (!RealOnly) {
        gf2o* D = (gf2o*)(static_cast<float2*>(g.D) + dBase);
        const int oc[2] = {dcol(l32), dcol(32 + l32)};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = wave * 16 + 8 * (i >> 2) + 4 * half + (i & 3);   // Re row r, Im row r + 16
            if (m0 + r >= g.M) continue;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int nn = c * 32 + l32;
                if (n0 + nn >= g.N) continue;
                D[r * dSm + oc[c]] = c ? f2v{acc1[i] * os, acc1[i + 8] * os}
                                          : f2v{acc0[i] * os, acc0[i + 8] * os};
            }
        }
    } else {
        const int nn = ct * 32 + l32;
        const bool nIn = n0 + nn < g.N;
        const int oc = dcol(nn);
        // fused blend: this lane's column offset in the group grid, or -1,
        // and in a first-axis band array, or -1
        const int64_t gc = (g.grp && nIn) ? grp_col(g, n0 + nn) : -1;
        int bside = 0, bst = 0;
        const int64_t bo = (g.grp && nIn && gc < 0) ? band_col(g, n0 + nn, batch, &bside, &bst) : -1;
        auto store = [&](auto map, auto* D, auto* G, auto* BL, auto* BH) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = rt * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
                if (m0 + r < g.M && nIn) {
                    const auto v = map(acc0[i] * os);
                    const int64_t gr = gc >= 0 ? grp_row(g, m0 + r) : -1;
                    if (gr >= 0) G[gr + gc] = v;
                    else if (bo >= 0) (bside ? BH : BL)[bo + (int64_t)(m0 + r) * bst] = v;
                    else D[r * dSm + oc] = v;
                }
            }
        };
        const float sc = (float)g.epi.scale, sh = (float)g.epi.shift;
        gfo* D32 = (gfo*)(static_cast<float*>(g.D) + dBase);
        gdo* D64 = (gdo*)(static_cast<double*>(g.D) + dBase);
        gfo* G32 = (gfo*)(static_cast<float*>(g.grp) + (int64_t)batch * g.grpStrideB);
        gdo* G64 = (gdo*)(static_cast<double*>(g.grp) + (int64_t)batch * g.grpStrideB);
        gfo* L32 = (gfo*)static_cast<float*>(g.bandLo);
        gfo* H32 = (gfo*)static_cast<float*>(g.bandHi);
        gdo* L64 = (gdo*)static_cast<double*>(g.bandLo);
        gdo* H64 = (gdo*)static_cast<double*>(g.bandHi);
        if (g.outDtype == RF_DTYPE_F64) {
            if (g.epi.margi == RF_FOM_NONE) store([](float u) { return (double)u; }, D64, G64, L64, H64);
            else if (g.epi.margi == RF_FOM_LOGNORMAL)
                store([=](float u) { return (double)expf(fmaf(u, sc, sh)); }, D64, G64, L64, H64);
            else store([=](float u) { return (double)fmaf(u, sc, sh); }, D64, G64, L64, H64);
        } else {
            if (g.epi.margi == RF_FOM_NONE) store([](float u) { return u; }, D32, G32, L32, H32);
            else if (g.epi.margi == RF_FOM_LOGNORMAL)
                store([=](float u) { return expf(fmaf(u, sc, sh)); }, D32, G32, L32, H32);
            else store([=](float u) { return fmaf(u, sc, sh); }, D32, G32, L32, H32);
        }
    }
}

// 128-row tiles of the complex product (used where the row axis is long,
// the first contraction): each wave owns two 16-complex-row groups and reads
// each B fragment once for both (8 fragment reads per 12 MFMAs instead of 6
// per 6); staging, operand split and epilogue as gemm_tile_h.
__device__ __forceinline__ void gemm_tile_c128(const rf_gemm_desc& g, int m0, int n0, int batch, HStage<2>* st) {
    constexpr int RW = 2;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB + (int64_t)m0 * g.aSm);
    const int64_t bCol0 = col_b(g, n0);
    const gf2* B = (const gf2*)(g.B + (int64_t)batch * g.bStrideB + bCol0);
    const int aSm = (int)g.aSm, aSk = (int)g.aSk, bSk = (int)g.bSk;
    f32x16 acc[RW][2];
#pragma unroll
    for (int rr = 0; rr < RW; rr++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            acc[rr][0][i] = 0.0f;
            acc[rr][1][i] = 0.0f;
        }
    // staging: A rows t/4 and t/4 + 64, k 4(t%4)..+3; B column t%64, k 4(t/64)..+3
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 63, bK = (tid >> 6) * 4;
    bool aIn[RW];
#pragma unroll
    for (int rr = 0; rr < RW; rr++) aIn[rr] = m0 + aRow + 64 * rr < g.M;
    const bool bIn = n0 + bCol < g.N;
    const int aOff = aRow * aSm, bOff = bIn ? (int)(col_b(g, n0 + bCol) - bCol0) : 0;
    const int kLast = g.K - 1;
    f2v ra[RW][4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ka = k0 + aK + i, kb = k0 + bK + i;
#pragma unroll
            for (int rr = 0; rr < RW; rr++)
                ra[rr][i] = (aIn[rr] && ka <= kLast) ? A[aOff + 64 * rr * aSm + ka * aSk] : f2v{0.f, 0.f};
            rb[i] = (bIn && kb <= kLast) ? B[kb * bSk + bOff] : f2v{0.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int rr = 0; rr < RW; rr++) {
            const int row = aRow + 64 * rr;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i] = ra[rr][i].x;
                im[i] = ra[rr][i].y;
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->a[0][row][aK]) = hi;
            *reinterpret_cast<hv4*>(&st->a[1][row][aK]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->a[2][row][aK]) = hi;
            *reinterpret_cast<hv4*>(&st->a[3][row][aK]) = lo;
            *reinterpret_cast<hv4*>(&st->a[4][row][aK]) = neg4(hi);   // split(-x) = -split(x)
            *reinterpret_cast<hv4*>(&st->a[5][row][aK]) = neg4(lo);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = rb[i].x;
            im[i] = rb[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->b[0][bCol][bK]) = hi;
        *reinterpret_cast<hv4*>(&st->b[1][bCol][bK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->b[2][bCol][bK]) = hi;
        *reinterpret_cast<hv4*>(&st->b[3][bCol][bK]) = lo;
    };
    // operand planes as in gemm_tile_h's complex product; row group rr of
    // this wave starts at row (2 wave + rr) 16
    const int isIm = l32 >> 4;
    const int aPlane = isIm ? (half ? 0 : 2) : (half ? 4 : 0);
    const int aRowL = wave * 16 * RW + (l32 & 15);
    const int bPlane = half ? 2 : 0;
    const bool c1 = n0 + 32 < g.N;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + wave * 16 * RW < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (c == 1 && !c1) break;                                      // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][32 * c + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][32 * c + l32][kk]);
#pragma unroll
                    for (int rr = 0; rr < RW; rr++) {
                        if (rr > 0 && m0 + (wave * RW + rr) * 16 >= g.M) break;   // uniform
                        const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL + 16 * rr][kk]);
                        const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL + 16 * rr][kk]);
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[rr][c], 0, 0, 0);
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[rr][c], 0, 0, 0);
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[rr][c], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dCol0 = col_d(g, n0);
    const int64_t dBase = (int64_t)batch * g.dStrideB + (int64_t)m0 * g.dSm + dCol0;
    const int dSm = (int)g.dSm;
    auto dcol = [&](int nn) { return n0 + nn < g.N ? (int)(col_d(g, n0 + nn) - dCol0) : 0; };
    gf2o* D = (gf2o*)(static_cast<float2*>(g.D) + dBase);
    const int oc[2] = {dcol(l32), dcol(32 + l32)};
#pragma unroll
    for (int rr = 0; rr < RW; rr++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = (wave * RW + rr) * 16 + 8 * (i >> 2) + 4 * half + (i & 3);   // Re row r, Im row r + 16
            if (m0 + r >= g.M) continue;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int nn = c * 32 + l32;
                if (n0 + nn >= g.N) continue;
                D[r * dSm + oc[c]] = f2v{acc[rr][c][i] * os, acc[rr][c][i + 8] * os};
            }
        }
}

// cos and sin of 2 pi U for U in [0, 1): FP32 sincospi of the FP32 part Uh
// and the first-order term of the remainder U - Uh (|U - Uh| <= 2^-25), a few
// FP32 ulps from the FP64 values -- the coefficients are FP32.
__device__ __forceinline__ void cis_rev(double U, float* c, float* s) {
    const float uh = (float)U;
    const float ul = (float)(U - (double)uh);
    float sh, ch;
    sincospif(2.0f * uh, &sh, &ch);
    const float d = 6.28318530717958648f * ul;
    *s = fmaf(d, ch, sh);
    *c = fmaf(-d, sh, ch);
}

// The folded coefficient of row (r, j3, q, j2) and column (h, j1) with its
// phase drawn from element drawBase + src of the stream (seed, tag 2): the
// one formula of the coefficient fold and of the first contraction that forms
// its coefficients itself.
__device__ __forceinline__ float2 fold_coef_draw(const rf_fold_desc& f, unsigned r, unsigned j3, unsigned q,
                                                 unsigned j2, unsigned j1, unsigned h, uint64_t seed,
                                                 int64_t drawBase) {
#pragma clang fp contract(off)
    unsigned b = 0;
    if (f.fold) b = f.nDim == 3 ? (q == 0 ? (h ? 3u : 0u) : (h ? 1u : 2u)) : h;
    const int64_t jb = (int64_t)b * f.kNu + ((int64_t)j3 * f.nk2 + j2) * f.nk1 + j1;   // b kNu + j
    const int64_t src = (int64_t)r * f.kNr + jb;
    const double a = f.amp[(int64_t)r * f.ampStrideR + jb] * f.ampMult;
    float cs, sn;
    cis_rev(rf_uniform(seed, 2u, (uint64_t)(drawBase + src)), &cs, &sn);
    return make_float2((float)(a * (double)cs), (float)(h ? -(a * (double)sn) : a * (double)sn));
}

// 64x160 tiles of the complex product (the first contraction when its
// column axis, the first grid axis, has 129..160 points): wave w owns complex
// rows 16 w.. and the five 32-column tiles, reading each A fragment once for
// all five; B (the phase table, small) is staged for all 160 columns.
// Staging, operand split and epilogue as gemm_tile_c128.
__device__ __forceinline__ void gemm_tile_c160(const rf_gemm_desc& g, int m0, int n0, int batch, HStage<5>* st) {
    constexpr int NC = kTN5 / 32;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB + (int64_t)m0 * g.aSm);
    const int64_t bCol0 = col_b(g, n0);
    const gf2* B = (const gf2*)(g.B + (int64_t)batch * g.bStrideB + bCol0);
    const int aSm = (int)g.aSm, aSk = (int)g.aSk, bSk = (int)g.bSk;
    f32x16 acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[c][i] = 0.0f;
    // staging: A row t/4, k 4(t%4)..+3; B columns t%128 (k 8(t/128)..+7) and
    // 128 + t%32 (k 2(t/32)..+1)
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 127, bK = (tid >> 7) * 8;
    const int bCol2 = 128 + (tid & 31), bK2 = (tid >> 5) * 2;
    const bool aIn = m0 + aRow < g.M, bIn = n0 + bCol < g.N, bIn2 = n0 + bCol2 < g.N;
    const int aOff = aRow * aSm;
    const int bOff = bIn ? (int)(col_b(g, n0 + bCol) - bCol0) : 0;
    const int bOff2 = bIn2 ? (int)(col_b(g, n0 + bCol2) - bCol0) : 0;
    const int kLast = g.K - 1;
    // a batch's fresh realizations: the coefficients of the current draw are
    // formed here (fold_coef_draw, as the fold would store them) instead of
    // read; the row's digits (j2, q, j3, r) once per tile
    const bool gen = g.genState != nullptr && g.genState[2] != 0;
    int64_t drawBase = 0;
    unsigned gr = 0, gj3 = 0, gq = 0, gj2 = 0;
    if (gen) {
        drawBase = g.genState[1] * (int64_t)g.genNmc * g.gen.kNr;
        const unsigned Q = (g.gen.fold && g.gen.nDim == 3) ? 2u : 1u;
        unsigned rem = (unsigned)(m0 + aRow);
        gj2 = rem % (unsigned)g.gen.nk2; rem /= (unsigned)g.gen.nk2;
        gq = rem % Q; rem /= Q;
        gj3 = g.gen.nDim == 3 ? rem % (unsigned)g.gen.nk3 : 0u;
        gr = g.gen.nDim == 3 ? rem / (unsigned)g.gen.nk3 : rem;
    }
    f2v ra[4], rb[8], rb2[2];
    auto fetch = [&](int k0) {
        if (gen) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int ka = k0 + aK + i;
                if (aIn && ka <= kLast) {
                    const float2 c = fold_coef_draw(g.gen, gr, gj3, gq, gj2, (unsigned)ka % (unsigned)g.gen.nk1,
                                                    (unsigned)ka / (unsigned)g.gen.nk1, g.genSeed, drawBase);
                    ra[i] = f2v{c.x, c.y};
                } else {
                    ra[i] = f2v{0.f, 0.f};
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int ka = k0 + aK + i;
                ra[i] = (aIn && ka <= kLast) ? A[aOff + ka * aSk] : f2v{0.f, 0.f};
            }
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int kb = k0 + bK + i;
            rb[i] = (bIn && kb <= kLast) ? B[kb * bSk + bOff] : f2v{0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int kb = k0 + bK2 + i;
            rb2[i] = (bIn2 && kb <= kLast) ? B[kb * bSk + bOff2] : f2v{0.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = ra[i].x;
            im[i] = ra[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[0][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[1][aRow][aK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[2][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[3][aRow][aK]) = lo;
        *reinterpret_cast<hv4*>(&st->a[4][aRow][aK]) = neg4(hi);   // split(-x) = -split(x)
        *reinterpret_cast<hv4*>(&st->a[5][aRow][aK]) = neg4(lo);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i] = rb[4 * h + i].x;
                im[i] = rb[4 * h + i].y;
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[0][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[1][bCol][bK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[2][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[3][bCol][bK + 4 * h]) = lo;
        }
        float re2[2] = {rb2[0].x, rb2[1].x}, im2[2] = {rb2[0].y, rb2[1].y};
        hv2 hi2, lo2;
        split2(re2, &hi2, &lo2);
        *reinterpret_cast<hv2*>(&st->b[0][bCol2][bK2]) = hi2;
        *reinterpret_cast<hv2*>(&st->b[1][bCol2][bK2]) = lo2;
        split2(im2, &hi2, &lo2);
        *reinterpret_cast<hv2*>(&st->b[2][bCol2][bK2]) = hi2;
        *reinterpret_cast<hv2*>(&st->b[3][bCol2][bK2]) = lo2;
    };
    const int isIm = l32 >> 4;
    const int aPlane = isIm ? (half ? 0 : 2) : (half ? 4 : 0);
    const int aRowL = wave * 16 + (l32 & 15);
    const int bPlane = half ? 2 : 0;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + wave * 16 < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
                const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    if (n0 + 32 * c >= g.N) break;                                 // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][32 * c + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][32 * c + l32][kk]);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dCol0 = col_d(g, n0);
    const int64_t dBase = (int64_t)batch * g.dStrideB + (int64_t)m0 * g.dSm + dCol0;
    const int dSm = (int)g.dSm;
    gf2o* D = (gf2o*)(static_cast<float2*>(g.D) + dBase);
    int oc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) oc[c] = n0 + 32 * c + l32 < g.N ? (int)(col_d(g, n0 + 32 * c + l32) - dCol0) : 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int r = wave * 16 + 8 * (i >> 2) + 4 * half + (i & 3);   // Re row r, Im row r + 16
        if (m0 + r >= g.M) continue;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (n0 + 32 * c + l32 >= g.N) continue;
            D[r * dSm + oc[c]] = f2v{acc[c][i] * os, acc[c][i + 8] * os};
        }
    }
}

// 64x128 tiles of the real part (the last contraction, whose column axis is
// the long (x2, x1) plane): wave w owns rows 32 (w / 2).. and the two 32-column
// tiles from 64 (w % 2), reading each A fragment once for both (6 fragment
// reads per 6 MFMAs instead of 4 per 3); half as many blocks, each with twice
// the MFMAs per 16-k chunk.  Staging, operand split and epilogue as gemm_tile_h.
__device__ __forceinline__ void gemm_tile_r128(const rf_gemm_desc& g, int m0, int n0, int batch, HStage<3>* st) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB + (int64_t)m0 * g.aSm);
    const int64_t bCol0 = col_b(g, n0);
    const gf2* B = (const gf2*)(g.B + (int64_t)batch * g.bStrideB + bCol0);
    const int aSm = (int)g.aSm, aSk = (int)g.aSk, bSk = (int)g.bSk;
    f32x16 acc[2];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        acc[0][i] = 0.0f;
        acc[1][i] = 0.0f;
    }
    const int rt = wave >> 1, ct = wave & 1;
    // staging: A row t/4, k 4(t%4)..+3; B column t%128, k 8(t/128)..+7
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 127, bK = (tid >> 7) * 8;
    const bool aIn = m0 + aRow < g.M, bIn = n0 + bCol < g.N;
    const int aOff = aRow * aSm, bOff = bIn ? (int)(col_b(g, n0 + bCol) - bCol0) : 0;
    const int kLast = g.K - 1;
    f2v ra[4], rb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ka = k0 + aK + i;
            ra[i] = (aIn && ka <= kLast) ? A[aOff + ka * aSk] : f2v{0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int kb = k0 + bK + i;
            rb[i] = (bIn && kb <= kLast) ? B[kb * bSk + bOff] : f2v{0.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = ra[i].x;
            im[i] = ra[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[0][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[1][aRow][aK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[2][aRow][aK]) = neg4(hi);   // split(-x) = -split(x)
        *reinterpret_cast<hv4*>(&st->a[3][aRow][aK]) = neg4(lo);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i] = rb[4 * h + i].x;
                im[i] = rb[4 * h + i].y;
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[0][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[1][bCol][bK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[2][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[3][bCol][bK + 4 * h]) = lo;
        }
    };
    // K' half 0 pairs Re A with Re B, half 1 -Im A with Im B
    const int aPlane = half ? 2 : 0, bPlane = half ? 2 : 0;
    const int aRowL = rt * 32 + l32;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + rt * 32 < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
                const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int col = 64 * ct + 32 * c;
                    if (n0 + col >= g.N) break;                                    // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][col + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][col + l32][kk]);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dCol0 = col_d(g, n0);
    const int64_t dBase = (int64_t)batch * g.dStrideB + (int64_t)m0 * g.dSm + dCol0;
    const int dSm = (int)g.dSm;
    auto dcol = [&](int nn) { return n0 + nn < g.N ? (int)(col_d(g, n0 + nn) - dCol0) : 0; };
    const int oc[2] = {dcol(64 * ct + l32), dcol(64 * ct + 32 + l32)};
    const bool nIn[2] = {n0 + 64 * ct + l32 < g.N, n0 + 64 * ct + 32 + l32 < g.N};
    // fused blend: the lane's column offsets in the group grid, or -1, and
    // in a first-axis band array, or -1
    int64_t gc[2], bo[2];
    int bside[2], bst[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const int nn = n0 + 64 * ct + 32 * c + l32;
        gc[c] = (g.grp && nIn[c]) ? grp_col(g, nn) : -1;
        bside[c] = 0;
        bst[c] = 0;
        bo[c] = (g.grp && nIn[c] && gc[c] < 0) ? band_col(g, nn, batch, &bside[c], &bst[c]) : -1;
    }
    auto store = [&](auto map, auto* D, auto* G, auto* BL, auto* BH) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rt * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
            if (m0 + r >= g.M) continue;
            const int64_t gr = (gc[0] >= 0 || gc[1] >= 0) ? grp_row(g, m0 + r) : -1;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (!nIn[c]) continue;
                const auto v = map(acc[c][i] * os);
                if (gr >= 0 && gc[c] >= 0) G[gr + gc[c]] = v;
                else if (bo[c] >= 0) (bside[c] ? BH : BL)[bo[c] + (int64_t)(m0 + r) * bst[c]] = v;
                else D[r * dSm + oc[c]] = v;
            }
        }
    };
    const float sc = (float)g.epi.scale, sh = (float)g.epi.shift;
    gfo* D32 = (gfo*)(static_cast<float*>(g.D) + dBase);
    gdo* D64 = (gdo*)(static_cast<double*>(g.D) + dBase);
    gfo* G32 = (gfo*)(static_cast<float*>(g.grp) + (int64_t)batch * g.grpStrideB);
    gdo* G64 = (gdo*)(static_cast<double*>(g.grp) + (int64_t)batch * g.grpStrideB);
    gfo* L32 = (gfo*)static_cast<float*>(g.bandLo);
    gfo* H32 = (gfo*)static_cast<float*>(g.bandHi);
    gdo* L64 = (gdo*)static_cast<double*>(g.bandLo);
    gdo* H64 = (gdo*)static_cast<double*>(g.bandHi);
    if (g.outDtype == RF_DTYPE_F64) {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return (double)u; }, D64, G64, L64, H64);
        else if (g.epi.margi == RF_FOM_LOGNORMAL)
            store([=](float u) { return (double)expf(fmaf(u, sc, sh)); }, D64, G64, L64, H64);
        else store([=](float u) { return (double)fmaf(u, sc, sh); }, D64, G64, L64, H64);
    } else {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return u; }, D32, G32, L32, H32);
        else if (g.epi.margi == RF_FOM_LOGNORMAL)
            store([=](float u) { return expf(fmaf(u, sc, sh)); }, D32, G32, L32, H32);
        else store([=](float u) { return fmaf(u, sc, sh); }, D32, G32, L32, H32);
    }
}

// Isotropic sum on a structured grid, per-realization directions.  With
// x = o + i step per axis, e^{i(k.x + phi)} factorises over the axes, so per
// realization r
//   u_r[l][(i, j)] = Re sum_k Z_k[l] B_k[(i, j)],   B_k[(i, j)] = W_k[i] Y_k[j],
//   W_k[i] = a_k e^{i(k_x x_i + phi_k)},  Y_k[j] = e^{i k_y y_j},  Z_k[l] = e^{i k_z z_l}
// (2-D: u_r[j][i] = Re sum_k Y_k[j] W_k[i]): one real-part GEMM with M the
// last axis, N the plane and K the shells, whose B operand is formed while
// staged -- one complex product per element -- instead of read.  Tables are
// [r][index][k], k fastest and padded to whole kTK chunks by zero entries:
// a thread's k run of a chunk is one 16-byte-aligned piece of a row, read
// by 16-byte loads from one address, with no masks; rows and columns past a
// tile's edge read the last row / column (finite values in accumulators
// that are never stored).  A 64x128 tile of the real part as gemm_tile_r128.
struct rf_iso_gemm {
    const float2* A;              // last-axis table [r][k][m]
    const float2* W;              // first-axis table with the coefficients [r][k][i]
    const float2* Y;              // second-axis table [r][k][j] (3-D; NULL in 2-D)
    int M, N, K, nx, ny;
    int64_t aStrideB, wStrideB, yStrideB;
    void* D;                      // [r][m][n]
    int outDtype;
    rf_epilogue epi;
    float outScale;
};
__device__ __forceinline__ void gemm_tile_iso(const rf_iso_gemm& g, int m0, int n0, int batch, HStage<3>* st) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)batch * g.aStrideB + (int64_t)m0 * g.K);
    const gf2* Wt = (const gf2*)(g.W + (int64_t)batch * g.wStrideB);
    const gf2* Yt = (const gf2*)(g.Y ? g.Y + (int64_t)batch * g.yStrideB : nullptr);
    f32x16 acc[2];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        acc[0][i] = 0.0f;
        acc[1][i] = 0.0f;
    }
    const int rt = wave >> 1, ct = wave & 1;
    // staging: A row t/4, k 4(t%4)..+3; B column t%128, k 8(t/128)..+7
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 127, bK = (tid >> 7) * 8;
    const int aRowC = m0 + aRow < g.M ? aRow : g.M - 1 - m0;
    const int bColC = n0 + bCol < g.N ? n0 + bCol : g.N - 1;
    const int bI = bColC % g.nx, bJ = bColC / g.nx;
    const bool hasY = Yt != nullptr;
    const gf2* aP = A + (int64_t)aRowC * g.K + aK;
    const gf2* wP = Wt + (int64_t)bI * g.K + bK;
    const gf2* yP = hasY ? Yt + (int64_t)bJ * g.K + bK : nullptr;
    f2v ra[4], rw[8], ry[8];
    auto fetch = [&](int k0) {
        ld_run<4>(aP + k0, ra);
        ld_run<8>(wP + k0, rw);
        if (hasY) ld_run<8>(yP + k0, ry);
        else {
#pragma unroll
            for (int i = 0; i < 8; i++) ry[i] = f2v{1.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = ra[i].x;
            im[i] = ra[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[0][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[1][aRow][aK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[2][aRow][aK]) = neg4(hi);   // split(-x) = -split(x)
        *reinterpret_cast<hv4*>(&st->a[3][aRow][aK]) = neg4(lo);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const f2v w = rw[4 * h + i], y = ry[4 * h + i];
                re[i] = fmaf(w.x, y.x, -w.y * y.y);
                im[i] = fmaf(w.x, y.y, w.y * y.x);
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[0][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[1][bCol][bK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[2][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[3][bCol][bK + 4 * h]) = lo;
        }
    };
    // K' half 0 pairs Re A with Re B, half 1 -Im A with Im B
    const int aPlane = half ? 2 : 0, bPlane = half ? 2 : 0;
    const int aRowL = rt * 32 + l32;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + rt * 32 < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
                const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int col = 64 * ct + 32 * c;
                    if (n0 + col >= g.N) break;                                    // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][col + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][col + l32][kk]);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dBase = (int64_t)batch * g.M * g.N + (int64_t)m0 * g.N + n0;
    const int nc[2] = {64 * ct + l32, 64 * ct + 32 + l32};
    const bool nIn[2] = {n0 + nc[0] < g.N, n0 + nc[1] < g.N};
    auto store = [&](auto map, auto* D) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rt * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
            if (m0 + r >= g.M) continue;
#pragma unroll
            for (int c = 0; c < 2; c++)
                if (nIn[c]) D[(int64_t)r * g.N + nc[c]] = map(acc[c][i] * os);
        }
    };
    const float sc = (float)g.epi.scale, sh = (float)g.epi.shift;
    gfo* D32 = (gfo*)(static_cast<float*>(g.D) + dBase);
    gdo* D64 = (gdo*)(static_cast<double*>(g.D) + dBase);
    if (g.outDtype == RF_DTYPE_F64) {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return (double)u; }, D64);
        else if (g.epi.margi == RF_FOM_LOGNORMAL) store([=](float u) { return (double)expf(fmaf(u, sc, sh)); }, D64);
        else store([=](float u) { return (double)fmaf(u, sc, sh); }, D64);
    } else {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return u; }, D32);
        else if (g.epi.margi == RF_FOM_LOGNORMAL) store([=](float u) { return expf(fmaf(u, sc, sh)); }, D32);
        else store([=](float u) { return fmaf(u, sc, sh); }, D32);
    }
}

// The isotropic grid GEMM on 128x128 tiles (a last axis longer than 64):
// each B element formed once per realization instead of once per 64 rows;
// wave w owns rows 64 (w / 2).. and columns 64 (w % 2).., four 32x32 MFMA
// tiles sharing two A and two B fragment reads per K-step.
struct IsoStage128 {
    _Float16 a[4][128][kHP];      // Re hi, Re lo, -Im hi, -Im lo
    _Float16 b[4][128][kHP];      // Re hi, Re lo, Im hi, Im lo
};
__device__ __forceinline__ void gemm_tile_iso128(const rf_iso_gemm& g, int m0, int n0, int batch, IsoStage128* st) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)batch * g.aStrideB + (int64_t)m0 * g.K);
    const gf2* Wt = (const gf2*)(g.W + (int64_t)batch * g.wStrideB);
    const gf2* Yt = (const gf2*)(g.Y ? g.Y + (int64_t)batch * g.yStrideB : nullptr);
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q >> 1][q & 1][i] = 0.0f;
    const int rt = wave >> 1, ct = wave & 1;
    // staging: A row t/2, k 8(t%2)..+7; B column t%128, k 8(t/128)..+7
    const int aRow = tid >> 1, aK = (tid & 1) * 8;
    const int bCol = tid & 127, bK = (tid >> 7) * 8;
    const int aRowC = m0 + aRow < g.M ? aRow : g.M - 1 - m0;
    const int bColC = n0 + bCol < g.N ? n0 + bCol : g.N - 1;
    const int bI = bColC % g.nx, bJ = bColC / g.nx;
    const bool hasY = Yt != nullptr;
    const gf2* aP = A + (int64_t)aRowC * g.K + aK;
    const gf2* wP = Wt + (int64_t)bI * g.K + bK;
    const gf2* yP = hasY ? Yt + (int64_t)bJ * g.K + bK : nullptr;
    f2v ra[8], rw[8], ry[8];
    auto fetch = [&](int k0) {
        ld_run<8>(aP + k0, ra);
        ld_run<8>(wP + k0, rw);
        if (hasY) ld_run<8>(yP + k0, ry);
        else {
#pragma unroll
            for (int i = 0; i < 8; i++) ry[i] = f2v{1.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i] = ra[4 * h + i].x;
                im[i] = ra[4 * h + i].y;
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->a[0][aRow][aK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->a[1][aRow][aK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->a[2][aRow][aK + 4 * h]) = neg4(hi);   // split(-x) = -split(x)
            *reinterpret_cast<hv4*>(&st->a[3][aRow][aK + 4 * h]) = neg4(lo);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const f2v w = rw[4 * h + i], y = ry[4 * h + i];
                re[i] = fmaf(w.x, y.x, -w.y * y.y);
                im[i] = fmaf(w.x, y.y, w.y * y.x);
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[0][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[1][bCol][bK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[2][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[3][bCol][bK + 4 * h]) = lo;
        }
    };
    // K' half 0 pairs Re A with Re B, half 1 -Im A with Im B
    const int aPlane = half ? 2 : 0, bPlane = half ? 2 : 0;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + rt * 64 < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
                hv8 ah[2], al[2];
#pragma unroll
                for (int rr = 0; rr < 2; rr++) {
                    ah[rr] = *reinterpret_cast<const hv8*>(&st->a[aPlane][rt * 64 + 32 * rr + l32][kk]);
                    al[rr] = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][rt * 64 + 32 * rr + l32][kk]);
                }
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int col = 64 * ct + 32 * c;
                    if (n0 + col >= g.N) break;                                    // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][col + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][col + l32][kk]);
#pragma unroll
                    for (int rr = 0; rr < 2; rr++) {
                        if (rr > 0 && m0 + rt * 64 + 32 >= g.M) break;            // uniform
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[rr], bh, acc[rr][c], 0, 0, 0);
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[rr], bl, acc[rr][c], 0, 0, 0);
                        acc[rr][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[rr], bh, acc[rr][c], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dBase = (int64_t)batch * g.M * g.N + (int64_t)m0 * g.N + n0;
    const int nc[2] = {64 * ct + l32, 64 * ct + 32 + l32};
    const bool nIn[2] = {n0 + nc[0] < g.N, n0 + nc[1] < g.N};
    auto store = [&](auto map, auto* D) {
#pragma unroll
        for (int rr = 0; rr < 2; rr++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int r = rt * 64 + 32 * rr + 8 * (i >> 2) + 4 * half + (i & 3);
                if (m0 + r >= g.M) continue;
#pragma unroll
                for (int c = 0; c < 2; c++)
                    if (nIn[c]) D[(int64_t)r * g.N + nc[c]] = map(acc[rr][c][i] * os);
            }
    };
    const float sc = (float)g.epi.scale, sh = (float)g.epi.shift;
    gfo* D32 = (gfo*)(static_cast<float*>(g.D) + dBase);
    gdo* D64 = (gdo*)(static_cast<double*>(g.D) + dBase);
    if (g.outDtype == RF_DTYPE_F64) {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return (double)u; }, D64);
        else if (g.epi.margi == RF_FOM_LOGNORMAL) store([=](float u) { return (double)expf(fmaf(u, sc, sh)); }, D64);
        else store([=](float u) { return (double)fmaf(u, sc, sh); }, D64);
    } else {
        if (g.epi.margi == RF_FOM_NONE) store([](float u) { return u; }, D32);
        else if (g.epi.margi == RF_FOM_LOGNORMAL) store([=](float u) { return expf(fmaf(u, sc, sh)); }, D32);
        else store([=](float u) { return fmaf(u, sc, sh); }, D32);
    }
}

// 64x128 tiles of the complex product (the middle contraction, whose
// columns (j3, x1) are long): wave w owns complex rows 16 w.. and four
// 32-column tiles, reading each A fragment once for all four (10 fragment
// reads per 12 MFMAs instead of 6 per 6).  Staging, operand split and
// epilogue as gemm_tile_h.
__device__ __forceinline__ void gemm_tile_cw(const rf_gemm_desc& g, int m0, int n0, int batch, HStage<4>* st) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, l32 = lane & 31;
    const gf2* A = (const gf2*)(g.A + (int64_t)((batch / g.aDiv) % g.aMod) * g.aStrideB + (int64_t)m0 * g.aSm);
    const int64_t bCol0 = col_b(g, n0);
    const gf2* B = (const gf2*)(g.B + (int64_t)batch * g.bStrideB + bCol0);
    const int aSm = (int)g.aSm, aSk = (int)g.aSk, bSk = (int)g.bSk;
    f32x16 acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[c][i] = 0.0f;
    // staging: A row t/4, k 4(t%4)..+3; B column t%128, k 8(t/128)..+7
    const int aRow = tid >> 2, aK = (tid & 3) * 4;
    const int bCol = tid & 127, bK = (tid >> 7) * 8;
    const bool aIn = m0 + aRow < g.M, bIn = n0 + bCol < g.N;
    const int aOff = aRow * aSm, bOff = bIn ? (int)(col_b(g, n0 + bCol) - bCol0) : 0;
    const int kLast = g.K - 1;
    f2v ra[4], rb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ka = k0 + aK + i;
            ra[i] = (aIn && ka <= kLast) ? A[aOff + ka * aSk] : f2v{0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int kb = k0 + bK + i;
            rb[i] = (bIn && kb <= kLast) ? B[kb * bSk + bOff] : f2v{0.f, 0.f};
        }
    };
    auto stash = [&]() {
        float re[4], im[4];
        hv4 hi, lo;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] = ra[i].x;
            im[i] = ra[i].y;
        }
        split4(re, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[0][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[1][aRow][aK]) = lo;
        split4(im, &hi, &lo);
        *reinterpret_cast<hv4*>(&st->a[2][aRow][aK]) = hi;
        *reinterpret_cast<hv4*>(&st->a[3][aRow][aK]) = lo;
        *reinterpret_cast<hv4*>(&st->a[4][aRow][aK]) = neg4(hi);   // split(-x) = -split(x)
        *reinterpret_cast<hv4*>(&st->a[5][aRow][aK]) = neg4(lo);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i] = rb[4 * h + i].x;
                im[i] = rb[4 * h + i].y;
            }
            split4(re, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[0][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[1][bCol][bK + 4 * h]) = lo;
            split4(im, &hi, &lo);
            *reinterpret_cast<hv4*>(&st->b[2][bCol][bK + 4 * h]) = hi;
            *reinterpret_cast<hv4*>(&st->b[3][bCol][bK + 4 * h]) = lo;
        }
    };
    // operand planes as in gemm_tile_h's complex product
    const int isIm = l32 >> 4;
    const int aPlane = isIm ? (half ? 0 : 2) : (half ? 4 : 0);
    const int aRowL = wave * 16 + (l32 & 15);
    const int bPlane = half ? 2 : 0;
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stash();
        __syncthreads();
        if (k0 + kTK < g.K) fetch(k0 + kTK);      // in flight during the MFMAs
        if (m0 + wave * 16 < g.M) {
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 8) {
                const hv8 ah = *reinterpret_cast<const hv8*>(&st->a[aPlane][aRowL][kk]);
                const hv8 al = *reinterpret_cast<const hv8*>(&st->a[aPlane + 1][aRowL][kk]);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    if (n0 + 32 * c >= g.N) break;                                 // uniform
                    const hv8 bh = *reinterpret_cast<const hv8*>(&st->b[bPlane][32 * c + l32][kk]);
                    const hv8 bl = *reinterpret_cast<const hv8*>(&st->b[bPlane + 1][32 * c + l32][kk]);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of a 32x32 tile: lane column l32, rows 8*(i/4) + 4*half + i%4
    const float os = g.outScale;
    const int64_t dCol0 = col_d(g, n0);
    const int64_t dBase = (int64_t)batch * g.dStrideB + (int64_t)m0 * g.dSm + dCol0;
    const int dSm = (int)g.dSm;
    auto dcol = [&](int nn) { return n0 + nn < g.N ? (int)(col_d(g, n0 + nn) - dCol0) : 0; };
    gf2o* D = (gf2o*)(static_cast<float2*>(g.D) + dBase);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int nn = 32 * c + l32;
        if (n0 + nn >= g.N) continue;
        const int oc = dcol(nn);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = wave * 16 + 8 * (i >> 2) + 4 * half + (i & 3);   // Re row r, Im row r + 16
            if (m0 + r < g.M) D[r * dSm + oc] = f2v{acc[c][i] * os, acc[c][i + 8] * os};
        }
    }
}

struct FStage {
    float2 a[kTK][kTM + 1];
    float2 b[kTK][kTN + 1];
};
template <int RW>
constexpr size_t kStageBytes = sizeof(HStage<RW>) > sizeof(FStage) ? sizeof(HStage<RW>) : sizeof(FStage);

// Tile kind RW (rf_gemm_desc::rw) > 1 only with FP16 operands: the host sets
// rw = 1 wherever it falls back to the FP32-operand tile.
template <bool RealOnly, int RW = 1>
__device__ __forceinline__ void run_tile(const rf_gemm_desc& g, int m0, int n0, int batch, char* smem) {
    if constexpr (RW == 2) {
        static_assert(!RealOnly, "kind 2 is the complex product");
        gemm_tile_c128(g, m0, n0, batch, reinterpret_cast<HStage<2>*>(smem));
    } else if constexpr (RW == 3) {
        static_assert(RealOnly, "kind 3 is the real part");
        gemm_tile_r128(g, m0, n0, batch, reinterpret_cast<HStage<3>*>(smem));
    } else if constexpr (RW == 4) {
        static_assert(!RealOnly, "kind 4 is the complex product");
        gemm_tile_cw(g, m0, n0, batch, reinterpret_cast<HStage<4>*>(smem));
    } else if constexpr (RW == 5) {
        static_assert(!RealOnly, "kind 5 is the complex product");
        gemm_tile_c160(g, m0, n0, batch, reinterpret_cast<HStage<5>*>(smem));
    } else if (g.split16) {
        gemm_tile_h<RealOnly>(g, m0, n0, batch, reinterpret_cast<HStage<1>*>(smem));
    } else {
        FStage* f = reinterpret_cast<FStage*>(smem);
        gemm_tile<RealOnly>(g, m0, n0, batch, f->a, f->b);
    }
}

// Tile order and placement.  A GEMM's tiles are numbered so that the tiles
// sharing the larger operand are consecutive: row tiles fastest when there
// are no more of them than column tiles (they share a column slab of B, the
// previous stage's output), column tiles fastest otherwise (they share a row
// slab of A, the coefficients).  Workgroups go round-robin to the 8 XCDs, each
// with its own L2; block b runs logical tile (b % 8) * P + b / 8, so every XCD
// walks a contiguous range of P tiles and a shared slab is re-read from its
// L2, not from HBM.
constexpr int kXcd = 8;

__device__ inline void tile_coords(const rf_gemm_desc& g, int64_t local, int* mI, int* nI, int* batch) {
    const int tm = tile_m(g), tn = tile_n(g);
    const int ntN = (g.N + tn - 1) / tn, ntM = (g.M + tm - 1) / tm;
    if (ntM <= ntN) {
        *mI = (int)(local % ntM);
        const int64_t r = local / ntM;
        *nI = (int)(r % ntN);
        *batch = (int)(r / ntN);
    } else {
        *nI = (int)(local % ntN);
        const int64_t r = local / ntN;
        *mI = (int)(r % ntM);
        *batch = (int)(r / ntM);
    }
}

__device__ inline int64_t xcd_tile(int64_t b, int64_t nTiles) {
    const int64_t per = (nTiles + kXcd - 1) / kXcd;
    return (b % kXcd) * per + b / kXcd;
}

template <bool RealOnly, int RW>
__device__ __forceinline__ void cgemm_body(const rf_gemm_desc& g, int64_t nTiles, char* smem) {
    const int64_t L = xcd_tile(blockIdx.x, nTiles);
    if (L >= nTiles) return;
    int mI, nI, batch;
    tile_coords(g, L, &mI, &nI, &batch);
    run_tile<RealOnly, RW>(g, mI * tile_m(g), nI * tile_n(g), batch, smem);
}

template <bool RealOnly, int RW>
__global__ __launch_bounds__(256) void k_cgemm(rf_gemm_desc g, int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<RW>];
    cgemm_body<RealOnly, RW>(g, nTiles, smem);
}

// Grouped form: the GEMMs of several fields in one launch; GEMM q owns
// logical tiles tile0 .. tile0 + tiles(q) - 1.
// RWMAX > 1: the GEMMs of the group with rw = RWMAX use that tile kind
template <bool RealOnly, int RWMAX>
__device__ __forceinline__ void cgemm_group_body(const rf_gemm_desc* __restrict__ gs, int nG, int64_t nTiles,
                                                 char* smem) {
    const int64_t b = xcd_tile(blockIdx.x, nTiles);
    if (b >= nTiles) return;
    int lo = 0, hi = nG - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gs[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const rf_gemm_desc g = gs[lo];                 // uniform: scalar registers
    int mI, nI, batch;
    tile_coords(g, b - g.tile0, &mI, &nI, &batch);
    if (RWMAX > 1 && g.rw == RWMAX) run_tile<RealOnly, RWMAX>(g, mI * tile_m(g), nI * tile_n(g), batch, smem);
    else run_tile<RealOnly, 1>(g, mI * kTM, nI * kTN, batch, smem);
}

template <bool RealOnly, int RWMAX>
__global__ __launch_bounds__(256) void k_cgemm_group(const rf_gemm_desc* __restrict__ gs, int nG, int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<RWMAX>];
    cgemm_group_body<RealOnly, RWMAX>(gs, nG, nTiles, smem);
}

// Kind 5 holds 80 accumulators per lane: asked for three waves per SIMD (its
// LDS allows three blocks per CU) the compiler keeps them in the VGPRs
// (160 registers, no spill) instead of 94 VGPRs + 80 AGPRs, which would
// allow two.  The other kinds keep their own allocation.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_cgemm_c160(rf_gemm_desc g,
                                                                                         int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<5>];
    cgemm_body<false, 5>(g, nTiles, smem);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_cgemm_group_c160(
    const rf_gemm_desc* __restrict__ gs, int nG, int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<5>];
    cgemm_group_body<false, 5>(gs, nG, nTiles, smem);
}

// Kind 3 asked for four waves per SIMD (its LDS allows four blocks per CU):
// the 32 accumulators live in the VGPRs (114 registers, no spill) instead of
// 86 VGPRs + 32 AGPRs; the localization step's last contraction ran ~1 %
// faster so.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_cgemm_r128(rf_gemm_desc g,
                                                                                         int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<3>];
    cgemm_body<true, 3>(g, nTiles, smem);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_cgemm_group_r128(
    const rf_gemm_desc* __restrict__ gs, int nG, int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[kStageBytes<3>];
    cgemm_group_body<true, 3>(gs, nG, nTiles, smem);
}

// One 64x128 tile of one realization's isotropic grid GEMM per block; tiles
// numbered row tiles fastest (they share a column slab of B), placed per XCD.
template <int TM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_iso_gemm(rf_iso_gemm g, int64_t nTiles) {
    __shared__ __attribute__((aligned(16))) char smem[TM == 128 ? sizeof(IsoStage128) : sizeof(HStage<3>)];
    const int64_t L = xcd_tile(blockIdx.x, nTiles);
    if (L >= nTiles) return;
    const int ntM = (g.M + TM - 1) / TM, ntN = (g.N + 2 * kTN - 1) / (2 * kTN);
    const int mI = (int)(L % ntM);
    const int64_t rest = L / ntM;
    const int nI = (int)(rest % ntN);
    const int batch = (int)(rest / ntN);
    if constexpr (TM == 128)
        gemm_tile_iso128(g, mI * TM, nI * 2 * kTN, batch, reinterpret_cast<IsoStage128*>(smem));
    else
        gemm_tile_iso(g, mI * TM, nI * 2 * kTN, batch, reinterpret_cast<HStage<3>*>(smem));
}

// Isotropic phase tables [r][index][k] for the axes of the grid, in FP64
// then rounded: axis 0 carries the coefficient, W = c a e^{i(k_x x + phi)}
// (c: ampMult times a power of two), the others e^{i k_a x}; k runs to kP
// (kN rounded up to whole chunks), zero past kN.
__global__ void k_iso_tables(int nDim, int64_t kN, int64_t kP, int Nmc, int n0, int n1, int n2,
                             const double* __restrict__ kIso,
                             const double* __restrict__ amp, const double* __restrict__ phase, double coef,
                             double o0, double o1, double o2, double s0, double s1, double s2, double i0, double i1,
                             double i2, float2* __restrict__ T0, float2* __restrict__ T1, float2* __restrict__ T2) {
    const int nT = n0 + n1 + (nDim == 3 ? n2 : 0);
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)Nmc * kP * nT) return;
    const int64_t k = e % kP;                        // k fastest, as stored
    const int64_t rt = e / kP;                       // r * nT + t
    const int t = (int)(rt % nT);
    const int64_t r = rt / nT;
    const int64_t rk = r * kN + k;
    const int a = t < n0 ? 0 : (t < n0 + n1 ? 1 : 2);
    const int idx = a == 0 ? t : (a == 1 ? t - n0 : t - n0 - n1);
    const int na = a == 0 ? n0 : (a == 1 ? n1 : n2);
    float2* T = a == 0 ? T0 : (a == 1 ? T1 : T2);
    const int64_t dst = (r * na + idx) * kP + k;     // [r][idx][k]
    if (k >= kN) {
        T[dst] = make_float2(0.f, 0.f);
        return;
    }
    const double o = a == 0 ? o0 : (a == 1 ? o1 : o2);
    const double st = a == 0 ? s0 : (a == 1 ? s1 : s2);
    const double inv = a == 0 ? i0 : (a == 1 ? i1 : i2);
    const double x = (o + st * (double)idx) * inv;
    double th = kIso[rk * nDim + a] * x;
    if (a == 0) th += phase[rk];
    double sn, cs;
    sincos(th, &sn, &cs);
    const double m = a == 0 ? amp[rk] * coef : 1.0;
    T[dst] = make_float2((float)(m * cs), (float)(m * sn));
}

// E[v][j][x] = exp(i s_v k_j x), k_j = j dk, x = (o + i step) / L, in FP64;
// variant v = 0 plain, v = 1 conjugate (negative sign).
__global__ void k_phase_table(int nk, int nx, double dk, double o, double step, double invL,
                              int variants, float2* __restrict__ E) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)nk * nx) return;
    const int j = (int)(e / nx), i = (int)(e % nx);
    const double x = (o + step * (double)i) * invL;
    const double t = (dk * (double)j) * x;
    double s, c;
    sincos(t, &s, &c);
    E[e] = make_float2((float)c, (float)s);
    if (variants > 1) E[(int64_t)nk * nx + e] = make_float2((float)c, (float)-s);
}

// Method 4: E[v][j][i] = (w_j cos(pi j (2i+1) / (2n)), 0), w_0 = 1, w_j = 2:
// the REDFT01 (DCT-III) of FFTW that gen_Std_Gauss_FFT applies along each
// axis (randomFieldND.f90:694-735), as a matrix; both variants equal.
__global__ void k_dct_table(int n, float2* __restrict__ E) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int j = (int)(e / n), i = (int)(e % n);
    const double w = j == 0 ? 1.0 : 2.0;
    const double c = w * cos(3.14159265358979323846 * (double)j * (2.0 * i + 1.0) / (2.0 * n));
    E[e] = make_float2((float)c, 0.0f);
    E[(int64_t)n * n + e] = make_float2((float)c, 0.0f);
}

// Sign-block folding.  The field is Re of the sum, and the tables of a
// negative sign are conjugates, Re(E3^- E2^s2 T) = Re(E3^+ E2^-s2 conj(T)):
// each block with s3 = - folds onto s3 = + with its stage-1 result
// conjugated, and conj(sum_j1 E1 C) = sum_j1 conj(E1) conj(C).  In 3-D
// (blocks b = 2 f2 + f3, f = 1 for a negative sign):
//   V_q[j3][j2][x1] = sum_{h,j1} E1^{h}[j1][x1] Cv[j3][q][j2][h][j1],
//     q = 0: Cv[.][0] = C_0, Cv[.][1] = conj(C_3);  q = 1: C_2, conj(C_1)
//   W[j3][x2][x1] = sum_{q,j2} E2^{q}[j2][x2] V_q[j3][j2][x1]
//   u[x3][x2][x1] = Re sum_j3 E3^+[j3][x3] W[j3][x2][x1]
// In 2-D: V[j2][x1] = sum_{h,j1} E1^h (C_0 | conj(C_1)), u = Re sum_j2 E2^+ V.
// The middle intermediate W is a quarter of the per-block one and the last
// contraction runs over nk3 instead of 4 nk3.  Without sign blocks (method 4)
// Cv = C.
//
// Cv[r][j3][q][j2][h][j1] = (conj if h) ampMult amp[r][b][j][j1] exp(i phase),
// stored K-major: element (row m = (r, j3, q, j2), k = (h, j1)) at k * rows + m,
// so the first contraction's operand loads run along rows (whole cache lines
// per wave) instead of 16 rows of 32 bytes each.
// ph == NULL: real coefficients (method 4), no phase factor.
// One block per 64 x 64 tile of (row m, k): the coefficients are formed in
// the source order (k fastest: the amplitude and phase reads are contiguous
// along j1) into LDS and written out K-major (m fastest).
// The phase is 2 pi U with U the phase uniform [r][src] (randomFieldND.f90:
// 278-279), or, Draw, element (*draw * Nmc + r) * kNr + src of the device
// stream (seed, tag 2) -- the uniform the plan's generator would have drawn.
// ampStrideR 0: every realization has the amplitudes of row 0.
constexpr int kFoldM = 64, kFoldK = 32;   // coefficient-fold tile: rows x k


// One 64-row x 32-k tile of the folded coefficients, transposed through LDS
// to the K-major layout C[k * rows + m]; rows < 2^31 (checked by the launchers).
template <bool Draw>
__device__ __forceinline__ void fold_tile(const rf_fold_desc& f, int bx, int by, uint64_t seed, int64_t drawBase,
                                          float2 (*tile)[kFoldM + 1]) {
#pragma clang fp contract(off)
    const int rows = (int)f.rows;
    const int m0 = bx * kFoldM;
    const int k0 = by * kFoldK;
    const unsigned nk1 = (unsigned)f.nk1, nk2 = (unsigned)f.nk2, nk3 = (unsigned)f.nk3;
    const unsigned Q = (f.fold && f.nDim == 3) ? 2u : 1u;
    // a thread keeps one k (256 is a multiple of kFoldK) and walks rows
    // mm0, mm0 + 8, ...: the row's digits (j2, q, j3, r) advance by carries
    // instead of a division per element
    static_assert(256 % kFoldK == 0, "one k per thread");
    constexpr int kRowStep = 256 / kFoldK;
    const int kk = threadIdx.x % kFoldK, mm0 = threadIdx.x / kFoldK;
    const int k = k0 + kk;
    if (k < f.K) {
        const unsigned j1 = (unsigned)k % nk1;
        const unsigned h = (unsigned)k / nk1;
        unsigned rem = (unsigned)(m0 + mm0);
        unsigned j2 = rem % nk2; rem /= nk2;
        unsigned q = rem % Q; rem /= Q;
        unsigned j3 = f.nDim == 3 ? rem % nk3 : 0u;
        unsigned r = f.nDim == 3 ? rem / nk3 : rem;
        for (int mm = mm0; mm < kFoldM; mm += kRowStep) {
            const int m = m0 + mm;
            if (m >= rows) break;
            if (Draw) {
                tile[kk][mm] = fold_coef_draw(f, r, j3, q, j2, j1, h, seed, drawBase);
            } else {
                unsigned b = 0;
                if (f.fold) b = f.nDim == 3 ? (q == 0 ? (h ? 3u : 0u) : (h ? 1u : 2u)) : h;
                const int64_t jb = (int64_t)b * f.kNu + ((int64_t)j3 * nk2 + j2) * nk1 + j1;   // b kNu + j
                const int64_t src = (int64_t)r * f.kNr + jb;
                const double a = f.amp[(int64_t)r * f.ampStrideR + jb] * f.ampMult;
                if (!f.uni) {
                    tile[kk][mm] = make_float2((float)a, 0.0f);
                } else {
                    float cs, sn;
                    cis_rev(f.uni[src], &cs, &sn);
                    tile[kk][mm] = make_float2((float)(a * (double)cs), (float)(h ? -(a * (double)sn) : a * (double)sn));
                }
            }
            // next row m + kRowStep: j2 fastest, then q, then j3 (3-D), then r
            j2 += kRowStep;
            while (j2 >= nk2) {
                j2 -= nk2;
                if (++q == Q) {
                    q = 0;
                    if (f.nDim != 3) {
                        ++r;
                    } else if (++j3 == nk3) {
                        j3 = 0;
                        ++r;
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kFoldM * kFoldK; e += 256) {
        const int mm = e % kFoldM, kk = e / kFoldM;
        const int m = m0 + mm;
        const int k = k0 + kk;
        if (m < rows && k < f.K) f.C[(int64_t)k * rows + m] = tile[kk][mm];
    }
}

// The coefficients of one grid execution from the plan's phase uniforms
// (f.uni) or, method 4, the real DCT coefficients (f.uni NULL).
__global__ __launch_bounds__(256) void k_coeffs_fold(rf_fold_desc f) {
    __shared__ float2 tile[kFoldK][kFoldM + 1];
    fold_tile<false>(f, blockIdx.x, blockIdx.y, 0, 0, tile);
}

// Every field of a batch in one launch, phases drawn in place: field i owns
// blocks block0 .. block0 + nbx * nby - 1 (ascending in i), block -> field by
// binary search; realizations *draw * Nmc .. of the field's stream.  Logical
// blocks are placed per XCD (xcd_tile): each XCD folds a contiguous run of
// fields, so a field's amplitudes are fetched into one L2, not eight.
__global__ __launch_bounds__(256) void k_coeffs_fold_draw(const rf_fold_item* __restrict__ items, int n,
                                                          const int64_t* __restrict__ draw, int64_t nBlocks) {
    __shared__ float2 tile[kFoldK][kFoldM + 1];
    const int64_t blk = xcd_tile(blockIdx.x, nBlocks);
    if (blk >= nBlocks) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block0 <= blk) lo = mid; else hi = mid - 1;
    }
    const rf_fold_item& it = items[lo];
    const int lb = (int)(blk - it.block0);
    const int64_t drawBase = (*draw) * (int64_t)it.Nmc * it.f.kNr;
    fold_tile<true>(it.f, lb % it.nbx, lb / it.nbx, it.seed, drawBase, tile);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int64_t rf_gemm_tiles(const rf_gemm_desc& g) {
    const int tm = tile_m(g), tn = tile_n(g);
    return (int64_t)((g.N + tn - 1) / tn) * ((g.M + tm - 1) / tm) * g.batches;
}

// blocks of a launch over nTiles logical tiles (a multiple of the XCD count)
inline int64_t xcd_blocks(int64_t nTiles) { return (nTiles + kXcd - 1) / kXcd * kXcd; }

hipError_t rf_launch_gemm(const rf_gemm_desc& g, hipStream_t s) {
    const int64_t nTiles = rf_gemm_tiles(g);
    if (nTiles <= 0) return hipSuccess;
    const int64_t blocks = xcd_blocks(nTiles);
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (g.realOut && g.rw == 3) hipLaunchKernelGGL(k_cgemm_r128, dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else if (g.realOut) hipLaunchKernelGGL((k_cgemm<true, 1>), dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else if (g.rw == 2) hipLaunchKernelGGL((k_cgemm<false, 2>), dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else if (g.rw == 4) hipLaunchKernelGGL((k_cgemm<false, 4>), dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else if (g.rw == 5) hipLaunchKernelGGL(k_cgemm_c160, dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else hipLaunchKernelGGL((k_cgemm<false, 1>), dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    return hipGetLastError();
}

hipError_t rf_launch_gemm_group(int realOut, const rf_gemm_desc* d_gs, int nG, int64_t nTiles, int rwMax,
                                hipStream_t s) {
    if (nG <= 0 || nTiles <= 0) return hipSuccess;
    const int64_t blocks = xcd_blocks(nTiles);
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (realOut && rwMax == 3)
        hipLaunchKernelGGL(k_cgemm_group_r128, dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    else if (realOut)
        hipLaunchKernelGGL((k_cgemm_group<true, 1>), dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    else if (rwMax == 2)
        hipLaunchKernelGGL((k_cgemm_group<false, 2>), dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    else if (rwMax == 4)
        hipLaunchKernelGGL((k_cgemm_group<false, 4>), dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    else if (rwMax == 5)
        hipLaunchKernelGGL(k_cgemm_group_c160, dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    else
        hipLaunchKernelGGL((k_cgemm_group<false, 1>), dim3((unsigned)blocks), dim3(256), 0, s, d_gs, nG, nTiles);
    return hipGetLastError();
}

// Scratch for one grid execution (float2 elements): tables (both sign
// variants per axis), folded coefficients, V and (3-D) W.
size_t rf_grid_scratch_elems(int nDim, const int32_t* kNStep, int nb, int Nmc, const int32_t* n) {
    int64_t tables = 0;
    for (int a = 0; a < nDim; a++) tables += 2LL * kNStep[a] * n[a];
    int64_t coeff = (int64_t)nb * Nmc;
    for (int a = 0; a < nDim; a++) coeff *= kNStep[a];
    const int fold = nb > 1, Q = (fold && nDim == 3) ? 2 : 1;
    int64_t v = (int64_t)Nmc * Q * n[0];
    for (int a = 1; a < nDim; a++) v *= kNStep[a];
    const int64_t w = nDim == 3 ? (int64_t)Nmc * kNStep[2] * n[1] * n[0] : 0;
    return (size_t)(tables + coeff + v + w);
}

// Builds the phase tables and coefficients in scratch (launches on s) and
// describes the GEMM stages: stage 1 rows (r, b, j_d..j_2) x j1 -> x1;
// stage 2 (3-D) per (r, b, j3): x2 x j2 times j2 x x1; last stage per r:
// x_d x (b, j_d) times (b, j_d) x plane, real part, epilogue, into d_out.
hipError_t rf_grid_prepare(int kind, int nDim, int Nmc, const int32_t* kNStep, const double* kDelta, int nb,
                           const double* origin, const double* step, const int32_t* n,
                           const double* corrL, const double* d_amp, int ampShared, const double* d_unif,
                           double ampMult, double coefMax, const rf_epilogue& epi, void* d_out, int outDtype,
                           float2* scratch, hipStream_t s, rf_grid_stages* st, rf_grid_coef* coef) {
    // power-of-two scales for the FP16 operands: coefficients into [1/2, 1),
    // the middle contraction's output below 2^14 (|V| <= eMax K1 and
    // |W| <= eMax^2 K1 K2 for |table| <= eMax); the last stage undoes both
    const int fold = nb > 1 ? 1 : 0;
    const int H = fold ? 2 : 1, Q = (fold && nDim == 3) ? 2 : 1;
    const double eMax = kind == 1 ? 2.0 : 1.0;
    double sc = 1.0;
    if (coefMax > 0.0 && std::isfinite(coefMax)) {
        int e = 0;
        std::frexp(coefMax, &e);
        sc = std::ldexp(1.0, -e);
    }
    const int K1 = H * kNStep[0];
    const int K2 = nDim == 3 ? Q * kNStep[1] : kNStep[1];
    const double bound1 = eMax * K1;
    const double bound2 = nDim == 3 ? eMax * eMax * K1 * (double)K2 : bound1;
    int m2 = 0;
    while (nDim == 3 && std::ldexp(bound2, -m2) > 16384.0) m2++;
    int split = bound1 <= 16384.0 ? 1 : 0;

    float2* p = scratch;
    float2* E[3];
    for (int a = 0; a < nDim; a++) {
        E[a] = p; p += 2LL * kNStep[a] * n[a];
        const int64_t cnt = (int64_t)kNStep[a] * n[a];
        if (kind == 1)
            hipLaunchKernelGGL(k_dct_table, dim3(blocks_for(cnt)), dim3(256), 0, s, n[a], E[a]);
        else
            hipLaunchKernelGGL(k_phase_table, dim3(blocks_for(cnt)), dim3(256), 0, s, kNStep[a], n[a],
                               kDelta[a], origin[a], step[a], 1.0 / corrL[a], 2, E[a]);
    }
    int64_t kNu = 1;
    for (int a = 0; a < nDim; a++) kNu *= kNStep[a];
    const int64_t nC = (int64_t)Nmc * nb * kNu;          // = Nmc x nk3 x Q x nk2 x H x nk1
    float2* Cf = p; p += nC;
    // a caller-owned coefficient buffer is reused while its scale matches
    // (the coefficients depend only on the plan and the scale)
    bool fresh = false;
    if (coef) {
        Cf = coef->buf;
        fresh = coef->key == ampMult * sc;
        coef->key = ampMult * sc;
    }
    *st = rf_grid_stages{};
    rf_fold_desc& fd = st->fold;
    fd.rows = nC / K1; fd.K = K1; fd.nDim = nDim; fd.fold = fold;
    fd.nk1 = kNStep[0]; fd.nk2 = nDim > 1 ? kNStep[1] : 1; fd.nk3 = nDim > 2 ? kNStep[2] : 1;
    fd.kNu = kNu; fd.kNr = (int64_t)nb * kNu;
    fd.amp = d_amp; fd.ampStrideR = ampShared ? 0 : fd.kNr;
    fd.uni = kind == 1 ? nullptr : d_unif;
    fd.ampMult = ampMult * sc; fd.C = Cf;
    if (fd.rows > INT32_MAX) return hipErrorInvalidValue;
    if (!fresh)
        hipLaunchKernelGGL(k_coeffs_fold,
                           dim3((unsigned)((fd.rows + kFoldM - 1) / kFoldM), (unsigned)((K1 + kFoldK - 1) / kFoldK)),
                           dim3(256), 0, s, fd);
    // stage 1: rows (r, j3, q, j2) x K (h, j1) -> V[r][j3][q][j2][x1]
    const int64_t rows1 = nC / K1;
    float2* V = p; p += rows1 * n[0];
    rf_gemm_desc& g = st->g[st->n++];
    g.A = Cf; g.aSm = 1; g.aSk = rows1; g.aStrideB = 0; g.aDiv = 1; g.aMod = 1;   // K-major
    g.B = E[0]; g.bSk = n[0]; g.bSn = 1; g.bStrideB = 0;     // K (h, j1): variant h of E1
    g.D = V; g.dSm = n[0]; g.dSn = 1; g.dStrideB = 0;
    g.M = (int)rows1; g.N = n[0]; g.K = K1; g.batches = 1; g.realOut = 0;
    // 128-row tiles over the long row axis; a first axis of 129..160 points
    // in one 160-column tile
    g.rw = n[0] > 2 * kTN && n[0] <= kTN5 ? 5 : rows1 >= 128 ? 2 : 1;
    g.outScale = 1.0f; g.split16 = split;
    const float2* last = V;
    int64_t plane = n[0];                      // points already resolved
    int64_t lastStrideB = (int64_t)kNStep[1] * n[0];
    if (nDim == 3) {
        // stage 2, batch r, columns (j3, x1):
        // W[j3][x2][x1] = sum_{q,j2} E2^q[j2][x2] V[j3][q][j2][x1]
        float2* W = p; p += (int64_t)Nmc * kNStep[2] * n[1] * n[0];
        rf_gemm_desc& h = st->g[st->n++];
        h.A = E[1]; h.aSm = 1; h.aSk = n[1]; h.aStrideB = 0; h.aDiv = 1; h.aMod = 1;
        h.B = V; h.bSk = n[0]; h.bSn = 1; h.bStrideB = (int64_t)kNStep[2] * K2 * n[0];
        h.D = W; h.dSm = n[0]; h.dSn = 1; h.dStrideB = (int64_t)kNStep[2] * n[1] * n[0];
        h.nLo = n[0]; h.bSnHi = (int64_t)K2 * n[0]; h.dSnHi = (int64_t)n[1] * n[0];
        h.M = n[1]; h.N = kNStep[2] * n[0]; h.K = K2; h.batches = Nmc; h.realOut = 0;
        h.rw = h.N >= 128 ? 4 : 1;            // 128-column tiles over the long (j3, x1) axis
        h.outScale = (float)std::ldexp(1.0, -m2); h.split16 = split;
        last = W;
        plane = (int64_t)n[0] * n[1];
        lastStrideB = (int64_t)kNStep[2] * plane;
    }
    // last stage per realization: u[x_d][plane] = Re sum_j E_d^+[j][x_d] last[j][plane]
    const int ad = nDim - 1;
    rf_gemm_desc& f = st->g[st->n++];
    f.A = E[ad]; f.aSm = 1; f.aSk = n[ad]; f.aStrideB = 0; f.aDiv = 1; f.aMod = 1;
    f.B = last; f.bSk = plane; f.bSn = 1; f.bStrideB = lastStrideB;
    f.D = d_out; f.dSm = plane; f.dSn = 1; f.dStrideB = plane * n[ad];
    f.M = n[ad]; f.N = (int)plane; f.K = kNStep[ad]; f.batches = Nmc; f.realOut = 1;
    f.outDtype = outDtype; f.epi = epi;
    f.outScale = (float)(1.0 / (sc * std::ldexp(1.0, -m2))); f.split16 = split;
    f.rw = plane >= 128 ? 3 : 1;              // 128-column tiles over the long plane axis
    // the wide tiles (kinds 2..4) exist with FP16 operands only: a K beyond
    // the FP16 range of the first contraction runs every stage on the FP32 tile
    if (!split)
        for (int j = 0; j < st->n; j++) st->g[j].rw = 1;
    // the split-precision tile addresses inside a tile with 32-bit offsets
    for (int i = 0; i < st->n; i++) {
        rf_gemm_desc& q = st->g[i];
        const int64_t lim = (int64_t)1 << 30;
        const int64_t tm = tile_m(q), tn = tile_n(q);
        if ((int64_t)q.K * q.aSk + tm * q.aSm >= lim || (int64_t)q.K * q.bSk + tn * q.bSn + q.bSnHi >= lim ||
            tm * q.dSm + tn * q.dSn + q.dSnHi >= lim)
            for (int j = 0; j < st->n; j++) {
                st->g[j].split16 = 0;
                st->g[j].rw = 1;                  // the FP32 tile is 64 x 64
            }
    }
    return hipGetLastError();
}

int rf_fold_item_blocks(const rf_fold_desc& f) {
    return (int)(((f.rows + kFoldM - 1) / kFoldM) * ((f.K + kFoldK - 1) / kFoldK));
}

void rf_fold_item_init(rf_fold_item* it, const rf_fold_desc& f, int Nmc, uint64_t seed, int64_t block0) {
    it->f = f;
    it->seed = seed;
    it->Nmc = Nmc;
    it->nbx = (int)((f.rows + kFoldM - 1) / kFoldM);
    it->block0 = block0;
}

hipError_t rf_launch_refold_draw(const rf_fold_item* d_items, int n, int64_t blocks, const int64_t* d_draw,
                                 hipStream_t s) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    if (!d_items || !d_draw || xcd_blocks(blocks) > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_coeffs_fold_draw, dim3((unsigned)xcd_blocks(blocks)), dim3(256), 0, s, d_items, n, d_draw,
                       blocks);
    return hipGetLastError();
}

namespace {
__global__ void k_draw_counter(int64_t* st, int64_t value, int gen) {
    if (threadIdx.x != 0) return;
    if (value >= 0) {
        st[0] = value;
        st[1] = value;
    } else if (value == -1) {
        st[1] = st[0];
        st[0] = st[0] + 1;
    } else if (value == -2) {
        st[0] = 1;
        st[1] = 0;
    }
    st[2] = gen;
}
}  // namespace

hipError_t rf_launch_draw_counter(int64_t* d_state, int64_t value, int gen, hipStream_t s) {
    hipLaunchKernelGGL(k_draw_counter, dim3(1), dim3(64), 0, s, d_state, value, gen);
    return hipGetLastError();
}

hipError_t rf_grid_run(const rf_grid_stages& st, hipStream_t s) {
    for (int i = 0; i < st.n; i++) {
        const hipError_t e = rf_launch_gemm(st.g[i], s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static int64_t iso_k_padded(int64_t kN) { return (kN + kTK - 1) / kTK * kTK; }

size_t rf_iso_grid_scratch_elems(int nDim, int64_t kN, int Nmc, const int32_t* n) {
    int64_t t = 0;
    for (int a = 0; a < nDim; a++) t += n[a];
    return (size_t)(Nmc * iso_k_padded(kN) * t);
}

hipError_t rf_launch_iso_grid(const rf_iso_grid_args& a, hipStream_t s) {
    if (a.nDim < 2 || a.nDim > 3 || a.Nmc < 1 || a.kN < 1) return hipErrorInvalidValue;
    // coefficients into [1/2, 1) by a power of two, undone by the epilogue
    double sc = 1.0;
    if (a.coefMax > 0.0 && std::isfinite(a.coefMax)) {
        int e = 0;
        std::frexp(a.coefMax, &e);
        sc = std::ldexp(1.0, -e);
    }
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.nDim == 3 ? a.n[2] : 1;
    const int64_t kP = iso_k_padded(a.kN);
    float2* T0 = a.scratch;
    float2* T1 = T0 + (int64_t)a.Nmc * kP * n0;
    float2* T2 = T1 + (int64_t)a.Nmc * kP * n1;
    const int64_t cnt = (int64_t)a.Nmc * kP * (n0 + n1 + (a.nDim == 3 ? n2 : 0));
    hipLaunchKernelGGL(k_iso_tables, dim3(blocks_for(cnt)), dim3(256), 0, s, a.nDim, a.kN, kP, a.Nmc, n0, n1, n2, a.kIso,
                       a.amp, a.phase, a.ampMult * sc, a.origin[0], a.origin[1], a.origin[2], a.step[0], a.step[1],
                       a.step[2], a.invL[0], a.invL[1], a.invL[2], T0, T1, T2);
    rf_iso_gemm g{};
    g.K = (int)kP;
    g.W = T0;
    g.nx = n0;
    g.wStrideB = kP * n0;
    if (a.nDim == 3) {
        g.A = T2; g.M = n2; g.aStrideB = kP * n2;
        g.Y = T1; g.ny = n1; g.yStrideB = kP * n1;
        g.N = n0 * n1;
    } else {
        g.A = T1; g.M = n1; g.aStrideB = kP * n1;
        g.Y = nullptr; g.ny = 1; g.yStrideB = 0;
        g.N = n0;
    }
    g.D = a.d_out;
    g.outDtype = a.outDtype;
    g.epi = a.epi;
    g.outScale = (float)(1.0 / sc);
    // a last axis longer than 64 on 128-row tiles (each B element formed once
    // per 128 rows)
    const int TM = g.M > kTM ? 128 : kTM;
    const int64_t nTiles = (int64_t)((g.M + TM - 1) / TM) * ((g.N + 2 * kTN - 1) / (2 * kTN)) * a.Nmc;
    const int64_t blocks = xcd_blocks(nTiles);
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (TM == 128) hipLaunchKernelGGL(k_iso_gemm<128>, dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    else hipLaunchKernelGGL(k_iso_gemm<kTM>, dim3((unsigned)blocks), dim3(256), 0, s, g, nTiles);
    return hipGetLastError();
}

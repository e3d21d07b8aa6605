// rf_sum.hip -- direct spectral sums u_r(x) = ampMult * sum_j a_rj cos(k_j.x + phi_rj)
// (randomFieldND.f90:184-337 methods 2/3, :368-514 method 1).
//
// Two precision forms:
//  * FP64: one thread per (point, realization), sequential sum over j in the
//    reference's order with FMA contraction off -- the on-GPU oracle form.
//  * FP32: one thread per point holding R realization accumulators.  The
//    argument k.x is formed in FP64, reduced to revolutions in [-1/2, 1/2)
//    once per (point, k) and reused for all R realizations; each term is then
//    one FP32 add, one hardware cosine (v_cos_f32 takes revolutions) and one
//    FMA.  Wavevector, phase and amplitude operands are wave-uniform and come
//    through scalar loads, so the inner loop is pure VALU/transcendental work.
#include "rf_internal.h"

namespace {

constexpr int kBlock = 256;

template <typename T>
__device__ inline void store_out(void* out, int64_t idx, T v, int dtype) {
    if (dtype == RF_DTYPE_F64) static_cast<double*>(out)[idx] = (double)v;
    else static_cast<float*>(out)[idx] = (float)v;
}

// randomFieldND.f90:171-176 fused after the sum.
__device__ inline double epilogue64(double u, const rf_epilogue& e) {
#pragma clang fp contract(off)
    if (e.margi == RF_FOM_NONE) return u;
    const double y = u * e.scale + e.shift;
    return (e.margi == RF_FOM_LOGNORMAL) ? exp(y) : y;
}
__device__ inline float epilogue32(float u, const rf_epilogue& e) {
    if (e.margi == RF_FOM_NONE) return u;
    const float y = fmaf(u, (float)e.scale, (float)e.shift);
    return (e.margi == RF_FOM_LOGNORMAL) ? expf(y) : y;
}

// ---------------------------------------------------------------- FP64 form
template <int D>
__global__ __launch_bounds__(kBlock) void k_sum_fp64(rf_sum_args a, const double* __restrict__ k,
                                                     int64_t kStrideR,
                                                     const double* __restrict__ amp,
                                                     const double* __restrict__ phase) {
#pragma clang fp contract(off)
    const int64_t i = a.i0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= a.i1) return;
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = a.x[i * D + d] / a.corrL[d];   // :59
    const double* kr = k + (int64_t)r * kStrideR;
    const double* am = amp + (int64_t)r * a.kN;
    const double* ph = phase + (int64_t)r * a.kN;
    double acc = 0.0;
    for (int64_t j = 0; j < a.kN; j++) {
        // dgemm.f90:246-253 (k.x then + phi), :310 cos, dgemm.f90:226-234
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) t = t + xn[d] * kr[j * D + d];
        const double c = cos(t + ph[j]);
        const double aj = am[j];
        if (aj != 0.0) acc = acc + aj * c;
    }
    const double u = epilogue64(a.ampMult * acc, a.epi);               // :330
    store_out(a.out, (int64_t)r * a.N + i, u, a.outDtype);
}

// ---------------------------------------------------------------- FP32 form
template <int D, int R>
__global__ __launch_bounds__(kBlock) void k_sum_fp32(rf_sum_args a, const double* __restrict__ kRev,
                                                     const float* __restrict__ phT,
                                                     const float* __restrict__ aT, int Rpad) {
    const int64_t i = a.i0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r0 = blockIdx.y * R;
    if (i >= a.i1) return;
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = a.x[i * D + d] / a.corrL[d];
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    const float* ph = phT + r0;
    const float* am = aT + r0;
    for (int64_t j = 0; j < a.kN; j++) {
        const double* kj = kRev + j * D;
        double t = kj[0] * xn[0];
#pragma unroll
        for (int d = 1; d < D; d++) t = fma(kj[d], xn[d], t);
        t = t - rint(t);                 // k.x mod 2pi, in revolutions
        const float tf = (float)t;
        const float* phj = ph + j * Rpad;
        const float* amj = am + j * Rpad;
#pragma unroll
        for (int r = 0; r < R; r++)
            acc[r] = fmaf(amj[r], __builtin_amdgcn_cosf(tf + phj[r]), acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r0 + r < a.Nmc)
            store_out(a.out, (int64_t)(r0 + r) * a.N + i, epilogue32(acc[r], a.epi), a.outDtype);
    }
}

// FP32 form of the isotropic sum with per-realization directions: the
// argument cannot be shared across realizations, so each term forms its own
// FP64 k.x.  kT[j][r][D] in revolutions, aT/pT[j][r] as in k_sum_fp32.
// Padded realizations of a chunk (zero amplitude) are computed and not stored:
// a per-realization exit would serialize the chunk's scalar loads and
// transcendentals (measured slower on the cfg 5 stream).
// Global-address-space views of the direction / phase / amplitude streams:
// their reads are uniform across a wave, and only through such pointers (as
// kernel arguments are, but not pointers read from a request record) does
// the compiler issue them as batched scalar loads rather than one flat
// vector load and wait per realization.
typedef const double __attribute__((address_space(1))) gdbl;
typedef const float __attribute__((address_space(1))) gflt;

template <int D, int R>
__device__ __forceinline__ void sum_fp32_iso_point(const rf_sum_args& a, gdbl* __restrict__ kT,
                                                   gflt* __restrict__ pT,
                                                   gflt* __restrict__ aT, int Rpad, int64_t i,
                                                   int r0) {
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = a.x[i * D + d] / a.corrL[d];
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    for (int64_t j = 0; j < a.kN; j++) {
        gdbl* kj = kT + (j * Rpad + r0) * D;
        gflt* pj = pT + j * Rpad + r0;
        gflt* aj = aT + j * Rpad + r0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            double t = kj[r * D] * xn[0];
#pragma unroll
            for (int d = 1; d < D; d++) t = fma(kj[r * D + d], xn[d], t);
            t = __builtin_amdgcn_fract(t);   // k.x mod 1 revolution: one v_fract_f64, exact
            acc[r] = fmaf(aj[r], __builtin_amdgcn_cosf((float)t + pj[r]), acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r0 + r < a.Nmc)
            store_out(a.out, (int64_t)(r0 + r) * a.N + i, epilogue32(acc[r], a.epi), a.outDtype);
    }
}

template <int D, int R>
__global__ __launch_bounds__(kBlock) void k_sum_fp32_iso(rf_sum_args a, const double* __restrict__ kT,
                                                         const float* __restrict__ pT,
                                                         const float* __restrict__ aT, int Rpad) {
    const int64_t i = a.i0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.i1) return;
    sum_fp32_iso_point<D, R>(a, (gdbl*)kT, (gflt*)pT, (gflt*)aT, Rpad, i, blockIdx.y * R);
}

// Grouped form: block b runs one tile = (field, block of kGroupIsoPts points,
// chunk of kGroupIsoR realizations), found by binary search over the fields'
// first tiles; same per-point operation order as k_sum_fp32_iso.
__global__ __launch_bounds__(kBlock) void k_group_iso(const rf_group_req* __restrict__ reqs, int nReq) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = nReq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const rf_group_req& q = reqs[lo];
    const int nChunk = q.Rpad / kGroupIsoR;
    const int64_t local = b - q.tile0;
    const int64_t i = (local / nChunk) * kGroupIsoPts + threadIdx.x;
    if (i >= q.a.N) return;
    sum_fp32_iso_point<3, kGroupIsoR>(q.a, (gdbl*)q.ops0, (gflt*)q.ops1, (gflt*)q.ops2, q.Rpad, i,
                                      (int)(local % nChunk) * kGroupIsoR);
}

template <int D>
hipError_t launch_fp64(const rf_sum_args& a, const double* k, int64_t ks, const double* amp,
                       const double* ph, hipStream_t s) {
    const int64_t n = a.i1 - a.i0;
    dim3 grid((unsigned)((n + kBlock - 1) / kBlock), (unsigned)a.Nmc);
    hipLaunchKernelGGL(k_sum_fp64<D>, grid, dim3(kBlock), 0, s, a, k, ks, amp, ph);
    return hipGetLastError();
}

}  // namespace

// Realization chunk of the FP32 kernels (accumulators per thread).
constexpr int kRChunk = 16;

hipError_t rf_launch_sum_fp64(const rf_sum_args& a, const double* d_k, int64_t kStrideR,
                              const double* d_amp, const double* d_phase, hipStream_t s) {
    if (a.i1 <= a.i0) return hipSuccess;
    switch (a.nDim) {
        case 1: return launch_fp64<1>(a, d_k, kStrideR, d_amp, d_phase, s);
        case 2: return launch_fp64<2>(a, d_k, kStrideR, d_amp, d_phase, s);
        default: return launch_fp64<3>(a, d_k, kStrideR, d_amp, d_phase, s);
    }
}

hipError_t rf_launch_sum_fp32(const rf_sum_args& a, const double* d_kRev, const float* d_phT,
                              const float* d_aT, int Rpad, hipStream_t s) {
    if (a.i1 <= a.i0) return hipSuccess;
    const int64_t n = a.i1 - a.i0;
    dim3 grid((unsigned)((n + kBlock - 1) / kBlock), (unsigned)(Rpad / kRChunk));
    switch (a.nDim) {
        case 1:
            hipLaunchKernelGGL((k_sum_fp32<1, kRChunk>), grid, dim3(kBlock), 0, s, a, d_kRev, d_phT, d_aT, Rpad);
            break;
        case 2:
            hipLaunchKernelGGL((k_sum_fp32<2, kRChunk>), grid, dim3(kBlock), 0, s, a, d_kRev, d_phT, d_aT, Rpad);
            break;
        default:
            hipLaunchKernelGGL((k_sum_fp32<3, kRChunk>), grid, dim3(kBlock), 0, s, a, d_kRev, d_phT, d_aT, Rpad);
            break;
    }
    return hipGetLastError();
}

hipError_t rf_launch_sum_fp32_iso_t(const rf_sum_args& a, const double* d_kT, const float* d_pT,
                                    const float* d_aT, int Rpad, hipStream_t s) {
    if (a.i1 <= a.i0) return hipSuccess;
    const int64_t n = a.i1 - a.i0;
    dim3 grid((unsigned)((n + kBlock - 1) / kBlock), (unsigned)(Rpad / kRChunk));
    if (a.nDim == 2)
        hipLaunchKernelGGL((k_sum_fp32_iso<2, kRChunk>), grid, dim3(kBlock), 0, s, a, d_kT, d_pT, d_aT, Rpad);
    else
        hipLaunchKernelGGL((k_sum_fp32_iso<3, kRChunk>), grid, dim3(kBlock), 0, s, a, d_kT, d_pT, d_aT, Rpad);
    return hipGetLastError();
}

int rf_fp32_rchunk() { return kRChunk; }

hipError_t rf_launch_group_iso(const rf_group_req* d_reqs, int nReq, int64_t nTiles, hipStream_t s) {
    static_assert(kGroupIsoPts == kBlock, "one thread per point");
    if (nTiles <= 0 || nReq <= 0) return hipSuccess;
    if (nTiles > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_group_iso, dim3((unsigned)nTiles), dim3(kBlock), 0, s, d_reqs, nReq);
    return hipGetLastError();
}

// rf_prep.hip -- O(N_k * Nmc) coefficient preparation on the device: phases
// from uniforms, the isotropic per-realization direction tables, and the
// operand layouts of the FP32 direct-sum kernels.  All FP64.
#include "rf_internal.h"
#include "rf_philox.h"

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void k_uniforms(uint64_t seed, uint32_t tag, uint64_t first, int64_t count,
                           double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < count) out[e] = rf_uniform(seed, tag, first + (uint64_t)e);
}

// randomFieldND.f90:278-279: phiK = 2.0D0*pi*U, with the reference's pi.
__global__ void k_phase(const double* __restrict__ u, int64_t n, double* __restrict__ ph) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < n) ph[e] = 2.0 * RF_PI_REF * u[e];
}

// randomFieldND.f90:420-490 per (realization, wavevector): direction, amplitude
// and phase of the isotropic sum.  Uniform rows per realization (draw order):
// 3-D phi, theta, gamma, psi; 2-D psi, theta, gamma.
//   REF_COMPAT: azimuth theta = 2 pi psi (psi already scaled, :428/:466),
//               gamma = sqrt(12.0 single) (U - 1/2) in 3-D, 2 pi U in 2-D.
//   FIXED:      theta = 2 pi U_theta, gamma = sqrt(12) (U - 1/2).
//   shared:     directions of realization 0 for every realization.
// Outputs kIso[r][j][nDim] (normalised units), amp[r][j] and phase[r][j].
__global__ void k_iso_prep(int nDim, int mode, int shared, int64_t kN, int Nmc,
                           const double* __restrict__ kRad, const double* __restrict__ Sk,
                           const double* __restrict__ u, const uint8_t* __restrict__ calc,
                           double* __restrict__ kIso, double* __restrict__ amp,
                           double* __restrict__ phase) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= kN * Nmc) return;
    const int r = (int)(e / kN);
    const int64_t j = e - (int64_t)r * kN;
    const int nv = (nDim == 3) ? 4 : 3;
    const double* U = u + (int64_t)r * nv * kN;
    const double* Ud = shared ? u : U;      // direction draws
    const double k = kRad[j];
    double* kv = kIso + e * nDim;
    const bool on = calc[r] != 0;
    if (nDim == 3) {
        const double psi = 2.0 * RF_PI_REF * U[3 * kN + j];
        const double psid = 2.0 * RF_PI_REF * Ud[3 * kN + j];
        const double th = (mode == RF_REF_COMPAT) ? 2.0 * RF_PI_REF * psid
                                                  : 2.0 * RF_PI_REF * Ud[1 * kN + j];
        const double ph = RF_PI_REF * Ud[0 * kN + j];
        // REF_COMPAT: float(sqrt(12.0)) = 3.4641015529632568359375 exactly
        const double sq12 = (mode == RF_REF_COMPAT) ? 3.4641015529632568359375 : 3.4641016151377544;
        const double gam = sq12 * (U[2 * kN + j] - 0.5);
        const double sph = sin(ph);
        kv[0] = cos(th) * sph * k;
        kv[1] = sin(th) * sph * k;
        kv[2] = cos(ph) * k;
        amp[e] = on ? sqrt(Sk[j] * sph * (k * k)) * gam : 0.0;
        phase[e] = psi;
    } else {
        const double psi = 2.0 * RF_PI_REF * U[0 * kN + j];
        const double psid = 2.0 * RF_PI_REF * Ud[0 * kN + j];
        const double th = (mode == RF_REF_COMPAT) ? 2.0 * RF_PI_REF * psid
                                                  : 2.0 * RF_PI_REF * Ud[1 * kN + j];
        const double gam = (mode == RF_REF_COMPAT) ? 2.0 * RF_PI_REF * U[2 * kN + j]
                                                   : 3.4641016151377544 * (U[2 * kN + j] - 0.5);
        kv[0] = cos(th) * k;
        kv[1] = sin(th) * k;
        amp[e] = on ? sqrt(Sk[j] * k) * gam : 0.0;
        phase[e] = psi;
    }
}

// Reduce x (radians) to revolutions in [-1/2, 1/2).
__device__ inline double to_rev(double rad) {
    const double t = rad * RF_INV_TWO_PI;
    return t - rint(t);
}

// FP32 direct-sum operands, wavevector-major so a wave reads the Rpad values
// of one wavevector with scalar loads: phT[j][r] = phase in revolutions,
// aT[j][r] = amp * ampMult (zero for padded realizations).
__global__ void k_prep_fp32(int64_t kN, int Nmc, int Rpad, const double* __restrict__ phase,
                            const double* __restrict__ amp, double ampMult,
                            float* __restrict__ phT, float* __restrict__ aT) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= kN * Rpad) return;
    const int64_t j = e / Rpad;
    const int r = (int)(e - j * Rpad);
    if (r < Nmc) {
        phT[e] = (float)to_rev(phase[(int64_t)r * kN + j]);
        aT[e] = (float)(amp[(int64_t)r * kN + j] * ampMult);
    } else {
        phT[e] = 0.0f;
        aT[e] = 0.0f;
    }
}

// Isotropic FP32 operands: kT[j][r][D] = k_rj / (2 pi), pT[j][r] = psi_rj in
// revolutions, aT[j][r] = amp_rj * ampMult; padded realizations are zero.
__global__ void k_prep_fp32_iso(int nDim, int64_t kN, int Nmc, int Rpad,
                                const double* __restrict__ kIso, const double* __restrict__ phase,
                                const double* __restrict__ amp, double ampMult,
                                double* __restrict__ kT, float* __restrict__ pT,
                                float* __restrict__ aT) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= kN * Rpad) return;
    const int64_t j = e / Rpad;
    const int r = (int)(e - j * Rpad);
    if (r < Nmc) {
        const int64_t s = (int64_t)r * kN + j;
        for (int d = 0; d < nDim; d++) kT[e * nDim + d] = kIso[s * nDim + d] * RF_INV_TWO_PI;
        pT[e] = (float)to_rev(phase[s]);
        aT[e] = (float)(amp[s] * ampMult);
    } else {
        for (int d = 0; d < nDim; d++) kT[e * nDim + d] = 0.0;
        pT[e] = 0.0f;
        aT[e] = 0.0f;
    }
}

// cos(2 pi U) for U in [0, 1) in FP32: cospi of the FP32 part Uh with the
// first-order term of the remainder U - Uh (|U - Uh| <= 2^-25), a few FP32
// ulps from the FP64 value (the grid form's coefficients are FP32).  Exact pi:
// the reference's literal (constants_RF.f90:5) differs by 1e-14 relative.
__device__ __forceinline__ float cos_rev(double U) {
    const float uh = (float)U;
    const float ul = (float)(U - (double)uh);
    float sh, ch;
    sincospif(2.0f * uh, &sh, &ch);
    return fmaf(-6.28318530717958648f * ul, sh, ch);
}

// gen_Std_Gauss_FFT coefficients (randomFieldND.f90:683-689) per realization:
// uniforms u[r][0 .. 3 kN - 2] = gammaK (kN), kN - 1 discarded, phiK (kN);
// amp = (gammaK - 0.5) * sqrt(S) * cos(2 pi phiK), phase 0; the cosine in FP32
// (cos_rev): the grid form folds these amplitudes into FP32 coefficients.
__global__ void k_fft_coeffs(int64_t kN, int Nmc, const double* __restrict__ Sk,
                             const double* __restrict__ u, const uint8_t* __restrict__ calc,
                             double* __restrict__ amp, double* __restrict__ phase) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= kN * Nmc) return;
    const int r = (int)(e / kN);
    const int64_t j = e - (int64_t)r * kN;
    const double* U = u + (int64_t)r * (3 * kN - 1);
    const double g = U[j] - 0.5;
    const double ph = U[2 * kN - 1 + j];
    amp[e] = calc[r] ? g * sqrt(Sk[j]) * (double)cos_rev(ph) : 0.0;
    phase[e] = 0.0;
}

// The same coefficients with the uniforms drawn in place: element
// first + r (3 kN - 1) + i of stream (seed, tag 2) is U[r][i] (the values
// k_uniforms writes); the discarded third is never drawn.  Phase stays 0.
__global__ void k_fft_coeffs_draw(int64_t kN, int Nmc, const double* __restrict__ Sk,
                                  const uint8_t* __restrict__ calc, uint64_t seed, uint64_t first,
                                  double* __restrict__ amp) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= kN * Nmc) return;
    const int r = (int)(e / kN);
    const int64_t j = e - (int64_t)r * kN;
    if (!calc[r]) {
        amp[e] = 0.0;
        return;
    }
    const uint64_t row = first + (uint64_t)r * (uint64_t)(3 * kN - 1);
    const double g = rf_uniform(seed, 2u, row + (uint64_t)j) - 0.5;
    const double ph = rf_uniform(seed, 2u, row + (uint64_t)(2 * kN - 1 + j));
    amp[e] = g * sqrt(Sk[j]) * (double)cos_rev(ph);
}

__global__ void k_scale(int64_t n, const double* __restrict__ a, double s, double* __restrict__ b) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < n) b[e] = a[e] * s;
}

}  // namespace

hipError_t rf_launch_uniforms_at(uint64_t seed, uint32_t tag, uint64_t first, int64_t count,
                                 double* d_out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_uniforms, dim3(grid_for(count)), dim3(kBlock), 0, s, seed, tag, first,
                       count, d_out);
    return hipGetLastError();
}

hipError_t rf_launch_uniforms(uint64_t seed, uint32_t tag, int64_t count, double* d_out,
                              hipStream_t s) {
    return rf_launch_uniforms_at(seed, tag, 0, count, d_out, s);
}

hipError_t rf_launch_phase_from_uniform(const double* d_u, int64_t n, double* d_phase,
                                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_phase, dim3(grid_for(n)), dim3(kBlock), 0, s, d_u, n, d_phase);
    return hipGetLastError();
}

hipError_t rf_launch_iso_prep(int nDim, int mode, int shared, int64_t kN, int Nmc,
                              const double* d_kRad, const double* d_Sk, const double* d_u,
                              const uint8_t* d_calc, double* d_kIso, double* d_amp,
                              double* d_phase, hipStream_t s) {
    const int64_t n = kN * Nmc;
    hipLaunchKernelGGL(k_iso_prep, dim3(grid_for(n)), dim3(kBlock), 0, s, nDim, mode, shared, kN,
                       Nmc, d_kRad, d_Sk, d_u, d_calc, d_kIso, d_amp, d_phase);
    return hipGetLastError();
}

hipError_t rf_launch_prep_fp32(int64_t kN, int Nmc, int Rpad, const double* d_phase,
                               const double* d_amp, double ampMult, float* d_phT, float* d_aT,
                               hipStream_t s) {
    const int64_t n = kN * Rpad;
    hipLaunchKernelGGL(k_prep_fp32, dim3(grid_for(n)), dim3(kBlock), 0, s, kN, Nmc, Rpad, d_phase,
                       d_amp, ampMult, d_phT, d_aT);
    return hipGetLastError();
}

hipError_t rf_launch_krev(int64_t count, const double* d_k, double* d_kRev, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(count)), dim3(kBlock), 0, s, count, d_k,
                       (double)RF_INV_TWO_PI, d_kRev);
    return hipGetLastError();
}

hipError_t rf_launch_prep_fp32_iso(int nDim, int64_t kN, int Nmc, int Rpad, const double* d_kIso,
                                   const double* d_phase, const double* d_amp, double ampMult,
                                   double* d_kT, float* d_pT, float* d_aT, hipStream_t s) {
    const int64_t n = kN * Rpad;
    hipLaunchKernelGGL(k_prep_fp32_iso, dim3(grid_for(n)), dim3(kBlock), 0, s, nDim, kN, Nmc, Rpad,
                       d_kIso, d_phase, d_amp, ampMult, d_kT, d_pT, d_aT);
    return hipGetLastError();
}

hipError_t rf_launch_fft_coeffs_draw(int64_t kN, int Nmc, const double* d_Sk, const uint8_t* d_calc,
                                     uint64_t seed, uint64_t first, double* d_amp, hipStream_t s) {
    const int64_t n = kN * Nmc;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fft_coeffs_draw, dim3(grid_for(n)), dim3(kBlock), 0, s, kN, Nmc, d_Sk, d_calc, seed,
                       first, d_amp);
    return hipGetLastError();
}

hipError_t rf_launch_fft_coeffs(int64_t kN, int Nmc, const double* d_Sk, const double* d_u,
                                const uint8_t* d_calc, double* d_amp, double* d_phase, hipStream_t s) {
    const int64_t n = kN * Nmc;
    hipLaunchKernelGGL(k_fft_coeffs, dim3(grid_for(n)), dim3(kBlock), 0, s, kN, Nmc, d_Sk, d_u, d_calc, d_amp,
                       d_phase);
    return hipGetLastError();
}

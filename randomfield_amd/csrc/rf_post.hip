// rf_post.hip -- HBM-bound kernels around the spectral sum:
//   marginal map        multiVariateTransformation (randomFieldND.f90:139-178)
//   FP64 moments        calculate_average_and_stdVar_MPI (statistics_RF.f90:18-87)
//   normalisation       normalize_randField (randomFieldND.f90:820-926)
//   overlap blend       add_RF_to_Group + generateUnityPartition_Matrix + the
//                       boundary correction (localization_RF.f90:199-263,705-857;
//                       calls_RF.f90:324-369), in gather form: every group point
//                       reads its <= 2^d covering subdomains, no atomics.
#include "rf_internal.h"
#include "rf_blend.h"

namespace {

constexpr int kBlock = 256;

template <typename T>
__global__ void k_marginal(int margi, double scale, double shift, T* __restrict__ d, int64_t n) {
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        const double y = (double)d[e] * scale + shift;                 // :171
        d[e] = (T)((margi == RF_FOM_LOGNORMAL) ? exp(y) : y);         // :175
    }
}

// Per-realization partial sums of u and u^2 in FP64; grid (nPart, Nmc).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_moments_partial(const T* __restrict__ d, int64_t N,
                                                            double* __restrict__ part) {
    __shared__ double s1[kBlock], s2[kBlock];
    const int r = blockIdx.y;
    const T* col = d + (int64_t)r * N;
    double a = 0.0, b = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
        const double v = (double)col[i];
        a += v;
        b = fma(v, v, b);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((int64_t)r * gridDim.x + blockIdx.x) * 2 + 0] = s1[0];
        part[((int64_t)r * gridDim.x + blockIdx.x) * 2 + 1] = s2[0];
    }
}

// Fixed-order final reduction of the partials: sums[r] = (sum u, sum u^2).
__global__ void k_moments_final(const double* __restrict__ part, int nPart, int Nmc,
                                double* __restrict__ sums) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= Nmc) return;
    double a = 0.0, b = 0.0;
    for (int p = 0; p < nPart; p++) {
        a += part[((int64_t)r * nPart + p) * 2 + 0];
        b += part[((int64_t)r * nPart + p) * 2 + 1];
    }
    sums[2 * r + 0] = a;
    sums[2 * r + 1] = b;
}

// u <- (u - shift_r) * mul_r
template <typename T>
__global__ void k_affine(T* __restrict__ d, int64_t N, const double* __restrict__ shift,
                         const double* __restrict__ mul) {
    const int r = blockIdx.y;
    T* col = d + (int64_t)r * N;
    const double m = shift[r], s = mul[r];
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride)
        col[i] = (T)(((double)col[i] - m) * s);
}

// ----------------------------------------------------------------- blend
template <typename T>
__global__ __launch_bounds__(kBlock) void k_blend(rf_blend_dev b, T* __restrict__ group) {
#pragma clang fp contract(off)
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= b.groupTot) return;
    int gi[3] = {0, 0, 0};
    {
        int64_t rem = g;
        for (int a = 0; a < b.nDim; a++) { gi[a] = (int)(rem % b.groupN[a]); rem /= b.groupN[a]; }
    }
    int cv[3][2], nc[3] = {1, 1, 1};
    cv[1][0] = cv[2][0] = 0;
    for (int a = 0; a < b.nDim; a++) {
        nc[a] = 0;
        for (int c = 0; c < 2; c++) {
            const int v = b.cover[a][2 * gi[a] + c];
            if (v >= 0) cv[a][nc[a]++] = v;
        }
    }
    // <= 8 covering subdomains, kept in ascending subdomain index so the sum
    // runs in the reference's order (loop over fields, calls_RF.f90:237)
    int sub[8];
    int64_t loc[8];
    double w[8];
    int nh = 0;
    for (int a2 = 0; a2 < nc[2]; a2++)
        for (int a1 = 0; a1 < nc[1]; a1++)
            for (int a0 = 0; a0 < nc[0]; a0++) {
                const int iv[3] = {cv[0][a0], cv[1][a1], cv[2][a2]};
                const int s = b.subOf[iv[0] + b.nInt[0] * (iv[1] + b.nInt[1] * iv[2])];
                if (s < 0) continue;
                int64_t l = 0, stride = 1;
                int pos[3] = {0, 0, 0};
                for (int a = 0; a < b.nDim; a++) {
                    pos[a] = gi[a] - b.intOrigin[a][iv[a]];
                    l += (int64_t)pos[a] * stride;
                    stride *= b.intSize[a][iv[a]];
                }
                // PU product in the reference's order: lower faces of axes
                // 1..d, then upper faces (localization_RF.f90:755-851)
                double pu = 1.0;
                for (int a = 0; a < b.nDim; a++) pu = pu * b.lowF[((int64_t)s * 3 + a) * b.maxSize + pos[a]];
                for (int a = 0; a < b.nDim; a++) pu = pu * b.upF[((int64_t)s * 3 + a) * b.maxSize + pos[a]];
                int q = nh++;
                while (q > 0 && sub[q - 1] > s) { sub[q] = sub[q - 1]; loc[q] = loc[q - 1]; w[q] = w[q - 1]; q--; }
                sub[q] = s; loc[q] = l; w[q] = pu;
            }
    double pu1 = 1.0;
    if (b.mode == RF_REF_COMPAT) {
        for (int a = 0; a < b.nDim; a++) pu1 = pu1 * b.pu1Low[a][gi[a]];
        for (int a = 0; a < b.nDim; a++) pu1 = pu1 * b.pu1Up[a][gi[a]];
    }
    for (int r = 0; r < b.Nmc; r++) {
        double acc = 0.0;
        for (int h = 0; h < nh; h++) {
            const T* F = static_cast<const T*>(b.fields[sub[h]]);
            const double v = (double)F[(int64_t)r * b.subTot[sub[h]] + loc[h]];
            acc = v * w[h] + acc;                        // :228 then :253
        }
        if (b.mode == RF_REF_COMPAT && r == 0) acc = acc / pu1;   // calls_RF.f90:355
        group[(int64_t)r * b.groupTot + g] = (T)acc;
    }
}

}  // namespace

hipError_t rf_launch_marginal(int margi, double scale, double shift, void* d, int64_t count,
                              int dtype, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_marginal<double>, dim3((unsigned)blocks), dim3(kBlock), 0, s, margi,
                           scale, shift, static_cast<double*>(d), count);
    else
        hipLaunchKernelGGL(k_marginal<float>, dim3((unsigned)blocks), dim3(kBlock), 0, s, margi,
                           scale, shift, static_cast<float*>(d), count);
    return hipGetLastError();
}

int rf_moments_parts(int64_t N) {
    int64_t p = (N + kBlock * 8 - 1) / (kBlock * 8);
    if (p < 1) p = 1;
    if (p > 256) p = 256;
    return (int)p;
}

hipError_t rf_launch_moments(const void* d, int64_t N, int Nmc, int dtype, double* d_part,
                             int nPart, double* d_sums, hipStream_t s) {
    dim3 grid((unsigned)nPart, (unsigned)Nmc);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_moments_partial<double>, grid, dim3(kBlock), 0, s,
                           static_cast<const double*>(d), N, d_part);
    else
        hipLaunchKernelGGL(k_moments_partial<float>, grid, dim3(kBlock), 0, s,
                           static_cast<const float*>(d), N, d_part);
    hipLaunchKernelGGL(k_moments_final, dim3((unsigned)((Nmc + 63) / 64)), dim3(64), 0, s, d_part,
                       nPart, Nmc, d_sums);
    return hipGetLastError();
}

hipError_t rf_launch_affine(void* d, int64_t N, int Nmc, int dtype, const double* d_shift,
                            const double* d_mul, hipStream_t s) {
    int64_t bx = (N + kBlock - 1) / kBlock;
    if (bx > 4096) bx = 4096;
    dim3 grid((unsigned)bx, (unsigned)Nmc);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_affine<double>, grid, dim3(kBlock), 0, s, static_cast<double*>(d), N,
                           d_shift, d_mul);
    else
        hipLaunchKernelGGL(k_affine<float>, grid, dim3(kBlock), 0, s, static_cast<float*>(d), N,
                           d_shift, d_mul);
    return hipGetLastError();
}

hipError_t rf_launch_blend_dev(const rf_blend_dev& b, void* d_group, int dtype, hipStream_t s) {
    const unsigned grid = (unsigned)((b.groupTot + kBlock - 1) / kBlock);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_blend<double>, dim3(grid), dim3(kBlock), 0, s, b,
                           static_cast<double*>(d_group));
    else
        hipLaunchKernelGGL(k_blend<float>, dim3(grid), dim3(kBlock), 0, s, b,
                           static_cast<float*>(d_group));
    return hipGetLastError();
}

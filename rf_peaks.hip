// Throughput ceilings of the units the generation kernels run on, measured on
// the device (the datasheet figures assume a clock the chip does not hold under
// every load):
//   f16 MFMA  v_mfma_f32_32x32x16_f16, 8 accumulators per wave, 2 waves per
//             SIMD, operands constant / changed every iteration (xorshift: the
//             power the matrix cores draw depends on operand toggling)
//   f32 MFMA  v_mfma_f32_32x32x2f32 (the grid form), operands changing
//   sin       v_sin_f32 + one FMA per evaluation, 8 independent chains
// Prints one JSON object.  Every kernel writes one value per thread to a
// buffer of exactly grid x block floats.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16x __attribute__((ext_vector_type(16)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 512;

__device__ inline uint32_t xs(uint32_t y) {
    y ^= y << 13;
    y ^= y >> 17;
    y ^= y << 5;
    return y;
}

template <int TOGGLE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_mfma_f16(float* out, int iters) {
    u4 ua, ub;
    uint32_t x = 0x9E3779B9u * (threadIdx.x + 1) + blockIdx.x * 0x85EBCA6Bu;
    for (int i = 0; i < 4; i++) {
        x = xs(x); ua[i] = x;
        x = xs(x); ub[i] = x;
    }
    f16x acc[8];
    for (int m = 0; m < 8; m++)
        for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
    u4 va = ua & 0x3BFF3BFFu, vb = ub & 0x3BFF3BFFu;
    for (int it = 0; it < iters; it++) {
        if (TOGGLE) {
            for (int i = 0; i < 4; i++) {
                ua[i] = xs(ua[i]);
                ub[i] = xs(ub[i]);
                va[i] = ua[i] & 0xBBFFBBFFu;
                vb[i] = ub[i] & 0xBBFFBBFFu;
            }
        }
        const h8 a = __builtin_bit_cast(h8, va), b = __builtin_bit_cast(h8, vb);
#pragma unroll
        for (int rep = 0; rep < 3; rep++)
#pragma unroll
            for (int m = 0; m < 8; m++) acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[m], 0, 0, 0);
    }
    float s = 0.0f;
    for (int m = 0; m < 8; m++)
        for (int r = 0; r < 16; r++) s += acc[m][r];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_mfma_f32(float* out, int iters) {
    uint32_t ua = 0x9E3779B9u * (threadIdx.x + 1) + blockIdx.x * 0x85EBCA6Bu, ub = xs(ua);
    f16x acc[4];
    for (int m = 0; m < 4; m++)
        for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
    for (int it = 0; it < iters; it++) {
        ua = xs(ua);
        ub = xs(ub);
        const float a = __builtin_bit_cast(float, (ua & 0x807FFFFFu) | 0x3F000000u);
        const float b = __builtin_bit_cast(float, (ub & 0x807FFFFFu) | 0x3F000000u);
#pragma unroll
        for (int rep = 0; rep < 4; rep++)
#pragma unroll
            for (int m = 0; m < 4; m++) acc[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[m], 0, 0, 0);
    }
    float s = 0.0f;
    for (int m = 0; m < 4; m++)
        for (int r = 0; r < 16; r++) s += acc[m][r];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sin(float* out, int iters) {
    float v[8];
    for (int i = 0; i < 8; i++) v[i] = 1e-3f * (threadIdx.x + i);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = __builtin_amdgcn_sinf(fmaf(v[i], 0.75f, 0.1f));
    }
    float s = 0.0f;
    for (int i = 0; i < 8; i++) s += v[i];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

template <typename F>
float time_ms(F launch) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0f;
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        (void)hipEventRecord(e0, 0);
        launch();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;   // the first run warms up
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

int main() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    const int blocks = cus * 8;
    float* d = nullptr;
    if (hipMalloc(&d, (size_t)blocks * kBlock * sizeof(float)) != hipSuccess) return 1;
    const int itF16 = 4000, itF32 = 4000, itSin = 20000;
    const double wavesTotal = (double)blocks * (kBlock / 64);
    const float mc = time_ms([&] { hipLaunchKernelGGL(k_mfma_f16<0>, dim3(blocks), dim3(kBlock), 0, 0, d, itF16); });
    const float mt = time_ms([&] { hipLaunchKernelGGL(k_mfma_f16<1>, dim3(blocks), dim3(kBlock), 0, 0, d, itF16); });
    const float m32 = time_ms([&] { hipLaunchKernelGGL(k_mfma_f32, dim3(blocks), dim3(kBlock), 0, 0, d, itF32); });
    const float ms = time_ms([&] { hipLaunchKernelGGL(k_sin, dim3(blocks), dim3(kBlock), 0, 0, d, itSin); });
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        std::printf("{\"error\": \"kernel failure\"}\n");
        return 2;
    }
    const double f16flop = wavesTotal * itF16 * 24 * 32768.0;       // 32x32x16 x 2 FLOP
    const double f32flop = wavesTotal * itF32 * 16 * 4096.0;        // 32x32x2 x 2 FLOP
    const double sins = (double)blocks * kBlock * itSin * 8.0;
    std::printf("{\"cus\": %d, \"mfma_f16_constant_TFLOPs\": %.1f, \"mfma_f16_toggling_TFLOPs\": %.1f, "
                "\"mfma_f32_toggling_TFLOPs\": %.1f, \"sin_fma_per_s\": %.4e, "
                "\"ms\": [%.3f, %.3f, %.3f, %.3f]}\n",
                cus, f16flop / (mc * 1e-3) / 1e12, f16flop / (mt * 1e-3) / 1e12, f32flop / (m32 * 1e-3) / 1e12,
                sins / (ms * 1e-3), mc, mt, m32, ms);
    (void)hipFree(d);
    return 0;
}

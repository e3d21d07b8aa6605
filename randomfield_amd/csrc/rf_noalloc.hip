// rf_noalloc.hip -- rf_normalize_device_noalloc: the normalisation of
// rf_normalize_device (normalize_randField per realization: subtract the mean,
// divide by the root mean square, FP64 sums) for callers that cannot allocate
// device memory (code under graph capture, tight serving loops).  Its scratch
// -- the reduction partials, the sums and the per-realization coefficients --
// lives in one file-scope device array instead of the context's scratch, and
// the coefficients are formed on the device, so a call allocates nothing and
// never waits for the host.  The moment and affine kernels are the ones
// rf_normalize_device launches, with the same partition of the sums.
//
// The array holds RF_NOALLOC_MAX_NMC (1024) realizations.  Every call on a device uses
// that device's one copy and leaves its contents behind; calls that overlap
// on one device (two streams, two host threads) race on it.
#include <mutex>

#include "rf_blend.h"
#include "rf_internal.h"

namespace {

constexpr int kMaxNmc = RF_NOALLOC_MAX_NMC;
constexpr int kMaxParts = 256;            // rf_moments_parts() never returns more
// partials [Nmc][parts][2], then sums [Nmc][2], shift [Nmc], mul [Nmc]
constexpr size_t kDoubles = (size_t)kMaxNmc * kMaxParts * 2 + (size_t)kMaxNmc * 4;
constexpr int kMaxDevices = 64;

std::mutex g_setup;
double* g_scratch[kMaxDevices] = {};      // device address of `scratch` per device

}  // namespace

static __device__ __align__(256) double scratch[kDoubles];

namespace {

// pass 0: the mean of each realization (randomFieldND.f90:869-875);
// pass 1: the inverse root mean square (:904-908)
__global__ void k_norm_coeffs(const double* __restrict__ sums, int Nmc, double count, int pass,
                              double* __restrict__ shift, double* __restrict__ mul) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= Nmc) return;
    if (pass == 0) {
        shift[r] = sums[2 * r] / count;
        mul[r] = 1.0;
    } else {
        shift[r] = 0.0;
        mul[r] = 1.0 / sqrt(sums[2 * r + 1] / count);
    }
}

}  // namespace

// Once per device, with that device current (rf_context_create): the device
// address of this device's copy of `scratch`.
hipError_t rf_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup);
    if (g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(p);
    return hipSuccess;
}

extern "C" int rf_normalize_device_noalloc(rf_context* ctx, void* d_field, int64_t N, int32_t Nmc,
                                           int32_t dtype, void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (N < 1 || Nmc < 1 || !d_field) return rf_set_error(ctx, RF_E_ARG, "bad field");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    if (Nmc > kMaxNmc)
        return rf_set_error(ctx, RF_E_ARG,
                            "rf_normalize_device_noalloc: Nmc " + std::to_string(Nmc) + " is above " +
                                std::to_string(kMaxNmc) + ", the realizations its static scratch holds");
    double* base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
    if (!base) return rf_set_error(ctx, RF_E_HIP, "rf_normalize_device_noalloc: device scratch not set up");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const int nPart = rf_moments_parts(N);
    if (nPart > kMaxParts) return rf_set_error(ctx, RF_E_ARG, "rf_normalize_device_noalloc: too many partials");
    double* part = base;
    double* sums = part + (size_t)Nmc * nPart * 2;
    double* shift = sums + (size_t)Nmc * 2;
    double* mul = shift + Nmc;
    const dim3 cgrid((unsigned)((Nmc + 63) / 64)), cblock(64);
    for (int pass = 0; pass < 2; pass++) {
        RF_HIP_TRY(ctx, rf_launch_moments(d_field, N, Nmc, dtype, part, nPart, sums, s));
        hipLaunchKernelGGL(k_norm_coeffs, cgrid, cblock, 0, s, sums, Nmc, (double)N, pass, shift, mul);
        RF_HIP_TRY(ctx, hipGetLastError());
        RF_HIP_TRY(ctx, rf_launch_affine(d_field, N, Nmc, dtype, shift, mul, s));
    }
    return RF_OK;
}

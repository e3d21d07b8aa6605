// rf_capi.cpp -- the C ABI of librf_hip.so (declared in include/rf_capi.h).
//
// A plan holds everything of one createRandomField call that does not depend
// on the points: the wavevector table, phases and amplitudes, resident in HBM
// in the layout each precision form reads.  rf_plan_execute is the hot path
// (randomFieldND.f90:184-514): one launch over the points, or one per point
// slab in REF_COMPAT, where the reference re-applies sqrt(Sk) per slab.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "rf_blend.h"
#include "rf_internal.h"
#include "rf_philox.h"

namespace {

thread_local std::string g_last_error;

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int plan_fail(rf_context* ctx, int code, const std::string& msg) { return rf_set_error(ctx, code, msg); }

bool is_iso_direct(const rf_params& p) { return p.method == RF_ISOTROPIC && !p.isoShared; }

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Form-specific operand build from d_phase / d_amp (Nmc x kN, FP64).
int build_form_operands(rf_field_plan* P, hipStream_t s) {
    rf_context* ctx = P->ctx;
    const int64_t kN = P->spec.info.kNTotal;
    const int Nmc = P->p.Nmc;
    const double ampMult = P->spec.info.ampMult;
    if (P->p.form == RF_FORM_FP32) {
        if (is_iso_direct(P->p)) {
            if (!P->d_kIsoT) RF_HIP_TRY(ctx, dmalloc(&P->d_kIsoT, (size_t)(kN * P->Rpad * P->p.nDim)));
            RF_HIP_TRY(ctx, rf_launch_prep_fp32_iso(P->p.nDim, kN, Nmc, P->Rpad, P->d_kIso, P->d_phase,
                                                    P->d_amp, ampMult, P->d_kIsoT, P->d_phT, P->d_aT, s));
        } else {
            RF_HIP_TRY(ctx, rf_launch_prep_fp32(kN, Nmc, P->Rpad, P->d_phase, P->d_amp, ampMult,
                                                P->d_phT, P->d_aT, s));
        }
    } else if (P->p.form == RF_FORM_MFMA_F16 || P->p.form == RF_FORM_MFMA_BF16) {
        RF_HIP_TRY(ctx, rf_launch_mfma_prep(P->p.form, P->p.nDim, kN, P->kNpad, Nmc, P->Rpad,
                                            P->d_kRev, P->d_phase, P->d_amp, kN, ampMult * P->bScale,
                                            P->d_B, s));
    }
    return RF_OK;
}

}  // namespace

int rf_set_error(rf_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    g_last_error = msg;
    return code;
}

extern "C" {

int rf_version(void) { return RF_CAPI_VERSION; }

double rf_uniform_host(uint64_t seed, uint32_t tag, uint64_t e) { return rf_uniform(seed, tag, e); }

void rf_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    rf_u32x4 c;
    for (int i = 0; i < 4; i++) c.v[i] = ctr[i];
    const rf_u32x4 o = rf_philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; i++) out[i] = o.v[i];
}

const char* rf_last_error(const rf_context* ctx) {
    if (ctx && !ctx->last_error.empty()) return ctx->last_error.c_str();
    return g_last_error.c_str();
}

int rf_context_create(rf_context** out, int device) {
    if (!out) return rf_set_error(nullptr, RF_E_ARG, "ctx out-pointer is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return rf_set_error(nullptr, RF_E_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return rf_set_error(nullptr, RF_E_ARG, "device index out of range");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return rf_set_error(nullptr, RF_E_NODEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return rf_set_error(nullptr, RF_E_NODEVICE,
                            std::string("librf_hip.so is built for gfx950 only, device is ") + prop.gcnArchName);
    if (hipSetDevice(device) != hipSuccess) return rf_set_error(nullptr, RF_E_HIP, "hipSetDevice failed");
    const hipError_t se = rf_noalloc_setup(device);
    if (se != hipSuccess)
        return rf_set_error(nullptr, RF_E_HIP, std::string("static scratch lookup failed: ") + hipGetErrorString(se));
    auto* c = new (std::nothrow) rf_context();
    if (!c) return rf_set_error(nullptr, RF_E_NOMEM, "context allocation failed");
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return rf_set_error(nullptr, RF_E_HIP, "hipStreamCreate failed");
    }
    *out = c;
    return RF_OK;
}

int rf_context_destroy(rf_context* ctx) {
    if (!ctx) return RF_OK;
    (void)hipSetDevice(ctx->device);
    for (auto& c : ctx->planCache) rf_plan_destroy(c.plan);   // they synchronize on the stream
    ctx->planCache.clear();
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    dfree(ctx->d_scratch);
    dfree(ctx->d_io);
    delete ctx;
    return RF_OK;
}

int rf_context_sync(rf_context* ctx) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    RF_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return RF_OK;
}

int rf_plan_spectrum(const rf_params* p, rf_spectrum_info* info) {
    if (!info) return rf_set_error(nullptr, RF_E_ARG, "info is NULL");
    std::string err;
    const int rc = rf_spectrum_build_info(p, info, &err);
    if (rc) return rf_set_error(nullptr, rc, err);
    return RF_OK;
}

int rf_spectrum_tables(const rf_params* p, const double* kUniform, double* kPoints, double* Sk) {
    rf_spectrum_info info;
    std::string err;
    int rc = rf_spectrum_build_info(p, &info, &err);
    if (rc) return rf_set_error(nullptr, rc, err);
    rf_spectrum spec;
    rc = rf_spectrum_build_tables(p, info, kUniform, &spec, &err);
    if (rc) return rf_set_error(nullptr, rc, err);
    if (kPoints) std::memcpy(kPoints, spec.k.data(), spec.k.size() * sizeof(double));
    if (Sk) std::memcpy(Sk, spec.Sk.data(), spec.Sk.size() * sizeof(double));
    return RF_OK;
}

int rf_plan_destroy(rf_field_plan* P) {
    if (!P) return RF_OK;
    if (P->ctx) {
        (void)hipSetDevice(P->ctx->device);
        (void)hipStreamSynchronize(P->ctx->stream);
    }
    dfree(P->d_k); dfree(P->d_kRad); dfree(P->d_kRev); dfree(P->d_kIso); dfree(P->d_kIsoT);
    dfree(P->d_phase); dfree(P->d_amp); dfree(P->d_uni); dfree(P->d_phT);
    dfree(P->d_aT); dfree(P->d_B); dfree(P->d_Sk); dfree(P->d_calc); dfree(P->d_gridScratch);
    if (P->d_gridCoef) (void)hipFree(P->d_gridCoef);
    delete P;
    return RF_OK;
}

int rf_plan_create(rf_context* ctx, const rf_params* p, const double* kUniform,
                   const double* phaseUniform, const uint8_t* calculate, rf_field_plan** out) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!out) return rf_set_error(ctx, RF_E_ARG, "plan out-pointer is NULL");
    *out = nullptr;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::string err;
    rf_spectrum_info info;
    int rc = rf_spectrum_build_info(p, &info, &err);
    if (rc) return rf_set_error(ctx, rc, err);

    std::unique_ptr<rf_field_plan, int (*)(rf_field_plan*)> P(new (std::nothrow) rf_field_plan(),
                                                              rf_plan_destroy);
    if (!P) return rf_set_error(ctx, RF_E_NOMEM, "plan allocation failed");
    P->ctx = ctx;
    P->p = *p;
    const hipStream_t s = ctx->stream;
    const int d = p->nDim;
    const int Nmc = p->Nmc;
    const int64_t kN = info.kNTotal;

    // wavevector uniforms (method 3): caller's draws or the device stream
    std::vector<double> kU;
    if (info.nKUniform > 0 && !kUniform) {
        double* d_tmp = nullptr;
        RF_HIP_TRY(ctx, dmalloc(&d_tmp, (size_t)info.nKUniform));
        hipError_t e = rf_launch_uniforms(p->seed, 1u, info.nKUniform, d_tmp, s);
        kU.resize((size_t)info.nKUniform);
        if (e == hipSuccess) e = hipMemcpyAsync(kU.data(), d_tmp, kU.size() * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        dfree(d_tmp);
        if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("wavevector draws: ") + hipGetErrorString(e));
        kUniform = kU.data();
    }
    rc = rf_spectrum_build_tables(p, info, kUniform, &P->spec, &err);
    if (rc) return rf_set_error(ctx, rc, err);

    P->calc.assign((size_t)Nmc, 1);
    if (calculate)
        for (int r = 0; r < Nmc; r++) P->calc[r] = calculate[r] ? 1 : 0;
    P->nCalc = (int)std::count(P->calc.begin(), P->calc.end(), (uint8_t)1);
    P->epi = rf_make_epilogue(p->margiFirst, p->fieldAvg, p->fieldVar);

    // phase uniforms: Nmc x nPhaseUniform, realization-major
    const int64_t nPU = info.nPhaseUniform * Nmc;
    RF_HIP_TRY(ctx, dmalloc(&P->d_uni, (size_t)nPU));
    if (phaseUniform)
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_uni, phaseUniform, (size_t)nPU * sizeof(double), hipMemcpyHostToDevice, s));
    else
        RF_HIP_TRY(ctx, rf_launch_uniforms(p->seed, 2u, nPU, P->d_uni, s));

    RF_HIP_TRY(ctx, dmalloc(&P->d_phase, (size_t)(kN * Nmc)));
    RF_HIP_TRY(ctx, dmalloc(&P->d_amp, (size_t)(kN * Nmc)));
    RF_HIP_TRY(ctx, dmalloc(&P->d_calc, (size_t)Nmc));
    RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_calc, P->calc.data(), (size_t)Nmc, hipMemcpyHostToDevice, s));

    const bool mfma = (p->form == RF_FORM_MFMA_F16 || p->form == RF_FORM_MFMA_BF16);
    P->Rpad = mfma ? rf_mfma_npad(Nmc) : round_up(Nmc, rf_fp32_rchunk());
    P->kNpad = mfma ? (kN + rf_mfma_kpad(p->form) - 1) / rf_mfma_kpad(p->form) * rf_mfma_kpad(p->form) : kN;

    double maxAmp = 0.0;
    if (p->method == RF_FFT) {
        // gen_Std_Gauss_FFT (randomFieldND.f90:683-689): real DCT coefficients,
        // evaluated once; the grid form is its only executor
        RF_HIP_TRY(ctx, dmalloc(&P->d_Sk, (size_t)kN));
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_Sk, P->spec.Sk.data(), (size_t)kN * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, rf_launch_fft_coeffs(kN, Nmc, P->d_Sk, P->d_uni, P->d_calc, P->d_amp, P->d_phase, s));
        P->coeffsReady = true;
        double sMax = 0.0;
        for (double v : P->spec.Sk) sMax = std::max(sMax, v);
        P->ampMax = 0.5 * std::sqrt(sMax);          // |gamma - 1/2| <= 1/2, |cos| <= 1
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));
        *out = P.release();
        return RF_OK;
    }
    if (p->method == RF_ISOTROPIC) {
        RF_HIP_TRY(ctx, dmalloc(&P->d_kIso, (size_t)(kN * Nmc * d)));
        RF_HIP_TRY(ctx, dmalloc(&P->d_kRad, (size_t)kN));
        RF_HIP_TRY(ctx, dmalloc(&P->d_Sk, (size_t)kN));
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_kRad, P->spec.k.data(), (size_t)kN * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_Sk, P->spec.Sk.data(), (size_t)kN * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, rf_launch_iso_prep(d, p->mode, p->isoShared, kN, Nmc, P->d_kRad, P->d_Sk, P->d_uni,
                                           P->d_calc, P->d_kIso, P->d_amp, P->d_phase, s));
        if (p->isoShared) {
            // realization 0's direction table is the shared wavevector table
            RF_HIP_TRY(ctx, dmalloc(&P->d_k, (size_t)(kN * d)));
            RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_k, P->d_kIso, (size_t)(kN * d) * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        // |amp| <= sqrt(S k^2) * max|gamma|
        for (int64_t j = 0; j < kN; j++) {
            const double k = P->spec.k[j];
            maxAmp = std::max(maxAmp, std::sqrt(P->spec.Sk[j] * (d == 3 ? k * k : k)) * (d == 3 ? 1.7320508075688772 : 2.0 * RF_PI_REF));
        }
    } else {
        RF_HIP_TRY(ctx, dmalloc(&P->d_k, (size_t)(kN * d)));
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_k, P->spec.k.data(), (size_t)(kN * d) * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, rf_launch_phase_from_uniform(P->d_uni, kN * Nmc, P->d_phase, s));
        // REF_COMPAT: the sqrt chain Sk^(2^-c), c = 1, 2, ... until fixed point
        if (p->mode == RF_REF_COMPAT) {
            std::vector<double> row(P->spec.Sk);
            for (int c = 0; c < 4096; c++) {
                bool changed = false;
                for (auto& v : row) {
                    const double nv = std::sqrt(v);
                    changed |= (nv != v);
                    v = nv;
                }
                P->sqrtChain.push_back(row);
                if (!changed) break;
            }
        } else {
            std::vector<double> amp((size_t)(kN * Nmc));
            for (int r = 0; r < Nmc; r++)
                for (int64_t j = 0; j < kN; j++)
                    amp[(size_t)(r * kN + j)] = P->calc[r] ? std::sqrt(P->spec.Sk[j]) : 0.0;
            RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_amp, amp.data(), amp.size() * sizeof(double), hipMemcpyHostToDevice, s));
            RF_HIP_TRY(ctx, hipStreamSynchronize(s));   // amp leaves scope
        }
        for (double v : P->spec.Sk) maxAmp = std::max(maxAmp, std::sqrt(v));
        if (p->mode == RF_REF_COMPAT) maxAmp = std::max(maxAmp, 1.0);
    }

    // operand buffers of the chosen form
    if (p->form == RF_FORM_FP32) {
        RF_HIP_TRY(ctx, dmalloc(&P->d_phT, (size_t)(kN * P->Rpad)));
        RF_HIP_TRY(ctx, dmalloc(&P->d_aT, (size_t)(kN * P->Rpad)));
    }
    if (p->form != RF_FORM_FP64 && !is_iso_direct(*p)) {
        // the MFMA form reads whole K-steps: zero wavevectors pad the tail
        RF_HIP_TRY(ctx, dmalloc(&P->d_kRev, (size_t)(P->kNpad * d)));
        if (P->kNpad > kN)
            RF_HIP_TRY(ctx, hipMemsetAsync(P->d_kRev, 0, (size_t)(P->kNpad * d) * sizeof(double), s));
        RF_HIP_TRY(ctx, rf_launch_krev(kN * d, P->d_k, P->d_kRev, s));
    }
    P->ampMax = maxAmp;
    if (mfma) {
        P->bScale = rf_mfma_bscale(P->spec.info.ampMult, maxAmp);
        RF_HIP_TRY(ctx, hipMalloc(&P->d_B, rf_mfma_bbytes(p->form, P->kNpad, P->Rpad)));
    }
    const bool perSlab = (p->mode == RF_REF_COMPAT && p->method != RF_ISOTROPIC);
    if (!perSlab) {
        rc = build_form_operands(P.get(), s);
        if (rc) return rc;
        P->coeffsReady = true;
    }
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));
    *out = P.release();
    return RF_OK;
}

int rf_plan_info(const rf_field_plan* P, rf_spectrum_info* info) {
    if (!P || !info) return rf_set_error(nullptr, RF_E_ARG, "NULL argument");
    *info = P->spec.info;
    return RF_OK;
}

int rf_plan_execute(rf_field_plan* P, const double* d_x, int64_t N, void* d_out, int32_t outDtype,
                    void* stream) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "plan is NULL");
    rf_context* ctx = P->ctx;
    if (N < 0) return rf_set_error(ctx, RF_E_ARG, "N must be >= 0");
    if (N == 0) return RF_OK;
    if (!d_x || !d_out) return rf_set_error(ctx, RF_E_ARG, "NULL device buffer");
    if (outDtype != RF_DTYPE_F64 && outDtype != RF_DTYPE_F32)
        return rf_set_error(ctx, RF_E_ARG, "unknown output dtype");
    if (P->p.method == RF_FFT)
        return rf_set_error(ctx, RF_E_UNSUPPORTED, "method 4 (FFT) generates structured grids only: rf_plan_execute_grid");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const rf_params& p = P->p;
    const int64_t kN = P->spec.info.kNTotal;

    rf_sum_args a{};
    a.x = d_x;
    a.N = N;
    a.nDim = p.nDim;
    a.Nmc = p.Nmc;
    a.kN = kN;
    for (int i = 0; i < 3; i++) a.corrL[i] = (i < p.nDim) ? p.corrL[i] : 1.0;
    a.ampMult = P->spec.info.ampMult;
    a.epi = P->epi;
    a.out = d_out;
    a.outDtype = outDtype;

    auto launch_range = [&](int64_t i0, int64_t i1) -> int {
        a.i0 = i0;
        a.i1 = i1;
        hipError_t e = hipSuccess;
        switch (p.form) {
            case RF_FORM_FP64:
                if (is_iso_direct(p)) e = rf_launch_sum_fp64(a, P->d_kIso, kN * p.nDim, P->d_amp, P->d_phase, s);
                else e = rf_launch_sum_fp64(a, P->d_k, 0, P->d_amp, P->d_phase, s);
                break;
            case RF_FORM_FP32:
                if (is_iso_direct(p)) e = rf_launch_sum_fp32_iso_t(a, P->d_kIsoT, P->d_phT, P->d_aT, P->Rpad, s);
                else e = rf_launch_sum_fp32(a, P->d_kRev, P->d_phT, P->d_aT, P->Rpad, s);
                break;
            default:
                e = rf_launch_sum_mfma(p.form, P->mfmaVariant, a, P->kNpad, P->Rpad, P->d_B, 1.0 / P->bScale, s);
                break;
        }
        if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("sum kernel launch: ") + hipGetErrorString(e));
        return RF_OK;
    };

    if (P->coeffsReady) return launch_range(0, N);

    // REF_COMPAT methods 2/3: point slabs (randomFieldND.f90:237-239) and the
    // per-pass sqrt exponent (:312); pass c = rank(n) * xNSlab + m + 1.
    int64_t xSlab = p.xSlab > 0 ? p.xSlab : (int64_t)((double)p.nDim * 1e8 / (double)kN);
    if (xSlab > N) xSlab = N;
    if (xSlab < 1) xSlab = 1;
    const int64_t xNSlab = (N + xSlab - 1) / xSlab;
    std::vector<double> amp((size_t)(kN * p.Nmc));
    const int64_t lastRow = (int64_t)P->sqrtChain.size() - 1;
    for (int64_t m = 0; m < xNSlab; m++) {
        int rank = 0;
        for (int r = 0; r < p.Nmc; r++) {
            double* dst = amp.data() + (size_t)r * kN;
            if (!P->calc[r]) {
                std::fill(dst, dst + kN, 0.0);
                continue;
            }
            const int64_t c = std::min<int64_t>((int64_t)rank * xNSlab + m, lastRow);
            std::memcpy(dst, P->sqrtChain[(size_t)c].data(), (size_t)kN * sizeof(double));
            rank++;
        }
        RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_amp, amp.data(), amp.size() * sizeof(double), hipMemcpyHostToDevice, s));
        int rc = build_form_operands(P, s);
        if (rc) return rc;
        const int64_t i0 = m * xSlab;
        const int64_t i1 = std::min(N, i0 + xSlab);
        rc = launch_range(i0, i1);
        if (rc) return rc;
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));   // amp is reused next slab
    }
    return RF_OK;
}

}  // extern "C"

namespace {

// Re-forms what a plan builds from its phase uniforms d_uni: phases (methods
// 2/3), directions, amplitudes and phases (method 1), DCT coefficients
// (method 4), the form's operand tables; the grid form re-folds on its next
// execution.
int plan_reform(rf_field_plan* P, hipStream_t s) {
    rf_context* ctx = P->ctx;
    const rf_params& p = P->p;
    const int64_t kN = P->spec.info.kNTotal;
    if (p.method == RF_FFT) {
        RF_HIP_TRY(ctx, rf_launch_fft_coeffs(kN, p.Nmc, P->d_Sk, P->d_uni, P->d_calc, P->d_amp, P->d_phase, s));
    } else if (p.method == RF_ISOTROPIC) {
        // new directions, amplitudes and phases from the same radii
        RF_HIP_TRY(ctx, rf_launch_iso_prep(p.nDim, p.mode, p.isoShared, kN, p.Nmc, P->d_kRad, P->d_Sk, P->d_uni,
                                           P->d_calc, P->d_kIso, P->d_amp, P->d_phase, s));
        if (p.isoShared) {
            RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_k, P->d_kIso, (size_t)(kN * p.nDim) * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
            if (P->d_kRev) RF_HIP_TRY(ctx, rf_launch_krev(kN * p.nDim, P->d_k, P->d_kRev, s));
        }
    } else {
        RF_HIP_TRY(ctx, rf_launch_phase_from_uniform(P->d_uni, kN * p.Nmc, P->d_phase, s));
    }
    if (P->coeffsReady && p.method != RF_FFT) {
        const int rc = build_form_operands(P, s);
        if (rc) return rc;
    }
    P->gridCoefKey = NAN;
    return RF_OK;
}

// A plan of the context's cache for (p, kU, calc), refreshed with the
// phase draws phU (host) or, phU NULL, the device stream's draw 0 -- the
// same plan rf_plan_create would build -- or a new plan put in the cache.
int cached_plan(rf_context* ctx, const rf_params* p, const double* kU, const double* phU, const uint8_t* calc,
                rf_field_plan** out) {
    *out = nullptr;
    if (p) {
        for (auto& c : ctx->planCache) {
            // equal parameter bytes fix the table sizes, so the stored lengths apply
            if (std::memcmp(&c.p, p, sizeof(rf_params)) != 0) continue;
            rf_field_plan* P = c.plan;
            const bool useKU = kU && P->spec.info.nKUniform > 0;   // kU is read by method 3 only
            if (useKU != c.hasKU) continue;
            if (useKU && std::memcmp(c.kU.data(), kU, c.kU.size() * sizeof(double)) != 0) continue;
            if ((calc != nullptr) != c.hasCalc) continue;
            if (calc && std::memcmp(c.calc.data(), calc, c.calc.size()) != 0) continue;
            RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
            const hipStream_t s = ctx->stream;
            int rc;
            if (phU) {
                const size_t nPU = (size_t)(P->spec.info.nPhaseUniform * p->Nmc);
                RF_HIP_TRY(ctx, hipMemcpyAsync(P->d_uni, phU, nPU * sizeof(double), hipMemcpyHostToDevice, s));
                rc = plan_reform(P, s);
                if (rc) (void)hipStreamSynchronize(s);   // the copy reads the caller's phU
            } else {
                rc = rf_plan_redraw(P, 0, s);
            }
            if (rc) return rc;
            c.lastUse = ++ctx->useClock;
            *out = P;
            return RF_OK;
        }
    }
    rf_field_plan* P = nullptr;
    const int rc = rf_plan_create(ctx, p, kU, phU, calc, &P);
    if (rc) return rc;
    if (ctx->planCache.size() >= kPlanCache) {
        auto lru = std::min_element(ctx->planCache.begin(), ctx->planCache.end(),
                                    [](const rf_cached_plan& x, const rf_cached_plan& y) { return x.lastUse < y.lastUse; });
        rf_plan_destroy(lru->plan);
        ctx->planCache.erase(lru);
    }
    rf_cached_plan c;
    c.p = *p;
    c.hasKU = kU && P->spec.info.nKUniform > 0;
    if (c.hasKU) c.kU.assign(kU, kU + P->spec.info.nKUniform);
    c.hasCalc = calc != nullptr;
    if (calc) c.calc.assign(calc, calc + p->Nmc);
    c.plan = P;
    c.lastUse = ++ctx->useClock;
    ctx->planCache.push_back(std::move(c));
    *out = P;
    return RF_OK;
}

// The context's staging buffer for the host-buffer entry points, at least bytes.
int io_buffer(rf_context* ctx, size_t bytes, void** out) {
    if (bytes > ctx->ioBytes) {
        RF_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        dfree(ctx->d_io);
        ctx->d_io = nullptr;
        ctx->ioBytes = 0;
        if (hipMalloc(&ctx->d_io, bytes) != hipSuccess) return rf_set_error(ctx, RF_E_NOMEM, "staging allocation failed");
        ctx->ioBytes = bytes;
    }
    *out = ctx->d_io;
    return RF_OK;
}

}  // namespace

extern "C" {

int rf_plan_redraw(rf_field_plan* P, int64_t drawIndex, void* stream) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "plan is NULL");
    rf_context* ctx = P->ctx;
    if (drawIndex < 0) return rf_set_error(ctx, RF_E_ARG, "drawIndex must be >= 0");
    const rf_params& p = P->p;
    const int64_t kN = P->spec.info.kNTotal;
    const int64_t nPU = P->spec.info.nPhaseUniform * p.Nmc;
    if (nPU > 0 && (uint64_t)drawIndex > (UINT64_MAX - (uint64_t)nPU) / (uint64_t)nPU)
        return rf_set_error(ctx, RF_E_ARG, "drawIndex too large for the 64-bit draw counter");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    // realizations d*Nmc .. d*Nmc + Nmc - 1 of the phase stream: element
    // (d*Nmc + r) * nPhaseUniform + j, the element plan creation draws for d = 0
    const uint64_t first = (uint64_t)drawIndex * (uint64_t)nPU;
    if (p.method == RF_FFT) {
        // drawn inside the coefficient kernel (a third of them is discarded)
        RF_HIP_TRY(ctx, rf_launch_fft_coeffs_draw(kN, p.Nmc, P->d_Sk, P->d_calc, p.seed, first, P->d_amp, s));
        P->gridCoefKey = NAN;
        return RF_OK;
    }
    RF_HIP_TRY(ctx, rf_launch_uniforms_at(p.seed, 2u, first, nPU, P->d_uni, s));
    return plan_reform(P, s);
}

int rf_plan_set_variant(rf_field_plan* P, int32_t variant) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "plan is NULL");
    if (variant < 0 || variant >= rf_mfma_num_variants())
        return rf_set_error(P->ctx, RF_E_ARG, "MFMA variant out of range");
    P->mfmaVariant = variant;
    return RF_OK;
}

int rf_plan_autotune(rf_field_plan* P, const double* d_x, int64_t N, void* d_out, int32_t outDtype,
                     void* stream, int32_t* chosen) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "plan is NULL");
    rf_context* ctx = P->ctx;
    if (P->p.form != RF_FORM_MFMA_F16 && P->p.form != RF_FORM_MFMA_BF16) {
        if (chosen) *chosen = 0;
        return RF_OK;                           // one launch shape only
    }
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RF_HIP_TRY(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return rf_set_error(ctx, RF_E_HIP, "hipEventCreate failed");
    }
    int best = P->mfmaVariant, rc = RF_OK;
    float bestMs = 0.0f;
    for (int v = 0; v < rf_mfma_num_variants() && rc == RF_OK; v++) {
        P->mfmaVariant = v;
        rc = rf_plan_execute(P, d_x, N, d_out, outDtype, s);            // warm-up
        if (rc) break;
        hipError_t e = hipEventRecord(e0, s);
        for (int rep = 0; rep < 2 && rc == RF_OK && e == hipSuccess; rep++)
            rc = rf_plan_execute(P, d_x, N, d_out, outDtype, s);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = rf_set_error(ctx, RF_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
            break;
        }
        if (v == 0 || ms < bestMs) {
            bestMs = ms;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->mfmaVariant = best;
    if (chosen) *chosen = best;
    return rc;
}

}  // extern "C"

namespace {

// The isotropic method on a grid: the per-realization wavevectors, amplitudes
// and phases of the plan (FIXED or REF_COMPAT draws alike) summed by the
// factorised form of rf_launch_iso_grid.
int execute_grid_iso(rf_field_plan* P, const double* origin, const double* step, const int32_t* nstep,
                     void* d_out, int32_t outDtype, void* stream) {
    rf_context* ctx = P->ctx;
    const rf_params& p = P->p;
    if (p.nDim < 2) return rf_set_error(ctx, RF_E_UNSUPPORTED, "the isotropic grid form needs nDim 2 or 3");
    int64_t N = 1, plane = 1;
    for (int a = 0; a < p.nDim; a++) {
        if (nstep[a] < 1) return rf_set_error(ctx, RF_E_ARG, "grid counts must be >= 1");
        if (!(step[a] > 0.0)) return rf_set_error(ctx, RF_E_ARG, "grid steps must be > 0");
        N *= nstep[a];
        if (a + 1 < p.nDim) plane *= nstep[a];
    }
    const int64_t kN = P->spec.info.kNTotal;
    if (plane > INT32_MAX || kN > INT32_MAX) return rf_set_error(ctx, RF_E_ARG, "grid or shell set too large for the grid form");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const size_t bytes = rf_iso_grid_scratch_elems(p.nDim, kN, p.Nmc, nstep) * sizeof(float2);
    if (bytes > P->gridScratchBytes) {
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));
        dfree(P->d_gridScratch);
        P->d_gridScratch = nullptr;
        P->gridScratchBytes = 0;
        RF_HIP_TRY(ctx, hipMalloc(&P->d_gridScratch, bytes));
        P->gridScratchBytes = bytes;
    }
    rf_iso_grid_args a{};
    a.nDim = p.nDim;
    a.Nmc = p.Nmc;
    a.kN = kN;
    for (int d = 0; d < 3; d++) {
        a.n[d] = d < p.nDim ? nstep[d] : 1;
        a.origin[d] = d < p.nDim ? origin[d] : 0.0;
        a.step[d] = d < p.nDim ? step[d] : 1.0;
        a.invL[d] = d < p.nDim ? 1.0 / p.corrL[d] : 1.0;
    }
    a.kIso = P->d_kIso;
    a.amp = P->d_amp;
    a.phase = P->d_phase;
    a.ampMult = P->spec.info.ampMult;
    a.coefMax = a.ampMult * P->ampMax;
    a.epi = P->epi;
    a.d_out = d_out;
    a.outDtype = outDtype;
    a.scratch = static_cast<float2*>(P->d_gridScratch);
    RF_HIP_TRY(ctx, rf_launch_iso_grid(a, s));
    return RF_OK;
}

}  // namespace

extern "C" {

int rf_plan_execute_grid(rf_field_plan* P, const double* origin, const double* step,
                         const int32_t* nstep, void* d_out, int32_t outDtype, void* stream) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "plan is NULL");
    rf_context* ctx = P->ctx;
    const rf_params& p = P->p;
    if (!origin || !step || !nstep || !d_out) return rf_set_error(ctx, RF_E_ARG, "NULL grid argument");
    if (outDtype != RF_DTYPE_F64 && outDtype != RF_DTYPE_F32)
        return rf_set_error(ctx, RF_E_ARG, "unknown output dtype");
    const bool fft = p.method == RF_FFT;
    if (p.method == RF_ISOTROPIC) return execute_grid_iso(P, origin, step, nstep, d_out, outDtype, stream);
    if (!fft && (p.method != RF_SHINOZUKA || p.mode != RF_FIXED || p.nDim < 2))
        return rf_set_error(ctx, RF_E_UNSUPPORTED,
                            "the grid form needs the Shinozuka lattice in FIXED mode, the isotropic method "
                            "(nDim 2 or 3), or method 4");
    for (int a = 0; a < p.nDim; a++) {
        if (nstep[a] < 1) return rf_set_error(ctx, RF_E_ARG, "grid counts must be >= 1");
        if (!(step[a] > 0.0)) return rf_set_error(ctx, RF_E_ARG, "grid steps must be > 0");
        if (fft && nstep[a] != p.gridN[a])
            return rf_set_error(ctx, RF_E_ARG, "method 4: the grid counts must equal params.gridN (the DCT sizes)");
    }
    {
        // every GEMM dimension of the three stages is an int32
        const int32_t* kn = P->spec.info.kNStep;
        int64_t rows1 = (int64_t)P->spec.info.nSignBlocks * p.Nmc, plane = 1;
        for (int a = 1; a < p.nDim; a++) rows1 *= kn[a];
        for (int a = 0; a + 1 < p.nDim; a++) plane *= nstep[a];
        if (rows1 > INT32_MAX || plane > INT32_MAX)
            return rf_set_error(ctx, RF_E_ARG, "grid or lattice too large for the grid form");
    }
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const rf_spectrum_info& in = P->spec.info;
    const size_t bytes = rf_grid_scratch_elems(p.nDim, in.kNStep, in.nSignBlocks, p.Nmc, nstep) * sizeof(float2);
    if (bytes > P->gridScratchBytes) {
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));
        dfree(P->d_gridScratch);
        P->d_gridScratch = nullptr;
        P->gridScratchBytes = 0;
        RF_HIP_TRY(ctx, hipMalloc(&P->d_gridScratch, bytes));
        P->gridScratchBytes = bytes;
    }
    double corrL[3] = {1.0, 1.0, 1.0};
    for (int a = 0; a < p.nDim; a++) corrL[a] = p.corrL[a];
    double ampMult = in.ampMult;
    if (fft) {
        // randomFieldND.f90:707 (2-D) and :732 (3-D): the transform is scaled by
        // sqrt(prod xStep) and, in 3-D, by 2 more; xStep in correlation lengths
        double v = 1.0;
        for (int a = 0; a < p.nDim; a++) v *= step[a] / corrL[a];
        ampMult = (p.nDim == 3 ? 2.0 : 1.0) * std::sqrt(v);
    }
    rf_grid_stages st;
    if (!P->d_gridCoef) {
        int64_t nC = (int64_t)in.nSignBlocks * p.Nmc;
        for (int a = 0; a < p.nDim; a++) nC *= in.kNStep[a];
        RF_HIP_TRY(ctx, hipMalloc((void**)&P->d_gridCoef, (size_t)nC * sizeof(float2)));
        P->gridCoefKey = NAN;
    }
    rf_grid_coef coef{P->d_gridCoef, P->gridCoefKey};
    RF_HIP_TRY(ctx, rf_grid_prepare(fft ? 1 : 0, p.nDim, p.Nmc, in.kNStep, in.kDelta, in.nSignBlocks, origin, step,
                                    nstep, corrL, P->d_amp, P->nCalc == p.Nmc && !fft, P->d_uni, ampMult,
                                    ampMult * P->ampMax, P->epi, d_out,
                                    outDtype,
                                    static_cast<float2*>(P->d_gridScratch), s, &st, &coef));
    P->gridCoefKey = coef.key;
    RF_HIP_TRY(ctx, rf_grid_run(st, s));
    return RF_OK;
}

int rf_generate_grid(rf_context* ctx, const rf_params* p, const double* origin, const double* step,
                     const int32_t* nstep, const double* phaseUniform, const uint8_t* calculate, void* out,
                     int32_t outDtype) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!p || !origin || !step || !nstep || !out) return rf_set_error(ctx, RF_E_ARG, "NULL grid argument");
    if (outDtype != RF_DTYPE_F64 && outDtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown output dtype");
    int64_t N = 1;
    for (int a = 0; a < p->nDim && a < 3; a++) N *= nstep[a] > 0 ? nstep[a] : 0;
    if (N < 1) return rf_set_error(ctx, RF_E_ARG, "grid counts must be >= 1");
    rf_field_plan* P = nullptr;
    int rc = cached_plan(ctx, p, nullptr, phaseUniform, calculate, &P);
    if (rc) return rc;
    const size_t bytes = (size_t)N * p->Nmc * (outDtype == RF_DTYPE_F64 ? 8 : 4);
    void* d_out = nullptr;
    rc = io_buffer(ctx, bytes, &d_out);
    if (rc) return rc;
    rc = rf_plan_execute_grid(P, origin, step, nstep, d_out, outDtype, ctx->stream);
    hipError_t e = hipSuccess;
    if (rc == RF_OK) e = hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t es = hipStreamSynchronize(ctx->stream);   // also when a launch failed: phU is the caller's
    if (rc) return rc;
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("rf_generate_grid: ") + hipGetErrorString(e));
    return RF_OK;
}

int rf_generate(rf_context* ctx, const rf_params* p, const double* x, int64_t N,
                const double* kUniform, const double* phaseUniform, const uint8_t* calculate,
                void* out, int32_t outDtype) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (N < 0 || (N > 0 && (!x || !out))) return rf_set_error(ctx, RF_E_ARG, "bad point/output buffers");
    rf_field_plan* P = nullptr;
    int rc = cached_plan(ctx, p, kUniform, phaseUniform, calculate, &P);
    if (rc) return rc;
    if (N == 0) return RF_OK;
    const size_t esz = (outDtype == RF_DTYPE_F64) ? 8 : 4;
    // points, then the outputs at the next 256-byte boundary, in one staging buffer
    const size_t xBytes = ((size_t)(N * p->nDim) * sizeof(double) + 255) & ~(size_t)255;
    const size_t outBytes = (size_t)N * p->Nmc * esz;
    void* d_io = nullptr;
    rc = io_buffer(ctx, xBytes + outBytes, &d_io);
    if (rc) return rc;
    double* d_x = static_cast<double*>(d_io);
    void* d_out = static_cast<char*>(d_io) + xBytes;
    hipError_t e = hipMemcpyAsync(d_x, x, (size_t)(N * p->nDim) * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) rc = rf_plan_execute(P, d_x, N, d_out, outDtype, ctx->stream);
    if (e == hipSuccess && rc == RF_OK) e = hipMemcpyAsync(out, d_out, outBytes, hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("rf_generate: ") + hipGetErrorString(e));
    return RF_OK;
}

int rf_marginal_device(rf_context* ctx, int32_t margiFirst, double fieldAvg, double fieldVar,
                       void* d_field, int64_t count, int32_t dtype, void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (margiFirst < RF_FOM_NONE || margiFirst > RF_FOM_LOGNORMAL) return rf_set_error(ctx, RF_E_ARG, "unknown margiFirst");
    if (margiFirst == RF_FOM_LOGNORMAL && !(fieldAvg > 0.0))
        return rf_set_error(ctx, RF_E_LOGNORMAL, "ERROR - when using lognormal fieldAvg should be a positive number");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    if (count < 0 || (count > 0 && !d_field)) return rf_set_error(ctx, RF_E_ARG, "bad field buffer");
    if (margiFirst == RF_FOM_NONE || count == 0) return RF_OK;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const rf_epilogue e = rf_make_epilogue(margiFirst, fieldAvg, fieldVar);
    RF_HIP_TRY(ctx, rf_launch_marginal(margiFirst, e.scale, e.shift, d_field, count, dtype,
                                       stream ? (hipStream_t)stream : ctx->stream));
    return RF_OK;
}

int rf_marginal_host(rf_context* ctx, int32_t margiFirst, double fieldAvg, double fieldVar,
                     double* field, int64_t count) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (count < 0 || (count > 0 && !field)) return rf_set_error(ctx, RF_E_ARG, "bad field buffer");
    if (margiFirst == RF_FOM_LOGNORMAL && !(fieldAvg > 0.0))
        return rf_set_error(ctx, RF_E_LOGNORMAL, "ERROR - when using lognormal fieldAvg should be a positive number");
    if (count == 0 || margiFirst == RF_FOM_NONE) return RF_OK;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* d = nullptr;
    RF_HIP_TRY(ctx, dmalloc(&d, (size_t)count));
    hipError_t e = hipMemcpyAsync(d, field, (size_t)count * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    int rc = RF_OK;
    if (e == hipSuccess) rc = rf_marginal_device(ctx, margiFirst, fieldAvg, fieldVar, d, count, RF_DTYPE_F64, ctx->stream);
    if (e == hipSuccess && rc == RF_OK)
        e = hipMemcpyAsync(field, d, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(d);
    if (rc) return rc;
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("rf_marginal_host: ") + hipGetErrorString(e));
    return RF_OK;
}

}  // extern "C"

// ---- statistics, normalisation and blend -----------------------------------

namespace {

int ensure_scratch(rf_context* ctx, size_t bytes) {
    if (ctx->scratch_bytes >= bytes) return RF_OK;
    dfree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    RF_HIP_TRY(ctx, hipMalloc(&ctx->d_scratch, bytes));
    ctx->scratch_bytes = bytes;
    return RF_OK;
}

// Sums of u and u^2 per realization into host sums[2*Nmc] (synchronous).
int moments(rf_context* ctx, const void* d, int64_t N, int Nmc, int dtype, hipStream_t s,
            std::vector<double>* sums) {
    const int nPart = rf_moments_parts(N);
    const size_t bytes = sizeof(double) * ((size_t)nPart * Nmc * 2 + (size_t)Nmc * 2);
    int rc = ensure_scratch(ctx, bytes);
    if (rc) return rc;
    double* part = static_cast<double*>(ctx->d_scratch);
    double* dsum = part + (size_t)nPart * Nmc * 2;
    RF_HIP_TRY(ctx, rf_launch_moments(d, N, Nmc, dtype, part, nPart, dsum, s));
    sums->resize((size_t)Nmc * 2);
    RF_HIP_TRY(ctx, hipMemcpyAsync(sums->data(), dsum, sizeof(double) * Nmc * 2, hipMemcpyDeviceToHost, s));
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));
    return RF_OK;
}

}  // namespace

extern "C" {

int rf_stats_device(rf_context* ctx, const void* d_field, int64_t N, int32_t Nmc, int32_t dtype,
                    double* avg, double* std, double* g) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (N < 1 || Nmc < 1 || !d_field) return rf_set_error(ctx, RF_E_ARG, "bad field");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<double> sums;
    int rc = moments(ctx, d_field, N, Nmc, dtype, ctx->stream, &sums);
    if (rc) return rc;
    // statistics_RF.f90:81-87 (biased, E[u^2] - E[u]^2)
    double t1 = 0.0, t2 = 0.0;
    for (int r = 0; r < Nmc; r++) {
        const double m = sums[2 * r] / (double)N;
        if (avg) avg[r] = m;
        if (std) std[r] = std::sqrt(sums[2 * r + 1] / (double)N - m * m);
        t1 += sums[2 * r];
        t2 += sums[2 * r + 1];
    }
    if (g) {
        g[0] = t1 / ((double)N * Nmc);
        g[1] = std::sqrt(t2 / ((double)N * Nmc) - g[0] * g[0]);
    }
    return RF_OK;
}

int rf_corr_device(rf_context* ctx, const void* d_field, int32_t dtype, int32_t nDim, const int32_t* nstep,
                   const double* xStep, int32_t realization, double avg, double std, double* R, double* corrL) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!d_field || !nstep || !R || nDim < 1 || nDim > 3) return rf_set_error(ctx, RF_E_ARG, "bad field or grid");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    if (!(std > 0.0) || realization < 0) return rf_set_error(ctx, RF_E_ARG, "std must be > 0, realization >= 0");
    int64_t N = 1, stride[3] = {1, 1, 1};
    int Mtot = 0, Mmax = 0;
    for (int a = 0; a < nDim; a++) {
        if (nstep[a] < 1) return rf_set_error(ctx, RF_E_ARG, "grid counts must be >= 1");
        if (a > 0) stride[a] = stride[a - 1] * nstep[a - 1];
        N *= nstep[a];
        Mtot += nstep[a] / 2;
        Mmax = std::max(Mmax, nstep[a] / 2);
    }
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = ctx->stream;
    const int nPart = rf_corr_parts(N);
    const size_t bytes = sizeof(double) * ((size_t)nPart * Mmax + (size_t)Mtot);
    int rc = ensure_scratch(ctx, bytes);
    if (rc) return rc;
    double* part = static_cast<double*>(ctx->d_scratch);
    double* dR = part + (size_t)nPart * Mmax;
    const size_t esz = dtype == RF_DTYPE_F64 ? 8 : 4;
    const void* u = static_cast<const char*>(d_field) + (size_t)realization * N * esz;
    int off = 0;
    for (int a = 0; a < nDim; a++) {
        const int M = nstep[a] / 2;
        RF_HIP_TRY(ctx, rf_launch_corr(u, dtype, N, stride[a], nstep[a], M, avg, std, part, nPart, dR + off, s));
        off += M;
    }
    RF_HIP_TRY(ctx, hipMemcpyAsync(R, dR, sizeof(double) * Mtot, hipMemcpyDeviceToHost, s));
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));
    // rebuild_corrL (statistics_RF.f90:438-445): twice the trapezoid of R
    if (corrL) {
        const double* r = R;
        for (int a = 0; a < nDim; a++) {
            const int M = nstep[a] / 2;
            double sum = 0.0;
            for (int m = 0; m < M; m++) sum += r[m];
            corrL[a] = M > 0 ? 2.0 * (sum - (r[0] + r[M - 1]) / 2.0) * (xStep ? xStep[a] : 1.0) : 0.0;
            r += M;
        }
    }
    return RF_OK;
}

}  // extern "C"

namespace {

int interpolate_common(rf_context* ctx, int32_t nDim, const int32_t* nstep, const int32_t* lo, const double* xMin,
                       const double* xStep, int32_t Nmc, const void* d_grid, int32_t gridDtype,
                       const double* d_nodes, int64_t nNodes, void* d_out, int32_t outDtype, void* stream,
                       int32_t* outside) {
    if (outside) *outside = 0;
    if (nNodes == 0) return RF_OK;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    int rc = ensure_scratch(ctx, sizeof(int) * 64);
    if (rc) return rc;
    int* flag = static_cast<int*>(ctx->d_scratch);
    RF_HIP_TRY(ctx, hipMemsetAsync(flag, 0, sizeof(int), s));
    RF_HIP_TRY(ctx, rf_launch_interp(nDim, nstep, lo, xMin, xStep, Nmc, d_grid, gridDtype, d_nodes, nNodes, d_out,
                                     outDtype, flag, s));
    if (outside) {
        int h = 0;
        RF_HIP_TRY(ctx, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));
        *outside = h;
    } else {
        RF_HIP_TRY(ctx, hipStreamSynchronize(s));   // the flag word is context scratch
    }
    return RF_OK;
}

int interpolate_args(rf_context* ctx, int32_t nDim, const double* xMin, const double* xStep, int32_t Nmc,
                     const void* d_grid, int32_t gridDtype, const double* d_nodes, int64_t nNodes, void* d_out,
                     int32_t outDtype) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (nDim < 1 || nDim > 3 || !xMin || !xStep || Nmc < 1 || nNodes < 0)
        return rf_set_error(ctx, RF_E_ARG, "bad interpolation grid");
    if (nNodes > 0 && (!d_grid || !d_nodes || !d_out)) return rf_set_error(ctx, RF_E_ARG, "NULL device buffer");
    if ((gridDtype != RF_DTYPE_F64 && gridDtype != RF_DTYPE_F32) || (outDtype != RF_DTYPE_F64 && outDtype != RF_DTYPE_F32))
        return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    for (int a = 0; a < nDim; a++)
        if (!(xStep[a] > 0.0)) return rf_set_error(ctx, RF_E_ARG, "grid step must be > 0 per axis");
    return RF_OK;
}

}  // namespace

extern "C" {

int rf_interpolate_device(rf_context* ctx, int32_t nDim, const int32_t* nstep, const double* xMin,
                          const double* xStep, int32_t Nmc, const void* d_grid, int32_t gridDtype,
                          const double* d_nodes, int64_t nNodes, void* d_out, int32_t outDtype, void* stream,
                          int32_t* outside) {
    int rc = interpolate_args(ctx, nDim, xMin, xStep, Nmc, d_grid, gridDtype, d_nodes, nNodes, d_out, outDtype);
    if (rc) return rc;
    if (!nstep) return rf_set_error(ctx, RF_E_ARG, "bad interpolation grid");
    for (int a = 0; a < nDim; a++)
        if (nstep[a] < 2) return rf_set_error(ctx, RF_E_ARG, "grid needs >= 2 points and step > 0 per axis");
    return interpolate_common(ctx, nDim, nstep, nullptr, xMin, xStep, Nmc, d_grid, gridDtype, d_nodes, nNodes,
                              d_out, outDtype, stream, outside);
}

int rf_interpolate_box_device(rf_context* ctx, int32_t nDim, const int32_t* minPos, const int32_t* maxPos,
                              const double* xMinGlob, const double* xStep, int32_t Nmc, const void* d_box,
                              int32_t gridDtype, const double* d_nodes, int64_t nNodes, void* d_out,
                              int32_t outDtype, void* stream, int32_t* outside) {
    int rc = interpolate_args(ctx, nDim, xMinGlob, xStep, Nmc, d_box, gridDtype, d_nodes, nNodes, d_out, outDtype);
    if (rc) return rc;
    if (!minPos || !maxPos) return rf_set_error(ctx, RF_E_ARG, "bad interpolation box");
    int32_t ext[3] = {1, 1, 1};
    for (int a = 0; a < nDim; a++) {
        if (minPos[a] < 1 || maxPos[a] < minPos[a]) return rf_set_error(ctx, RF_E_ARG, "bad interpolation box");
        ext[a] = maxPos[a] - minPos[a] + 1;
    }
    return interpolate_common(ctx, nDim, ext, minPos, xMinGlob, xStep, Nmc, d_box, gridDtype, d_nodes, nNodes,
                              d_out, outDtype, stream, outside);
}

int rf_normalize_device(rf_context* ctx, void* d_field, int64_t N, int32_t Nmc, int32_t dtype,
                        void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (N < 1 || Nmc < 1 || !d_field) return rf_set_error(ctx, RF_E_ARG, "bad field");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    std::vector<double> sums;
    int rc = moments(ctx, d_field, N, Nmc, dtype, s, &sums);
    if (rc) return rc;
    // randomFieldND.f90:869-875: subtract the mean
    std::vector<double> hv((size_t)Nmc * 2);
    for (int r = 0; r < Nmc; r++) { hv[r] = sums[2 * r] / (double)N; hv[Nmc + r] = 1.0; }
    double* dv = nullptr;
    RF_HIP_TRY(ctx, dmalloc(&dv, hv.size()));
    hipError_t e = hipMemcpyAsync(dv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = rf_launch_affine(d_field, N, Nmc, dtype, dv, dv + Nmc, s);
    if (e != hipSuccess) { dfree(dv); return rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e)); }
    rc = moments(ctx, d_field, N, Nmc, dtype, s, &sums);
    if (rc) { dfree(dv); return rc; }
    // :904-908: divide by the root mean square
    for (int r = 0; r < Nmc; r++) { hv[r] = 0.0; hv[Nmc + r] = 1.0 / std::sqrt(sums[2 * r + 1] / (double)N); }
    e = hipMemcpyAsync(dv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = rf_launch_affine(d_field, N, Nmc, dtype, dv, dv + Nmc, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(dv);
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    return RF_OK;
}

int rf_normalize_window_device(rf_context* ctx, void* d_field, int32_t nDim, const int32_t* gridN,
                               const int32_t* minPos, const int32_t* maxPos, int32_t Nmc, int32_t dtype,
                               void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!d_field || !gridN || !minPos || !maxPos || nDim < 1 || nDim > 3 || Nmc < 1)
        return rf_set_error(ctx, RF_E_ARG, "bad field or window");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    rf_window w{};
    int64_t N = 1;
    w.count = 1;
    for (int a = 0; a < 3; a++) {
        w.n[a] = a < nDim ? gridN[a] : 1;
        w.lo[a] = a < nDim ? minPos[a] - 1 : 0;
        w.ext[a] = a < nDim ? (int64_t)maxPos[a] - minPos[a] + 1 : 1;
        if (w.n[a] < 1 || w.lo[a] < 0 || w.ext[a] < 1 || w.lo[a] + w.ext[a] > w.n[a])
            return rf_set_error(ctx, RF_E_ARG, "window outside the grid (1 <= minPos <= maxPos <= gridN)");
        N *= w.n[a];
        w.count *= w.ext[a];
    }
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    int nPart = rf_moments_parts(w.count);
    const size_t bytes = sizeof(double) * ((size_t)nPart * 2 + 2 + 2 * (size_t)Nmc);
    int rc = ensure_scratch(ctx, bytes);
    if (rc) return rc;
    double* part = static_cast<double*>(ctx->d_scratch);
    double* dsum = part + (size_t)nPart * 2;
    double* dv = dsum + 2;                       // shift[Nmc], mul[Nmc]
    std::vector<double> hv((size_t)Nmc * 2);
    double hs[2];
    // randomFieldND.f90:856-875: window mean of realization 1, subtracted
    // from every realization
    RF_HIP_TRY(ctx, rf_launch_window_moments(d_field, dtype, w, part, nPart, dsum, s));
    RF_HIP_TRY(ctx, hipMemcpyAsync(hs, dsum, sizeof(hs), hipMemcpyDeviceToHost, s));
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));
    const double avg = hs[0] / (double)w.count;
    for (int r = 0; r < Nmc; r++) { hv[r] = avg; hv[Nmc + r] = 1.0; }
    RF_HIP_TRY(ctx, hipMemcpyAsync(dv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s));
    RF_HIP_TRY(ctx, rf_launch_affine(d_field, N, Nmc, dtype, dv, dv + Nmc, s));
    // :891-908: window root mean square of realization 1, every realization
    // divided by it
    RF_HIP_TRY(ctx, rf_launch_window_moments(d_field, dtype, w, part, nPart, dsum, s));
    RF_HIP_TRY(ctx, hipMemcpyAsync(hs, dsum, sizeof(hs), hipMemcpyDeviceToHost, s));
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));
    const double sd = std::sqrt(hs[1] / (double)w.count);
    for (int r = 0; r < Nmc; r++) { hv[r] = 0.0; hv[Nmc + r] = 1.0 / sd; }
    RF_HIP_TRY(ctx, hipMemcpyAsync(dv, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s));
    RF_HIP_TRY(ctx, rf_launch_affine(d_field, N, Nmc, dtype, dv, dv + Nmc, s));
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));     // the coefficients are context scratch
    return RF_OK;
}

int rf_normalize_window_host(rf_context* ctx, double* field, int32_t nDim, const int32_t* gridN,
                             const int32_t* minPos, const int32_t* maxPos, int32_t Nmc) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!field || !gridN || nDim < 1 || nDim > 3 || Nmc < 1) return rf_set_error(ctx, RF_E_ARG, "bad field");
    int64_t N = 1;
    for (int a = 0; a < nDim; a++) {
        if (gridN[a] < 1) return rf_set_error(ctx, RF_E_ARG, "grid counts must be >= 1");
        N *= gridN[a];
    }
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)N * Nmc * sizeof(double);
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return rf_set_error(ctx, RF_E_NOMEM, "field allocation failed");
    hipError_t e = hipMemcpyAsync(d, field, bytes, hipMemcpyHostToDevice, ctx->stream);
    int rc = e == hipSuccess ? rf_normalize_window_device(ctx, d, nDim, gridN, minPos, maxPos, Nmc, RF_DTYPE_F64,
                                                          nullptr)
                             : rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    if (rc == RF_OK) {
        e = hipMemcpyAsync(field, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

int rf_device_uniforms(rf_context* ctx, uint64_t seed, uint32_t tag, int64_t count, double* d_out,
                       void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (count < 0 || (count > 0 && !d_out)) return rf_set_error(ctx, RF_E_ARG, "bad output");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    RF_HIP_TRY(ctx, rf_launch_uniforms(seed, tag, count, d_out, stream ? (hipStream_t)stream : ctx->stream));
    return RF_OK;
}

}  // extern "C"

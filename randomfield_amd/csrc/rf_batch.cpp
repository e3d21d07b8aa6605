// rf_batch.cpp -- batched structured-grid generation (all subdomains of a
// localization group in three grouped launches) and graph capture of a
// stream's launches.
//
// The reference generates each subdomain on its own MPI rank
// (localization_RF.f90:127-192 lays them out, calls_RF.f90:250-296 runs one
// createRandomField per rank); here each subdomain is one field of a batch.
// rf_grid_batch_create does everything that allocates or depends only on the
// geometry and the phases (phase tables, coefficients, GEMM descriptors);
// rf_grid_batch_execute then enqueues three grouped GEMM launches and nothing
// else, so it can be captured into a graph and replayed.
#include <cstring>
#include <cmath>
#include <memory>
#include <new>
#include <vector>

#include "rf_internal.h"

struct rf_grid_batch {
    rf_context* ctx = nullptr;
    char* d_mem = nullptr;              // scratch of every field, then descriptors
    rf_gemm_desc* d_desc[3] = {nullptr, nullptr, nullptr};
    int nDesc[3] = {0, 0, 0};
    int64_t tiles[3] = {0, 0, 0};
    int realOut[3] = {0, 0, 1};
    int rwMax[3] = {1, 1, 1};           // wide tile kind used by a stage's GEMMs (1: none)
};

struct rf_graph {
    rf_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {
size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

extern "C" {

int rf_grid_batch_create(rf_context* ctx, int32_t n, rf_field_plan* const* plans, const double* origins,
                         const double* steps, const int32_t* nsteps, void* const* d_outs, int32_t outDtype,
                         rf_grid_batch** out) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!out) return rf_set_error(ctx, RF_E_ARG, "batch out-pointer is NULL");
    *out = nullptr;
    if (n < 1 || !plans || !origins || !steps || !nsteps || !d_outs)
        return rf_set_error(ctx, RF_E_ARG, "NULL or empty batch argument");
    if (outDtype != RF_DTYPE_F64 && outDtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown output dtype");
    // validate and size
    std::vector<size_t> off((size_t)n);
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        const rf_field_plan* P = plans[i];
        if (!P || !d_outs[i]) return rf_set_error(ctx, RF_E_ARG, "NULL plan or output in batch");
        if (P->ctx != ctx) return rf_set_error(ctx, RF_E_ARG, "plan of another context in batch");
        const rf_params& p = P->p;
        const bool fft = p.method == RF_FFT;
        if (!fft && (p.method != RF_SHINOZUKA || p.mode != RF_FIXED || p.nDim < 2))
            return rf_set_error(ctx, RF_E_UNSUPPORTED,
                                "the grid form needs the Shinozuka lattice, FIXED mode and nDim 2 or 3, or method 4");
        const int32_t* ns = nsteps + 3 * i;
        const rf_spectrum_info& in = P->spec.info;
        int64_t rows1 = (int64_t)in.nSignBlocks * p.Nmc, plane = 1;
        for (int a = 0; a < p.nDim; a++) {
            if (ns[a] < 1 || !(steps[3 * i + a] > 0.0)) return rf_set_error(ctx, RF_E_ARG, "bad grid counts or steps");
            if (fft && ns[a] != p.gridN[a])
                return rf_set_error(ctx, RF_E_ARG, "method 4: the grid counts must equal params.gridN (the DCT sizes)");
            if (a > 0) rows1 *= in.kNStep[a];
            if (a + 1 < p.nDim) plane *= ns[a];
        }
        if (rows1 > INT32_MAX || plane > INT32_MAX)
            return rf_set_error(ctx, RF_E_ARG, "grid or lattice too large for the grid form");
        off[(size_t)i] = total;
        total += align256(rf_grid_scratch_elems(p.nDim, in.kNStep, in.nSignBlocks, p.Nmc, ns) * sizeof(float2));
    }
    const size_t descOff = total;
    total += align256(sizeof(rf_gemm_desc) * 3 * (size_t)n);
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<rf_grid_batch> B(new (std::nothrow) rf_grid_batch());
    if (!B) return rf_set_error(ctx, RF_E_NOMEM, "batch allocation failed");
    B->ctx = ctx;
    if (hipMalloc((void**)&B->d_mem, total) != hipSuccess)
        return rf_set_error(ctx, RF_E_NOMEM, "batch scratch allocation failed");
    const hipStream_t s = ctx->stream;
    std::vector<rf_gemm_desc> stage[3];
    for (int i = 0; i < n; i++) {
        const rf_field_plan* P = plans[i];
        const rf_params& p = P->p;
        const rf_spectrum_info& in = P->spec.info;
        double o[3] = {0, 0, 0}, st[3] = {1, 1, 1}, corrL[3] = {1, 1, 1};
        int32_t ns[3] = {1, 1, 1};
        for (int a = 0; a < p.nDim; a++) {
            o[a] = origins[3 * i + a];
            st[a] = steps[3 * i + a];
            ns[a] = nsteps[3 * i + a];
            corrL[a] = p.corrL[a];
        }
        const bool fft = p.method == RF_FFT;
        double ampMult = in.ampMult;
        if (fft) {
            // the DCT scale of randomFieldND.f90:707 and :732, as in rf_plan_execute_grid
            double v = 1.0;
            for (int a = 0; a < p.nDim; a++) v *= st[a] / corrL[a];
            ampMult = (p.nDim == 3 ? 2.0 : 1.0) * std::sqrt(v);
        }
        rf_grid_stages gs;
        const hipError_t e = rf_grid_prepare(fft ? 1 : 0, p.nDim, p.Nmc, in.kNStep, in.kDelta, in.nSignBlocks, o, st,
                                             ns, corrL, P->d_amp, P->d_phase, ampMult, ampMult * P->ampMax, P->epi,
                                             d_outs[i], outDtype,
                                             reinterpret_cast<float2*>(B->d_mem + off[(size_t)i]), s, &gs);
        if (e != hipSuccess) {
            (void)hipFree(B->d_mem);
            return rf_set_error(ctx, RF_E_HIP, std::string("batch tables: ") + hipGetErrorString(e));
        }
        // stage slots: 0 first contraction, 1 middle (3-D), 2 last (real)
        stage[0].push_back(gs.g[0]);
        if (gs.n == 3) stage[1].push_back(gs.g[1]);
        stage[2].push_back(gs.g[gs.n - 1]);
    }
    std::vector<rf_gemm_desc> all;
    size_t dOff = 0;
    for (int k = 0; k < 3; k++) {
        int64_t t = 0;
        for (auto& g : stage[k]) {
            g.tile0 = t;
            t += rf_gemm_tiles(g);
            if (g.rw > B->rwMax[k]) B->rwMax[k] = g.rw;
        }
        B->tiles[k] = t;
        B->nDesc[k] = (int)stage[k].size();
        B->d_desc[k] = reinterpret_cast<rf_gemm_desc*>(B->d_mem + descOff) + dOff;
        all.insert(all.end(), stage[k].begin(), stage[k].end());
        dOff += stage[k].size();
        if (t > 0x7fffffffLL) {
            (void)hipFree(B->d_mem);
            return rf_set_error(ctx, RF_E_ARG, "batch too large for one grouped launch");
        }
    }
    hipError_t e = hipMemcpyAsync(B->d_mem + descOff, all.data(), all.size() * sizeof(rf_gemm_desc),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipFree(B->d_mem);
        return rf_set_error(ctx, RF_E_HIP, std::string("batch descriptors: ") + hipGetErrorString(e));
    }
    *out = B.release();
    return RF_OK;
}

int rf_grid_batch_execute(rf_grid_batch* B, void* stream) {
    if (!B) return rf_set_error(nullptr, RF_E_ARG, "batch is NULL");
    rf_context* ctx = B->ctx;
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    for (int k = 0; k < 3; k++)
        RF_HIP_TRY(ctx, rf_launch_gemm_group(B->realOut[k], B->d_desc[k], B->nDesc[k], B->tiles[k], B->rwMax[k], s));
    return RF_OK;
}

int rf_grid_batch_destroy(rf_grid_batch* B) {
    if (!B) return RF_OK;
    (void)hipSetDevice(B->ctx->device);
    (void)hipFree(B->d_mem);
    delete B;
    return RF_OK;
}

int rf_graph_capture_begin(rf_context* ctx, void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!stream) return rf_set_error(ctx, RF_E_ARG, "capture needs an explicit stream");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    RF_HIP_TRY(ctx, hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return RF_OK;
}

int rf_graph_capture_end(rf_context* ctx, void* stream, rf_graph** out) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!stream || !out) return rf_set_error(ctx, RF_E_ARG, "NULL stream or graph out-pointer");
    *out = nullptr;
    std::unique_ptr<rf_graph> G(new (std::nothrow) rf_graph());
    if (!G) return rf_set_error(ctx, RF_E_NOMEM, "graph allocation failed");
    G->ctx = ctx;
    RF_HIP_TRY(ctx, hipStreamEndCapture((hipStream_t)stream, &G->graph));
    const hipError_t e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(G->graph);
        return rf_set_error(ctx, RF_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *out = G.release();
    return RF_OK;
}

int rf_graph_launch(rf_graph* G, void* stream) {
    if (!G) return rf_set_error(nullptr, RF_E_ARG, "graph is NULL");
    const hipStream_t s = stream ? (hipStream_t)stream : G->ctx->stream;
    RF_HIP_TRY(G->ctx, hipGraphLaunch(G->exec, s));
    return RF_OK;
}

int rf_graph_destroy(rf_graph* G) {
    if (!G) return RF_OK;
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    delete G;
    return RF_OK;
}

}  // extern "C"

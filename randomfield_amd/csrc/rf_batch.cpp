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

#include "rf_blend.h"
#include "rf_internal.h"

struct rf_grid_batch {
    rf_context* ctx = nullptr;
    char* d_mem = nullptr;              // scratch of every field, then descriptors
    rf_gemm_desc* d_desc[3] = {nullptr, nullptr, nullptr};
    int nDesc[3] = {0, 0, 0};
    int64_t tiles[3] = {0, 0, 0};
    int realOut[3] = {0, 0, 1};
    int rwMax[3] = {1, 1, 1};           // wide tile kind used by a stage's GEMMs (1: none)
    // fresh realizations: every field's coefficient fold, its seed and Nmc,
    // and the device draw counter (realizations counter * Nmc .. of each
    // field's device stream)
    std::vector<rf_fold_desc> folds;
    std::vector<uint64_t> seeds;
    std::vector<int> nmc;
    bool redrawable = true;
    int64_t* d_draw = nullptr;          // draw state: [0] next, [1] current, [2] coefficient mode
    bool genActive = false;             // the last redraw left the first contraction forming the coefficients
    rf_fold_item* d_items = nullptr;    // the folds as one grouped launch
    int64_t foldBlocks = 0;
    // host copies of the stage descriptors and the tile kinds the set-up
    // chose (the autotuner switches a stage between those and 64 x 64)
    std::vector<rf_gemm_desc> stage[3];
    std::vector<int> kind0[3];
    int outDtype = RF_DTYPE_F32;
};

struct rf_graph {
    rf_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {
size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// tile offsets, tile counts and widest kind of every stage, uploaded
hipError_t upload_stages(rf_grid_batch* B, hipStream_t s) {
    std::vector<rf_gemm_desc> all;
    for (int k = 0; k < 3; k++) {
        int64_t t = 0;
        B->rwMax[k] = 1;
        for (auto& g : B->stage[k]) {
            g.tile0 = t;
            t += rf_gemm_tiles(g);
            if (g.rw > B->rwMax[k]) B->rwMax[k] = g.rw;
        }
        B->tiles[k] = t;
        all.insert(all.end(), B->stage[k].begin(), B->stage[k].end());
    }
    for (int k = 0; k < 3; k++)
        if (B->tiles[k] > 0x7fffffffLL) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(B->d_desc[0], all.data(), all.size() * sizeof(rf_gemm_desc), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}
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
    const size_t drawOff = total;
    total += 256;
    const size_t itemOff = total;
    total += align256(sizeof(rf_fold_item) * (size_t)n);
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<rf_grid_batch> B(new (std::nothrow) rf_grid_batch());
    if (!B) return rf_set_error(ctx, RF_E_NOMEM, "batch allocation failed");
    B->ctx = ctx;
    B->outDtype = outDtype;
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
                                             ns, corrL, P->d_amp, P->nCalc == p.Nmc && !fft, P->d_uni, ampMult,
                                             ampMult * P->ampMax, P->epi,
                                             d_outs[i], outDtype,
                                             reinterpret_cast<float2*>(B->d_mem + off[(size_t)i]), s, &gs);
        if (e != hipSuccess) {
            (void)hipFree(B->d_mem);
            return rf_set_error(ctx, RF_E_HIP, std::string("batch tables: ") + hipGetErrorString(e));
        }
        B->folds.push_back(gs.fold);
        B->seeds.push_back(p.seed);
        B->nmc.push_back(p.Nmc);
        if (fft) B->redrawable = false;
        // stage slots: 0 first contraction, 1 middle (3-D), 2 last (real)
        stage[0].push_back(gs.g[0]);
        if (gs.n == 3) stage[1].push_back(gs.g[1]);
        stage[2].push_back(gs.g[gs.n - 1]);
    }
    B->d_draw = reinterpret_cast<int64_t*>(B->d_mem + drawOff);
    if (B->redrawable)
        for (size_t i = 0; i < stage[0].size(); i++) {
            // the first contraction can form fresh coefficients itself
            stage[0][i].gen = B->folds[i];
            stage[0][i].genSeed = B->seeds[i];
            stage[0][i].genNmc = B->nmc[i];
            stage[0][i].genState = B->d_draw;
        }
    size_t dOff = 0;
    for (int k = 0; k < 3; k++) {
        B->stage[k] = stage[k];
        for (const auto& g : stage[k]) B->kind0[k].push_back(g.rw);
        B->nDesc[k] = (int)stage[k].size();
        B->d_desc[k] = reinterpret_cast<rf_gemm_desc*>(B->d_mem + descOff) + dOff;
        dOff += stage[k].size();
    }
    hipError_t e = upload_stages(B.get(), s);
    if (e == hipErrorInvalidValue) {
        (void)hipFree(B->d_mem);
        return rf_set_error(ctx, RF_E_ARG, "batch too large for one grouped launch");
    }
    // draw 0, the plans' own realizations, folded at creation; next draw 1
    if (e == hipSuccess) e = rf_launch_draw_counter(B->d_draw, -2, 0, s);
    std::vector<rf_fold_item> items((size_t)n);
    if (B->redrawable) {
        for (int i = 0; i < n; i++) {
            rf_fold_item_init(&items[(size_t)i], B->folds[(size_t)i], B->nmc[(size_t)i], B->seeds[(size_t)i],
                              B->foldBlocks);
            B->foldBlocks += rf_fold_item_blocks(B->folds[(size_t)i]);
        }
        B->d_items = reinterpret_cast<rf_fold_item*>(B->d_mem + itemOff);
        if (e == hipSuccess)
            e = hipMemcpyAsync(B->d_items, items.data(), items.size() * sizeof(rf_fold_item), hipMemcpyHostToDevice, s);
    }
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

int rf_grid_batch_autotune(rf_grid_batch* B, void* stream, int32_t* chosen) {
    if (!B) return rf_set_error(nullptr, RF_E_ARG, "batch is NULL");
    rf_context* ctx = B->ctx;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (B->genActive) {
        // the current realizations' coefficients stored, as every tile kind
        // reads them, and the first contraction back to reading them
        RF_HIP_TRY(ctx, rf_launch_refold_draw(B->d_items, (int)B->folds.size(), B->foldBlocks, B->d_draw + 1, s));
        RF_HIP_TRY(ctx, rf_launch_draw_counter(B->d_draw, -3, 0, s));
        B->genActive = false;
    }
    int rc = rf_grid_batch_execute(B, stream);          // every stage's inputs in place
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RF_HIP_TRY(ctx, hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return rf_set_error(ctx, RF_E_HIP, hipGetErrorString(he));
    }
    auto time_stage = [&](int k, float* ms) -> hipError_t {
        hipError_t e = rf_launch_gemm_group(B->realOut[k], B->d_desc[k], B->nDesc[k], B->tiles[k], B->rwMax[k], s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int i = 0; i < 3 && e == hipSuccess; i++)
            e = rf_launch_gemm_group(B->realOut[k], B->d_desc[k], B->nDesc[k], B->tiles[k], B->rwMax[k], s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
        return e;
    };
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; k++) {
        bool wide = false;
        for (int r : B->kind0[k]) wide |= r > 1;
        if (!wide) continue;
        // candidates: the set-up's wide kinds, 128x64 tiles where it chose
        // the 64x160 whole-axis tile, and 64 x 64 tiles everywhere
        auto kind_of = [&](int opt, int i) {
            const int r = B->kind0[k][i];
            return opt == 0 ? r : opt == 1 ? (r == 5 ? 2 : r) : 1;
        };
        bool narrower = false;
        for (int r : B->kind0[k]) narrower |= r == 5;
        float best = 0.0f;
        int bestOpt = 0;
        for (int opt = 0; opt < 3 && e == hipSuccess; opt++) {
            if (opt == 1 && !narrower) continue;
            for (size_t i = 0; i < B->stage[k].size(); i++) B->stage[k][i].rw = kind_of(opt, (int)i);
            e = upload_stages(B, s);
            float ms = 0.0f;
            if (e == hipSuccess) e = time_stage(k, &ms);
            if (e == hipSuccess && (opt == 0 || ms < best)) {
                best = ms;
                bestOpt = opt;
            }
        }
        for (size_t i = 0; i < B->stage[k].size(); i++) B->stage[k][i].rw = kind_of(bestOpt, (int)i);
    }
    if (e == hipSuccess) e = upload_stages(B, s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("batch autotune: ") + hipGetErrorString(e));
    if (chosen)
        for (int k = 0; k < 3; k++) chosen[k] = B->rwMax[k];
    return RF_OK;
}

int rf_grid_batch_redraw(rf_grid_batch* B, int64_t drawIndex, void* stream) {
    if (!B) return rf_set_error(nullptr, RF_E_ARG, "batch is NULL");
    rf_context* ctx = B->ctx;
    if (!B->redrawable)
        return rf_set_error(ctx, RF_E_UNSUPPORTED, "rf_grid_batch_redraw: method 4 fields have no phases to redraw");
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    // when every field's first contraction runs on 64x160 tiles, those form
    // the coefficients of the current draw themselves (no fold, no
    // coefficient round trip through HBM); otherwise the grouped fold
    // stores them
    bool gen = !B->stage[0].empty();
    for (const auto& g : B->stage[0]) gen = gen && g.rw == 5 && g.genState;
    RF_HIP_TRY(ctx, rf_launch_draw_counter(B->d_draw, drawIndex >= 0 ? drawIndex : -1, gen ? 1 : 0, s));
    if (!gen)
        RF_HIP_TRY(ctx, rf_launch_refold_draw(B->d_items, (int)B->folds.size(), B->foldBlocks, B->d_draw + 1, s));
    B->genActive = gen;
    return RF_OK;
}

int rf_grid_batch_fuse_blend(rf_grid_batch* B, rf_blend_plan* P, void* d_group) {
    if (!B || !P) return rf_set_error(B ? B->ctx : nullptr, RF_E_ARG, "batch or blend plan is NULL");
    rf_context* ctx = B->ctx;
    if (P->ctx != ctx) return rf_set_error(ctx, RF_E_ARG, "blend plan of another context");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = ctx->stream;
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));          // no launch may read the tables being replaced
    std::vector<rf_gemm_desc>& last = B->stage[2];
    auto drop_boxes = [&]() {
        if (P->d_boxes) (void)hipFree(P->d_boxes);
        P->d_boxes = nullptr;
        P->nBoxes = 0;
        P->boxPoints = 0;
        if (P->d_subFused) (void)hipFree(P->d_subFused);
        P->d_subFused = nullptr;
        if (P->d_bands) (void)hipFree(P->d_bands);
        P->d_bands = nullptr;
        P->sep.sub = P->subPlain;
    };
    if (!d_group) {                                    // back to the separate blend
        for (auto& g : last) {
            g.grp = nullptr;
            g.bandLo = g.bandHi = nullptr;
        }
        P->sep.skipSingle = 0;
        P->fusedGroup = nullptr;
        drop_boxes();
        RF_HIP_TRY(ctx, upload_stages(B, s));
        return RF_OK;
    }
    if (!P->separable || P->mode != RF_FIXED)
        return rf_set_error(ctx, RF_E_UNSUPPORTED, "the fused blend needs a FIXED-mode tensor-product layout");
    if (P->dtype != B->outDtype) return rf_set_error(ctx, RF_E_ARG, "blend and batch element types differ");
    const int d = P->nDim;
    if (d < 2) return rf_set_error(ctx, RF_E_UNSUPPORTED, "the fused blend needs a 2-D or 3-D layout");
    int64_t gTot = 1;
    for (int a = 0; a < d; a++) gTot *= P->groupN[a];
    auto single = [&](int a, int gi) { return P->cover[a][(size_t)gi * 2 + 1] < 0; };
    // the local index range along axis a of subdomain box [o, o + n) that no
    // other subdomain covers (one run: the overlaps sit at the box's ends)
    auto run = [&](int a, int o, int n, int* lo, int* hi) -> bool {
        int l = 0, h = n;
        while (l < n && !single(a, o + l)) l++;
        while (h > l && !single(a, o + h - 1)) h--;
        for (int i = l; i < h; i++)
            if (!single(a, o + i)) return false;
        *lo = l;
        *hi = h;
        return true;
    };
    std::vector<rf_gemm_desc> fused = last;
    for (auto& g : fused) {
        if (!g.split16) return rf_set_error(ctx, RF_E_UNSUPPORTED, "fused blend: FP32-operand last stage");
        int s0 = -1;
        for (size_t k = 0; k < P->fields.size(); k++)
            if (P->fields[k] == g.D) { s0 = (int)k; break; }
        if (s0 < 0) return rf_set_error(ctx, RF_E_ARG, "a batch output is not a field of the blend plan");
        const int32_t* o = &P->origins[(size_t)s0 * d];
        const int32_t* n = &P->sizes[(size_t)s0 * d];
        const int64_t plane = d == 3 ? (int64_t)n[0] * n[1] : n[0];
        if (g.M != n[d - 1] || g.N != plane) return rf_set_error(ctx, RF_E_ARG, "field and blend box sizes differ");
        int xl, xh, yl = 0, yh = 1, zl, zh;
        bool ok = run(0, o[0], n[0], &xl, &xh) && run(d - 1, o[d - 1], n[d - 1], &zl, &zh);
        if (d == 3) ok = ok && run(1, o[1], n[1], &yl, &yh);
        if (!ok) return rf_set_error(ctx, RF_E_UNSUPPORTED, "fused blend: a subdomain's own points are not one box");
        const int64_t planeG = d == 3 ? (int64_t)P->groupN[0] * P->groupN[1] : P->groupN[0];
        g.grp = d_group;
        g.grpStrideB = gTot;
        g.grpNx = n[0];
        g.grpXlo = xl; g.grpXhi = xh; g.grpYlo = yl; g.grpYhi = yh; g.grpZlo = zl; g.grpZhi = zh;
        g.grpG0 = P->groupN[0];
        g.grpCol0 = o[0] + (d == 3 ? (int64_t)P->groupN[0] * o[1] : 0);
        g.grpRow0 = (int64_t)o[d - 1] * planeG;
        g.grpPlaneG = planeG;
        g.grpNy = d == 3 ? n[1] : 1;
        g.bandLo = g.bandHi = nullptr;
        g.bandLoStrideB = (int64_t)g.grpNy * n[d - 1] * xl;            // realization strides
        g.bandHiStrideB = (int64_t)g.grpNy * n[d - 1] * (n[0] - xh);
    }
    // the first-axis overlap bands of every field in arrays of their own,
    // [Nmc][rows][band width], 256-byte aligned
    const size_t esz = P->dtype == RF_DTYPE_F64 ? 8 : 4;
    std::vector<size_t> offLo(fused.size()), offHi(fused.size());
    size_t bandBytes = 0;
    auto take = [&](int64_t elems) {
        const size_t at = bandBytes;
        bandBytes += ((size_t)elems * esz + 255) & ~(size_t)255;
        return at;
    };
    for (size_t i = 0; i < fused.size(); i++) {
        offLo[i] = take(fused[i].bandLoStrideB * fused[i].batches);
        offHi[i] = take(fused[i].bandHiStrideB * fused[i].batches);
    }
    char* d_bands = nullptr;
    if (bandBytes && hipMalloc((void**)&d_bands, bandBytes) != hipSuccess)
        return rf_set_error(ctx, RF_E_NOMEM, "fused blend band arrays");
    for (size_t i = 0; i < fused.size(); i++) {
        if (fused[i].bandLoStrideB > 0) fused[i].bandLo = d_bands + offLo[i];
        if (fused[i].bandHiStrideB > 0) fused[i].bandHi = d_bands + offHi[i];
    }
    // the blend's subdomain table with the band arrays
    const int64_t nCells = (int64_t)P->sep.nInt[0] * P->sep.nInt[1] * P->sep.nInt[2];
    std::vector<rf_blend_sub> subs((size_t)nCells);
    rf_blend_sub* d_subFused = nullptr;
    hipError_t es = hipMemcpy(subs.data(), P->subPlain, subs.size() * sizeof(rf_blend_sub), hipMemcpyDeviceToHost);
    for (auto& S : subs) {
        for (const auto& g : fused)
            if (S.field && S.field == g.D) {
                S.bandLo = g.bandLo;
                S.bandHi = g.bandHi;
                S.bandLoTot = g.bandLoStrideB;
                S.bandHiTot = g.bandHiStrideB;
                S.wLo = g.grpXlo;
                S.wHi = g.grpNx - g.grpXhi;
            }
    }
    if (es == hipSuccess) es = hipMalloc((void**)&d_subFused, subs.size() * sizeof(rf_blend_sub));
    if (es == hipSuccess)
        es = hipMemcpy(d_subFused, subs.data(), subs.size() * sizeof(rf_blend_sub), hipMemcpyHostToDevice);
    if (es != hipSuccess) {
        if (d_subFused) (void)hipFree(d_subFused);
        if (d_bands) (void)hipFree(d_bands);
        return rf_set_error(ctx, RF_E_HIP, std::string("fused blend tables: ") + hipGetErrorString(es));
    }
    // the overlap bands: per axis, each run of group indices two intervals cover
    std::vector<rf_blend_box> boxes;
    int64_t first = 0;
    for (int a = 0; a < d; a++) {
        for (int gi = 0; gi < P->groupN[a];) {
            if (single(a, gi)) { gi++; continue; }
            int ge = gi;
            while (ge < P->groupN[a] && !single(a, ge)) ge++;
            rf_blend_box bx{};
            for (int c = 0; c < 3; c++) {
                bx.lo[c] = 0;
                bx.n[c] = c < d ? P->groupN[c] : 1;
            }
            bx.lo[a] = gi;
            bx.n[a] = ge - gi;
            bx.first = first;
            first += (int64_t)bx.n[0] * bx.n[1] * bx.n[2];
            boxes.push_back(bx);
            gi = ge;
        }
    }
    rf_blend_box* d_boxes = nullptr;
    auto free_new = [&]() {
        if (d_boxes) (void)hipFree(d_boxes);
        (void)hipFree(d_subFused);
        if (d_bands) (void)hipFree(d_bands);
    };
    if (!boxes.empty()) {
        if (hipMalloc((void**)&d_boxes, boxes.size() * sizeof(rf_blend_box)) != hipSuccess) {
            free_new();
            return rf_set_error(ctx, RF_E_NOMEM, "fused blend boxes");
        }
        const hipError_t e = hipMemcpy(d_boxes, boxes.data(), boxes.size() * sizeof(rf_blend_box),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            free_new();
            return rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
        }
    }
    const std::vector<rf_gemm_desc> before = last;
    last = fused;
    const hipError_t e = upload_stages(B, s);
    if (e != hipSuccess) {
        last = before;
        free_new();
        return rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    }
    drop_boxes();
    P->d_boxes = d_boxes;
    P->nBoxes = (int)boxes.size();
    P->boxPoints = first;
    P->d_subFused = d_subFused;
    P->d_bands = d_bands;
    P->sep.sub = d_subFused;
    P->sep.skipSingle = 1;
    P->fusedGroup = d_group;
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

// rf_server.cpp -- serving loop for field requests (rf_server_* in
// include/rf_capi.h): HBM admission, continuous batching, grouped launches.
//
// One worker thread owns the device side.  Each loop it
//   1. retires batches whose completion event has fired (frees their
//      transient memory, marks their requests done, wakes waiters),
//   2. if a stream slot is free, admits queued requests in arrival order
//      while their memory fits the budget (head-of-line: a request that does
//      not fit waits for memory, later ones wait behind it), closes the batch
//      at cfg.maxBatch requests or cfg.batchTerms terms, and launches it on
//      the slot's stream: copies, per-request coefficient preparation (the
//      same kernels rf_plan_create runs), then one grouped sum launch per
//      method,
//   3. otherwise sleeps on a condition variable (short timeout while batches
//      are in flight, so completions are seen within ~50 us).
// The mutex is never held across a blocking HIP call.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rf_arena.h"
#include "rf_internal.h"

namespace {

constexpr size_t kNone = ~(size_t)0;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device-resident lattice of one method, shared by every request.
struct Lattice {
    rf_spectrum_info info{};
    double* d_k = nullptr;    // method 1: shell radii [kN]; method 2: wavevectors / 2pi [kNpad][3]
    double* d_Sk = nullptr;   // method 1: spectral density [kN]
    double* d_amp = nullptr;  // method 2: sqrt(Sk) [kN] (one row for every realization)
    int64_t kNpad = 0;
    double bScale = 1.0;      // method 2: panel scale
};

struct Req {
    int64_t id = 0;
    rf_request r{};
    int status = 1;           // 1 queued / running, 0 done, < 0 failed
    double tSubmit = 0.0, tLaunch = 0.0, tDone = 0.0;
    int64_t batch = -1;
    size_t oOut = kNone, oX = kNone, oScr = kNone;
    int Rpad = 0;
    size_t bytesOut = 0, bytesX = 0, bytesScr = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    int64_t batch = -1;
    std::vector<int64_t> ids;
    rf_group_req* d_table = nullptr;   // [2 * maxBatch]: method 2 then method 1
    rf_group_req* h_table = nullptr;   // pinned staging of the same
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

struct rf_server {
    rf_context* ctx = nullptr;
    rf_server_config cfg{};
    Lattice lat[2];                    // [0] method 1, [1] method 2
    char* d_arena = nullptr;
    rf_arena arena;
    uint8_t* d_ones = nullptr;         // calculate flags, maxNmc ones
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Req> reqs;
    int64_t nextId = 1, nextBatch = 0;
    bool stop = false;
    uint64_t gen = 0;                  // bumped by submit, release and shutdown
    std::thread worker;
    rf_server_stats st{};
    std::string error;                 // first asynchronous failure

    Lattice& L(int method) { return lat[method == RF_SHINOZUKA ? 1 : 0]; }

    // device bytes of one request: output, point copy, scratch
    void sizes(Req& q) {
        const Lattice& l = L(q.r.method);
        const int64_t N = q.r.N, kN = l.info.kNTotal;
        const int Nmc = q.r.Nmc;
        q.bytesOut = align256((size_t)N * Nmc * sizeof(float));
        q.bytesX = q.r.xOnDevice ? 0 : align256((size_t)N * 3 * sizeof(double));
        if (q.r.method == RF_SHINOZUKA) {
            q.Rpad = (Nmc + 31) / 32 * 32;       // whole 32-row tiles (any count)
            q.bytesScr = 2 * align256((size_t)Nmc * kN * sizeof(double)) +   // uniforms, phases
                         align256(rf_mfma_bbytes(cfg.form, l.kNpad, q.Rpad));
        } else {
            q.Rpad = (Nmc + kGroupIsoR - 1) / kGroupIsoR * kGroupIsoR;
            q.bytesScr = align256((size_t)Nmc * 4 * kN * 8)        // uniforms
                         + align256((size_t)Nmc * kN * 3 * 8)      // directions
                         + 2 * align256((size_t)Nmc * kN * 8)      // amp, phase
                         + align256((size_t)kN * q.Rpad * 3 * 8)   // kT
                         + 2 * align256((size_t)kN * q.Rpad * 4);  // pT, aT
        }
    }

    double terms(const Req& q) { return (double)q.r.N * q.r.Nmc * (double)L(q.r.method).info.kNTotal; }

    bool alloc_req(Req& q) {
        size_t o = 0;
        if (!arena.alloc(q.bytesOut, &o)) return false;
        q.oOut = o;
        if (q.bytesX) {
            if (!arena.alloc(q.bytesX, &o)) {
                arena.release(q.oOut);
                q.oOut = kNone;
                return false;
            }
            q.oX = o;
        }
        if (!arena.alloc(q.bytesScr, &o)) {
            arena.release(q.oOut);
            q.oOut = kNone;
            if (q.oX != kNone) arena.release(q.oX);
            q.oX = kNone;
            return false;
        }
        q.oScr = o;
        return true;
    }

    void free_transient(Req& q) {
        if (q.oX != kNone) arena.release(q.oX);
        if (q.oScr != kNone) arena.release(q.oScr);
        q.oX = q.oScr = kNone;
    }

    int launch(Slot& sl, const std::vector<Req*>& batch, std::string& err);
    void run();
};

#define SRV_TRY(expr)                                                            \
    do {                                                                         \
        const hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess) {                                                  \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return RF_E_HIP;                                                     \
        }                                                                        \
    } while (0)

// Enqueues the batch's work on the slot's stream.  Called without the mutex;
// the requests it reads are not touched by other threads while admitted.
int rf_server::launch(Slot& sl, const std::vector<Req*>& batch, std::string& err) {
    const hipStream_t s = sl.stream;
    const int maxB = cfg.maxBatch;
    int nM = 0, nI = 0;
    int64_t tilesM = 0, tilesI = 0;
    for (Req* q : batch) {
        const rf_request& r = q->r;
        const Lattice& l = L(r.method);
        const int64_t kN = l.info.kNTotal;
        char* scr = d_arena + q->oScr;
        const double* d_x = r.xOnDevice ? r.x : reinterpret_cast<const double*>(d_arena + q->oX);
        if (!r.xOnDevice)
            SRV_TRY(hipMemcpyAsync(d_arena + q->oX, r.x, (size_t)r.N * 3 * sizeof(double),
                                   hipMemcpyHostToDevice, s));
        rf_group_req g{};
        g.a.x = d_x;
        g.a.N = r.N;
        g.a.i0 = 0;
        g.a.i1 = r.N;
        for (int d = 0; d < 3; d++) {
            g.a.corrL[d] = r.corrL[d];
            g.a.invL[d] = 1.0 / r.corrL[d];
        }
        g.a.nDim = 3;
        g.a.Nmc = r.Nmc;
        g.a.kN = kN;
        g.a.ampMult = l.info.ampMult;
        g.a.epi = rf_make_epilogue(r.margiFirst, r.fieldAvg, r.fieldVar);
        g.a.out = d_arena + q->oOut;
        g.a.outDtype = RF_DTYPE_F32;
        const int64_t nU = (int64_t)r.Nmc * kN;
        if (r.method == RF_SHINOZUKA) {
            // phases 2 pi U from the (seed, 2) stream, then the panels
            double* d_u = reinterpret_cast<double*>(scr);
            double* d_ph = reinterpret_cast<double*>(scr + align256((size_t)nU * sizeof(double)));
            char* d_B = scr + 2 * align256((size_t)nU * sizeof(double));
            SRV_TRY(rf_launch_uniforms(r.seed, 2u, nU, d_u, s));
            SRV_TRY(rf_launch_phase_from_uniform(d_u, nU, d_ph, s));
            SRV_TRY(rf_launch_mfma_prep(cfg.form, 3, kN, l.kNpad, r.Nmc, q->Rpad, l.d_k, d_ph, l.d_amp, 0,
                                        l.info.ampMult * l.bScale, d_B, s));
            const int mtTotal = q->Rpad / 32;
            g.ops0 = d_B;
            g.ops1 = d_B + (size_t)(l.kNpad / 8) * mtTotal * 128 * 16;
            g.nSteps = l.kNpad / 8;
            g.mtTotal = mtTotal;
            g.outScale = (float)(1.0 / l.bScale);
            g.tile0 = tilesM;
            tilesM += rf_group_tiles_mfma(r.N, mtTotal);
            sl.h_table[nM++] = g;
        } else {
            char* p = scr;
            double* d_u = reinterpret_cast<double*>(p);
            p += align256((size_t)r.Nmc * 4 * kN * 8);
            double* d_kIso = reinterpret_cast<double*>(p);
            p += align256((size_t)nU * 3 * 8);
            double* d_amp = reinterpret_cast<double*>(p);
            p += align256((size_t)nU * 8);
            double* d_phase = reinterpret_cast<double*>(p);
            p += align256((size_t)nU * 8);
            double* d_kT = reinterpret_cast<double*>(p);
            p += align256((size_t)kN * q->Rpad * 3 * 8);
            float* d_pT = reinterpret_cast<float*>(p);
            p += align256((size_t)kN * q->Rpad * 4);
            float* d_aT = reinterpret_cast<float*>(p);
            SRV_TRY(rf_launch_uniforms(r.seed, 2u, (int64_t)r.Nmc * 4 * kN, d_u, s));
            SRV_TRY(rf_launch_iso_prep(3, RF_FIXED, 0, kN, r.Nmc, l.d_k, l.d_Sk, d_u, d_ones, d_kIso, d_amp,
                                       d_phase, s));
            SRV_TRY(rf_launch_prep_fp32_iso(3, kN, r.Nmc, q->Rpad, d_kIso, d_phase, d_amp, l.info.ampMult, d_kT,
                                            d_pT, d_aT, s));
            g.ops0 = d_kT;
            g.ops1 = d_pT;
            g.ops2 = d_aT;
            g.Rpad = q->Rpad;
            g.tile0 = tilesI;
            tilesI += rf_group_tiles_iso(r.N, q->Rpad);
            sl.h_table[maxB + nI++] = g;
        }
    }
    SRV_TRY(hipMemcpyAsync(sl.d_table, sl.h_table, sizeof(rf_group_req) * 2 * maxB, hipMemcpyHostToDevice, s));
    SRV_TRY(rf_launch_group_mfma(cfg.form, sl.d_table, nM, tilesM, s));
    SRV_TRY(rf_launch_group_iso(sl.d_table + maxB, nI, tilesI, s));
    SRV_TRY(hipEventRecord(sl.done, s));
    return RF_OK;
}

void rf_server::run() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // 1. retire completed batches
        bool retired = false;
        int inflight = 0;
        for (Slot& sl : slots) {
            if (!sl.busy) continue;
            const hipError_t e = hipEventQuery(sl.done);
            if (e == hipErrorNotReady) {
                inflight++;
                continue;
            }
            const double t = now_s();
            for (int64_t id : sl.ids) {
                auto it = reqs.find(id);
                if (it == reqs.end()) continue;
                Req& q = it->second;
                free_transient(q);
                q.tDone = t;
                if (e == hipSuccess) {
                    q.status = 0;
                    st.completed++;
                    st.pointRealizations += (double)q.r.N * q.r.Nmc;
                } else {
                    q.status = RF_E_HIP;
                    st.failed++;
                    if (error.empty()) error = std::string("batch: ") + hipGetErrorString(e);
                }
            }
            sl.ids.clear();
            sl.busy = false;
            retired = true;
        }
        if (retired) cvDone.notify_all();
        if (stop && inflight == 0) break;

        // 2. admit and launch into a free slot
        Slot* slot = nullptr;
        if (!stop)
            for (Slot& sl : slots)
                if (!sl.busy) {
                    slot = &sl;
                    break;
                }
        if (slot && !pending.empty()) {
            std::vector<Req*> batch;
            double work = 0.0;
            while (!pending.empty() && (int)batch.size() < cfg.maxBatch &&
                   (batch.empty() || work < cfg.batchTerms)) {
                auto it = reqs.find(pending.front());
                if (it == reqs.end()) {
                    pending.pop_front();
                    continue;
                }
                Req& q = it->second;
                if (!alloc_req(q)) break;   // waits for memory (head of line)
                pending.pop_front();
                work += terms(q);
                batch.push_back(&q);
            }
            if (!batch.empty()) {
                const int64_t b = nextBatch++;
                slot->busy = true;
                slot->batch = b;
                slot->ids.clear();
                const double t = now_s();
                for (Req* q : batch) {
                    q->batch = b;
                    q->tLaunch = t;
                    slot->ids.push_back(q->id);
                }
                st.batches++;
                st.maxBatchSeen = std::max(st.maxBatchSeen, (int32_t)batch.size());
                st.maxInflightSeen = std::max(st.maxInflightSeen, (int32_t)(inflight + 1));
                st.arenaUsed = (int64_t)arena.used();
                st.arenaPeak = (int64_t)arena.peak();
                std::string err;
                lk.unlock();
                const int rc = launch(*slot, batch, err);
                if (rc != RF_OK) (void)hipStreamSynchronize(slot->stream);   // drain what was enqueued
                lk.lock();
                if (rc != RF_OK) {
                    if (error.empty()) error = err;
                    for (Req* q : batch) {
                        free_transient(*q);
                        if (q->oOut != kNone) arena.release(q->oOut);
                        q->oOut = kNone;
                        q->status = rc;
                        q->tDone = now_s();
                        st.failed++;
                    }
                    slot->ids.clear();
                    slot->busy = false;
                    cvDone.notify_all();
                }
                continue;
            }
        }
        // 3. sleep: poll while batches are in flight, else until a submit,
        // release or shutdown changes the state
        if (inflight > 0) {
            cvWork.wait_for(lk, std::chrono::microseconds(50));
        } else if (!stop) {
            const uint64_t g = gen;
            cvWork.wait(lk, [&] { return stop || gen != g; });
        }
    }
}

namespace {

int build_lattice(rf_server* S, int method, Lattice& l) {
    rf_context* ctx = S->ctx;
    rf_params p{};
    p.nDim = 3;
    p.method = method;
    p.corrMod = RF_CM_GAUSSIAN;
    p.mode = RF_FIXED;
    p.form = method == RF_SHINOZUKA ? S->cfg.form : RF_FORM_FP32;
    p.kSampling = RF_KSAMP_BOX;
    p.Nmc = 1;
    for (int d = 0; d < 3; d++) {
        p.corrL[d] = 1.0;
        p.xRange[d] = S->cfg.xRangeNorm[d];
    }
    p.fieldVar = 1.0;
    std::string err;
    int rc = rf_spectrum_build_info(&p, &l.info, &err);
    if (rc) return rf_set_error(ctx, rc, err);
    rf_spectrum spec;
    rc = rf_spectrum_build_tables(&p, l.info, nullptr, &spec, &err);
    if (rc) return rf_set_error(ctx, rc, err);
    const int64_t kN = l.info.kNTotal;
    const hipStream_t s = ctx->stream;
    if (method == RF_SHINOZUKA) {
        l.kNpad = (kN + rf_mfma_kpad(p.form) - 1) / rf_mfma_kpad(p.form) * rf_mfma_kpad(p.form);
        std::vector<double> kRev((size_t)l.kNpad * 3, 0.0), amp((size_t)kN);
        double maxAmp = 0.0;
        for (int64_t j = 0; j < kN; j++) {
            for (int d = 0; d < 3; d++) kRev[(size_t)j * 3 + d] = spec.k[(size_t)j * 3 + d] * RF_INV_TWO_PI;
            amp[(size_t)j] = std::sqrt(spec.Sk[(size_t)j]);
            maxAmp = std::max(maxAmp, amp[(size_t)j]);
        }
        l.bScale = rf_mfma_bscale(l.info.ampMult, maxAmp);
        RF_HIP_TRY(ctx, hipMalloc((void**)&l.d_k, kRev.size() * sizeof(double)));
        RF_HIP_TRY(ctx, hipMalloc((void**)&l.d_amp, amp.size() * sizeof(double)));
        RF_HIP_TRY(ctx, hipMemcpyAsync(l.d_k, kRev.data(), kRev.size() * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, hipMemcpyAsync(l.d_amp, amp.data(), amp.size() * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        RF_HIP_TRY(ctx, hipMalloc((void**)&l.d_k, (size_t)kN * sizeof(double)));
        RF_HIP_TRY(ctx, hipMalloc((void**)&l.d_Sk, (size_t)kN * sizeof(double)));
        RF_HIP_TRY(ctx, hipMemcpyAsync(l.d_k, spec.k.data(), (size_t)kN * sizeof(double), hipMemcpyHostToDevice, s));
        RF_HIP_TRY(ctx, hipMemcpyAsync(l.d_Sk, spec.Sk.data(), (size_t)kN * sizeof(double), hipMemcpyHostToDevice, s));
    }
    RF_HIP_TRY(ctx, hipStreamSynchronize(s));   // host tables leave scope
    return RF_OK;
}

void free_server(rf_server* S) {
    for (Slot& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.d_table) (void)hipFree(sl.d_table);
        if (sl.h_table) (void)hipHostFree(sl.h_table);
    }
    for (Lattice& l : S->lat) {
        if (l.d_k) (void)hipFree(l.d_k);
        if (l.d_Sk) (void)hipFree(l.d_Sk);
        if (l.d_amp) (void)hipFree(l.d_amp);
    }
    if (S->d_ones) (void)hipFree(S->d_ones);
    if (S->d_arena) (void)hipFree(S->d_arena);
    delete S;
}

}  // namespace

extern "C" {

int rf_server_create(rf_context* ctx, const rf_server_config* cfg, rf_server** out) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!out) return rf_set_error(ctx, RF_E_ARG, "server out-pointer is NULL");
    *out = nullptr;
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    rf_server* S = new (std::nothrow) rf_server();
    if (!S) return rf_set_error(ctx, RF_E_NOMEM, "server allocation failed");
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    rf_server_config& c = S->cfg;
    if (c.maxInflight <= 0) c.maxInflight = 3;
    if (c.maxBatch <= 0) c.maxBatch = 64;
    if (c.form == 0) c.form = RF_FORM_MFMA_F16;
    if (c.maxNmc <= 0) c.maxNmc = 1024;
    if (!(c.batchTerms > 0.0)) c.batchTerms = 2e11;
    for (int d = 0; d < 3; d++)
        if (!(c.xRangeNorm[d] > 0.0)) c.xRangeNorm[d] = 10.0;
    if (c.maxInflight > 8 || (c.form != RF_FORM_MFMA_F16 && c.form != RF_FORM_MFMA_BF16)) {
        delete S;
        return rf_set_error(ctx, RF_E_ARG, "bad server config (maxInflight 1..8, form MFMA F16/BF16)");
    }
    int rc = build_lattice(S, RF_ISOTROPIC, S->lat[0]);
    if (!rc) rc = build_lattice(S, RF_SHINOZUKA, S->lat[1]);
    if (rc) {
        free_server(S);
        return rc;
    }
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
        free_server(S);
        return rf_set_error(ctx, RF_E_HIP, "hipMemGetInfo failed");
    }
    if (c.hbmBudget <= 0) c.hbmBudget = (int64_t)(freeB / 2);
    if ((size_t)c.hbmBudget > freeB) {
        free_server(S);
        return rf_set_error(ctx, RF_E_NOMEM, "server budget exceeds the free device memory");
    }
    if (hipMalloc((void**)&S->d_arena, (size_t)c.hbmBudget) != hipSuccess) {
        free_server(S);
        return rf_set_error(ctx, RF_E_NOMEM, "server arena allocation failed");
    }
    S->arena.init((size_t)c.hbmBudget, 256);
    S->st.arenaCapacity = (int64_t)S->arena.capacity();
    std::vector<uint8_t> ones((size_t)c.maxNmc, 1);
    if (hipMalloc((void**)&S->d_ones, ones.size()) != hipSuccess ||
        hipMemcpy(S->d_ones, ones.data(), ones.size(), hipMemcpyHostToDevice) != hipSuccess) {
        free_server(S);
        return rf_set_error(ctx, RF_E_HIP, "server flag table");
    }
    S->slots.resize((size_t)c.maxInflight);
    for (Slot& sl : S->slots) {
        const size_t tb = sizeof(rf_group_req) * 2 * (size_t)c.maxBatch;
        if (hipStreamCreate(&sl.stream) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void**)&sl.d_table, tb) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_table, tb, hipHostMallocDefault) != hipSuccess) {
            free_server(S);
            return rf_set_error(ctx, RF_E_HIP, "server stream slot set-up failed");
        }
        std::memset(sl.h_table, 0, tb);
    }
    S->worker = std::thread([S] { S->run(); });
    *out = S;
    return RF_OK;
}

int rf_server_destroy(rf_server* S) {
    if (!S) return RF_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->gen++;
        for (int64_t id : S->pending) {
            auto it = S->reqs.find(id);
            if (it == S->reqs.end()) continue;
            it->second.status = RF_E_ARG;
            it->second.tDone = now_s();
        }
        S->pending.clear();
    }
    S->cvWork.notify_all();
    S->cvDone.notify_all();
    if (S->worker.joinable()) S->worker.join();
    free_server(S);
    return RF_OK;
}

int rf_server_submit(rf_server* S, const rf_request* r, int64_t* id) {
    if (!S || !r || !id) return rf_set_error(nullptr, RF_E_ARG, "NULL argument");
    if (r->method != RF_SHINOZUKA && r->method != RF_ISOTROPIC)
        return rf_set_error(S->ctx, RF_E_UNSUPPORTED, "server methods: 1 (isotropic) and 2 (Shinozuka)");
    if (r->N < 1 || !r->x) return rf_set_error(S->ctx, RF_E_ARG, "request needs N >= 1 points");
    if (r->Nmc < 1 || r->Nmc > S->cfg.maxNmc) return rf_set_error(S->ctx, RF_E_ARG, "request Nmc out of range");
    for (int d = 0; d < 3; d++)
        if (!(r->corrL[d] > 0.0)) return rf_set_error(S->ctx, RF_E_ARG, "corrL must be > 0");
    if (r->margiFirst < RF_FOM_NONE || r->margiFirst > RF_FOM_LOGNORMAL)
        return rf_set_error(S->ctx, RF_E_ARG, "unknown marginal");
    if (r->margiFirst == RF_FOM_LOGNORMAL && !(r->fieldAvg > 0.0))
        return rf_set_error(S->ctx, RF_E_LOGNORMAL, "lognormal marginal needs fieldAvg > 0");
    Req q;
    q.r = *r;
    S->sizes(q);
    if (q.bytesOut + q.bytesX + q.bytesScr > S->arena.capacity())
        return rf_set_error(S->ctx, RF_E_NOMEM, "request exceeds the server's memory budget");
    if (rf_group_tiles_mfma(r->N, q.Rpad / 32) > 0x7fffffffLL || rf_group_tiles_iso(r->N, q.Rpad) > 0x7fffffffLL)
        return rf_set_error(S->ctx, RF_E_ARG, "request too large for one grouped launch");
    q.tSubmit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return rf_set_error(S->ctx, RF_E_ARG, "server is shutting down");
        const int64_t i = S->nextId++;
        q.id = i;
        S->reqs.emplace(i, q);
        S->pending.push_back(i);
        S->st.submitted++;
        S->gen++;
        *id = i;
    }
    S->cvWork.notify_one();
    return RF_OK;
}

int rf_server_wait(rf_server* S, int64_t id, double timeoutSec, rf_request_result* res) {
    if (!S || !res) return rf_set_error(nullptr, RF_E_ARG, "NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return rf_set_error(S->ctx, RF_E_ARG, "unknown request id");
    auto ready = [&] {
        auto j = S->reqs.find(id);
        return j == S->reqs.end() || j->second.status != 1;
    };
    if (timeoutSec < 0.0) S->cvDone.wait(lk, ready);
    else S->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSec), ready);
    it = S->reqs.find(id);
    if (it == S->reqs.end()) return rf_set_error(S->ctx, RF_E_ARG, "request released while waiting");
    const Req& q = it->second;
    res->status = q.status;
    res->Nmc = q.r.Nmc;
    res->N = q.r.N;
    res->d_out = (q.status == 0 && q.oOut != kNone) ? reinterpret_cast<const float*>(S->d_arena + q.oOut) : nullptr;
    res->tSubmit = q.tSubmit;
    res->tLaunch = q.tLaunch;
    res->tDone = q.tDone;
    res->batch = q.batch;
    return RF_OK;
}

int rf_server_read(rf_server* S, int64_t id, float* host) {
    if (!S || !host) return rf_set_error(nullptr, RF_E_ARG, "NULL argument");
    const float* src = nullptr;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end() || it->second.status != 0)
            return rf_set_error(S->ctx, RF_E_ARG, "request is unknown or not completed");
        src = reinterpret_cast<const float*>(S->d_arena + it->second.oOut);
        bytes = (size_t)it->second.r.N * it->second.r.Nmc * sizeof(float);
    }
    // the output stays allocated until rf_server_release (caller's order)
    RF_HIP_TRY(S->ctx, hipSetDevice(S->ctx->device));
    RF_HIP_TRY(S->ctx, hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
    return RF_OK;
}

int rf_server_release(rf_server* S, int64_t id) {
    if (!S) return rf_set_error(nullptr, RF_E_ARG, "server is NULL");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return rf_set_error(S->ctx, RF_E_ARG, "unknown request id");
        if (it->second.status == 1) return rf_set_error(S->ctx, RF_E_ARG, "request still queued or running");
        if (it->second.oOut != kNone) S->arena.release(it->second.oOut);
        S->reqs.erase(it);
        S->st.arenaUsed = (int64_t)S->arena.used();
        S->gen++;
    }
    S->cvWork.notify_one();   // memory freed: queued requests may fit now
    return RF_OK;
}

int rf_server_get_stats(rf_server* S, rf_server_stats* out) {
    if (!S || !out) return rf_set_error(nullptr, RF_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    S->st.arenaUsed = (int64_t)S->arena.used();
    S->st.arenaPeak = (int64_t)S->arena.peak();
    *out = S->st;
    return RF_OK;
}

}  // extern "C"

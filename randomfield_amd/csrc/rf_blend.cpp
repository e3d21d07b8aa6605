// rf_blend.cpp -- host side of rf_blend_device: analyses the subdomain boxes
// (a tensor-product arrangement, as prepare_Localization's setGrid of
// subdivision origins produces, localization_RF.f90:151-155), splits each
// axis into segments of constant cover, lists the covering subdomains of
// every segment product (cell) in ascending order, tabulates the separable
// partition-of-unity factors, and launches the gather-form blend kernel.
//
// Partition of unity (generateUnityPartition_Matrix, localization_RF.f90:705-857):
// the weight of a subdomain point is a product over tapered faces of a
// per-axis factor, so it is stored per (subdomain, axis, position).
//   REF_COMPAT  lower face position s = 1..n: v = (n+1-s)/(n+2); upper face
//               (index N-n+s): v = s/(n+2); factor (1 + cos(pi v))/2 (:800-813),
//               applied linearly (:228).  Afterwards realization 1 of the
//               group is divided by the same taper on every group face
//               (calls_RF.f90:338-355; one rank has no MPI neighbours).
//   FIXED       v = (n+1-s)/(n+1) resp. s/(n+1), factor sqrt((1+cos(pi v))/2):
//               two faces meeting in an n-point overlap sum their squared
//               factors to exactly one, so independent unit-variance
//               subdomains blend to unit variance.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <utility>

#include "rf_blend.h"
#include "rf_internal.h"

namespace {

double face_factor(int s, int n, bool upper, int mode) {
    const int q = upper ? s : (n + 1 - s);
    if (mode == RF_REF_COMPAT) {
        const double v = std::fabs((double)q / (double)(n + 2));
        return (1.0 + std::cos(RF_PI_REF * v)) / 2.0;
    }
    const double v = (double)q / (double)(n + 1);
    return std::sqrt((1.0 + std::cos(RF_PI_REF * v)) / 2.0);
}

// factors of an N-point axis: lower face at positions 0..n-1, upper face at
// N-n..N-1 (1.0 where untapered)
void axis_factors(int N, int n, bool lower, bool upper, int mode, double* low, double* up) {
    for (int p = 0; p < N; p++) { low[p] = 1.0; up[p] = 1.0; }
    if (n < 1) return;
    if (lower)
        for (int s = 1; s <= n; s++) low[s - 1] = face_factor(s, n, false, mode);
    if (upper)
        for (int s = 1; s <= n; s++) up[N - n + s - 1] = face_factor(s, n, true, mode);
}

}  // namespace

extern "C" int rf_blend_plan_create(rf_context* ctx, int32_t nDim, const int32_t* groupN, int32_t Nmc,
                                    int32_t nSub, const int32_t* origins, const int32_t* sizes,
                                    const int32_t* nOvl, const int32_t* faceMasks, int32_t mode,
                                    const void* const* d_fields, int32_t dtype, rf_blend_plan** out) {
    if (out) *out = nullptr;
    if (!out) return rf_set_error(ctx, RF_E_ARG, "plan out-pointer is NULL");
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (nDim < 1 || nDim > 3 || Nmc < 1 || nSub < 1) return rf_set_error(ctx, RF_E_ARG, "bad blend sizes");
    if (!groupN || !origins || !sizes || !nOvl || !faceMasks || !d_fields)
        return rf_set_error(ctx, RF_E_ARG, "NULL blend argument");
    if (mode != RF_REF_COMPAT && mode != RF_FIXED) return rf_set_error(ctx, RF_E_ARG, "unknown mode");
    if (dtype != RF_DTYPE_F64 && dtype != RF_DTYPE_F32) return rf_set_error(ctx, RF_E_ARG, "unknown dtype");
    int64_t gTot = 1;
    for (int a = 0; a < nDim; a++) {
        if (groupN[a] < 1) return rf_set_error(ctx, RF_E_ARG, "group size must be >= 1");
        if (nOvl[a] < 0) return rf_set_error(ctx, RF_E_ARG, "overlap points must be >= 0");
        gTot *= groupN[a];
    }
    // distinct intervals per axis
    std::vector<std::pair<int, int>> ints[3];
    std::vector<int> subInt((size_t)nSub * 3, 0);
    int maxSize = 1;
    for (int s = 0; s < nSub; s++) {
        for (int a = 0; a < nDim; a++) {
            const int o = origins[s * nDim + a], z = sizes[s * nDim + a];
            if (z < 1 || o < 0 || o + z > groupN[a])
                return rf_set_error(ctx, RF_E_ARG, "subdomain box outside the group grid");
            for (int f = 0; f < 2; f++)
                if ((faceMasks[s] >> (2 * a + f)) & 1)
                    if (nOvl[a] > z)
                        return rf_set_error(ctx, RF_E_ARG, "subdomain thinner than its tapered overlap");
            maxSize = std::max(maxSize, z);
            ints[a].push_back({o, z});
        }
    }
    int nInt[3] = {1, 1, 1};
    for (int a = 0; a < nDim; a++) {
        std::sort(ints[a].begin(), ints[a].end());
        ints[a].erase(std::unique(ints[a].begin(), ints[a].end()), ints[a].end());
        nInt[a] = (int)ints[a].size();
    }
    for (int s = 0; s < nSub; s++)
        for (int a = 0; a < nDim; a++) {
            const std::pair<int, int> key{origins[s * nDim + a], sizes[s * nDim + a]};
            subInt[(size_t)s * 3 + a] = (int)(std::lower_bound(ints[a].begin(), ints[a].end(), key) - ints[a].begin());
        }
    const int64_t cells = (int64_t)nInt[0] * nInt[1] * nInt[2];
    std::vector<int32_t> subOf((size_t)cells, -1);
    for (int s = 0; s < nSub; s++) {
        const int64_t c = subInt[(size_t)s * 3] + (int64_t)nInt[0] * (subInt[(size_t)s * 3 + 1] + (int64_t)nInt[1] * subInt[(size_t)s * 3 + 2]);
        if (subOf[(size_t)c] >= 0) return rf_set_error(ctx, RF_E_UNSUPPORTED, "two subdomains share the same box");
        subOf[(size_t)c] = s;
    }
    // cover tables: <= 2 intervals per group index, ascending
    std::vector<int32_t> cover[3];
    for (int a = 0; a < nDim; a++) {
        cover[a].assign((size_t)groupN[a] * 2, -1);
        for (int iv = 0; iv < nInt[a]; iv++) {
            for (int g = ints[a][iv].first; g < ints[a][iv].first + ints[a][iv].second; g++) {
                int32_t* c = &cover[a][(size_t)g * 2];
                if (c[0] < 0) c[0] = iv;
                else if (c[1] < 0) c[1] = iv;
                else return rf_set_error(ctx, RF_E_UNSUPPORTED, "more than two subdomains overlap along one axis");
            }
        }
    }
    // segments: runs of group indices with the same covering intervals
    std::vector<int32_t> segOf[3];
    std::vector<std::pair<int32_t, int32_t>> segCover[3];
    int nSeg[3] = {1, 1, 1};
    for (int a = 0; a < 3; a++) {
        if (a >= nDim) {
            segOf[a].assign(1, 0);
            segCover[a].push_back({0, -1});
            continue;
        }
        segOf[a].resize((size_t)groupN[a]);
        for (int g = 0; g < groupN[a]; g++) {
            const std::pair<int32_t, int32_t> key{cover[a][(size_t)g * 2], cover[a][(size_t)g * 2 + 1]};
            if (segCover[a].empty() || segCover[a].back() != key) segCover[a].push_back(key);
            segOf[a][(size_t)g] = (int32_t)segCover[a].size() - 1;
        }
        nSeg[a] = (int)segCover[a].size();
    }
    // cells: the covering subdomains of each segment product, ascending
    const int64_t nCell = (int64_t)nSeg[0] * nSeg[1] * nSeg[2];
    std::vector<int32_t> cellOff;
    std::vector<rf_blend_entry> entries;
    cellOff.reserve((size_t)nCell + 1);
    for (int64_t c = 0; c < nCell; c++) {
        const int c0 = (int)(c % nSeg[0]), c1 = (int)((c / nSeg[0]) % nSeg[1]), c2 = (int)(c / ((int64_t)nSeg[0] * nSeg[1]));
        cellOff.push_back((int32_t)entries.size());
        const int32_t iv[3][2] = {{segCover[0][c0].first, segCover[0][c0].second},
                                  {segCover[1][c1].first, segCover[1][c1].second},
                                  {segCover[2][c2].first, segCover[2][c2].second}};
        std::vector<int32_t> subs;
        for (int k2 = 0; k2 < 2; k2++)
            for (int k1 = 0; k1 < 2; k1++)
                for (int k0 = 0; k0 < 2; k0++) {
                    if (iv[0][k0] < 0 || iv[1][k1] < 0 || iv[2][k2] < 0) continue;
                    const int32_t sb = subOf[(size_t)(iv[0][k0] + (int64_t)nInt[0] * (iv[1][k1] + (int64_t)nInt[1] * iv[2][k2]))];
                    if (sb >= 0) subs.push_back(sb);
                }
        std::sort(subs.begin(), subs.end());
        if ((int)subs.size() > kBlendMaxEntries)
            return rf_set_error(ctx, RF_E_UNSUPPORTED, "more than 2^d subdomains cover a group point");
        for (int32_t sb : subs) {
            rf_blend_entry e{};
            int64_t n[3] = {1, 1, 1};
            for (int ax = 0; ax < nDim; ax++) {
                n[ax] = sizes[sb * nDim + ax];
                e.o[ax] = origins[sb * nDim + ax];
            }
            e.field = d_fields[sb];
            e.subTot = n[0] * n[1] * n[2];
            e.stride1 = n[0];
            e.stride2 = n[0] * n[1];
            e.base = -((int64_t)e.o[0] + (int64_t)e.o[1] * e.stride1 + (int64_t)e.o[2] * e.stride2);
            e.s = sb;
            entries.push_back(e);
        }
    }
    cellOff.push_back((int32_t)entries.size());
    // taper factor tables
    std::vector<double> lowF((size_t)nSub * 3 * maxSize, 1.0), upF((size_t)nSub * 3 * maxSize, 1.0);
    for (int s = 0; s < nSub; s++)
        for (int a = 0; a < nDim; a++)
            axis_factors(sizes[s * nDim + a], nOvl[a], (faceMasks[s] >> (2 * a)) & 1,
                         (faceMasks[s] >> (2 * a + 1)) & 1, mode,
                         &lowF[((size_t)s * 3 + a) * maxSize], &upF[((size_t)s * 3 + a) * maxSize]);
    std::vector<double> pu1[3][2];
    for (int a = 0; a < nDim; a++) {
        pu1[a][0].assign((size_t)groupN[a], 1.0);
        pu1[a][1].assign((size_t)groupN[a], 1.0);
        if (mode == RF_REF_COMPAT)
            axis_factors(groupN[a], std::min(nOvl[a], groupN[a]), true, true, mode, pu1[a][0].data(), pu1[a][1].data());
    }

    // separable arrangement: the tapered faces of every subdomain along axis
    // a are a function of its interval along a
    bool separable = true;
    std::vector<int> ivMask[3];
    for (int a = 0; a < nDim; a++) ivMask[a].assign((size_t)nInt[a], -1);
    for (int s = 0; s < nSub && separable; s++)
        for (int a = 0; a < nDim; a++) {
            const int m = (faceMasks[s] >> (2 * a)) & 3;
            int& slot = ivMask[a][(size_t)subInt[(size_t)s * 3 + a]];
            if (slot < 0) slot = m;
            else if (slot != m) separable = false;
        }
    std::vector<rf_blend_axis> axT[3];
    std::vector<rf_blend_sub> subT;
    if (separable) {
        for (int a = 0; a < 3; a++) {
            if (a >= nDim) {
                rf_blend_axis u{};
                u.iv[0] = 0; u.iv[1] = -1; u.sz[0] = u.sz[1] = 1; u.f[0] = u.f[1] = 1.0; u.pu1 = 1.0;
                axT[a].push_back(u);
                continue;
            }
            axT[a].resize((size_t)groupN[a]);
            for (int g = 0; g < groupN[a]; g++) {
                rf_blend_axis& e = axT[a][(size_t)g];
                e = rf_blend_axis{};
                e.pu1 = pu1[a][0][(size_t)g] * pu1[a][1][(size_t)g];
                for (int c = 0; c < 2; c++) {
                    const int iv = cover[a][(size_t)g * 2 + c];
                    e.iv[c] = iv;
                    e.sz[c] = 1;
                    e.f[c] = 0.0;
                    if (iv < 0) continue;
                    const int o = ints[a][(size_t)iv].first, z = ints[a][(size_t)iv].second;
                    const int m = ivMask[a][(size_t)iv] < 0 ? 0 : ivMask[a][(size_t)iv];
                    std::vector<double> lo((size_t)z), up((size_t)z);
                    axis_factors(z, nOvl[a], m & 1, (m >> 1) & 1, mode, lo.data(), up.data());
                    e.pos[c] = g - o;
                    e.sz[c] = z;
                    e.f[c] = lo[(size_t)(g - o)] * up[(size_t)(g - o)];
                }
            }
        }
        subT.assign((size_t)cells, rf_blend_sub{nullptr, 0});
        for (int64_t c = 0; c < cells; c++) {
            const int sb = subOf[(size_t)c];
            if (sb < 0) continue;
            int64_t t = 1;
            for (int a = 0; a < nDim; a++) t *= sizes[sb * nDim + a];
            subT[(size_t)c] = rf_blend_sub{d_fields[sb], t};
        }
    }

    // one device block for every table
    std::vector<char> blob;
    auto put = [&](const void* src, size_t bytes) {
        const size_t off = (blob.size() + 15) / 16 * 16;
        blob.resize(off + bytes);
        if (bytes) std::memcpy(blob.data() + off, src, bytes);
        return off;
    };
    size_t oSeg[3] = {0, 0, 0}, oP1[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int a = 0; a < 3; a++) oSeg[a] = put(segOf[a].data(), segOf[a].size() * 4);
    for (int a = 0; a < nDim; a++) {
        oP1[a][0] = put(pu1[a][0].data(), pu1[a][0].size() * 8);
        oP1[a][1] = put(pu1[a][1].data(), pu1[a][1].size() * 8);
    }
    const size_t oCell = put(cellOff.data(), cellOff.size() * 4);
    const size_t oEnt = put(entries.data(), entries.size() * sizeof(rf_blend_entry));
    const size_t oLow = put(lowF.data(), lowF.size() * 8);
    const size_t oUp = put(upF.data(), upF.size() * 8);
    size_t oAx[3] = {0, 0, 0}, oSubT = 0;
    if (separable) {
        for (int a = 0; a < 3; a++) oAx[a] = put(axT[a].data(), axT[a].size() * sizeof(rf_blend_axis));
        oSubT = put(subT.data(), subT.size() * sizeof(rf_blend_sub));
    }

    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<rf_blend_plan> P(new (std::nothrow) rf_blend_plan());
    if (!P) return rf_set_error(ctx, RF_E_NOMEM, "blend plan allocation failed");
    P->ctx = ctx;
    P->dtype = dtype;
    P->nDim = nDim;
    P->mode = mode;
    for (int a = 0; a < nDim; a++) {
        P->groupN[a] = groupN[a];
        P->cover[a] = cover[a];
    }
    P->origins.assign(origins, origins + (size_t)nSub * nDim);
    P->sizes.assign(sizes, sizes + (size_t)nSub * nDim);
    P->fields.assign(d_fields, d_fields + nSub);
    RF_HIP_TRY(ctx, hipMalloc((void**)&P->d_blob, blob.size()));
    const hipError_t e = hipMemcpy(P->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(P->d_blob);
        return rf_set_error(ctx, RF_E_HIP, std::string("blend tables: ") + hipGetErrorString(e));
    }
    rf_blend_dev& b = P->b;
    const char* d_blob = P->d_blob;
    b.nDim = nDim; b.Nmc = Nmc; b.mode = mode;
    for (int a = 0; a < 3; a++) {
        b.groupN[a] = a < nDim ? groupN[a] : 1;
        b.nSeg[a] = nSeg[a];
        b.seg[a] = reinterpret_cast<const int32_t*>(d_blob + oSeg[a]);
        if (a < nDim) {
            b.pu1Low[a] = reinterpret_cast<const double*>(d_blob + oP1[a][0]);
            b.pu1Up[a] = reinterpret_cast<const double*>(d_blob + oP1[a][1]);
        }
    }
    b.groupTot = gTot;
    b.cellOff = reinterpret_cast<const int32_t*>(d_blob + oCell);
    b.entries = reinterpret_cast<const rf_blend_entry*>(d_blob + oEnt);
    b.lowF = reinterpret_cast<const double*>(d_blob + oLow);
    b.upF = reinterpret_cast<const double*>(d_blob + oUp);
    b.maxSize = maxSize;
    if (separable) {
        P->separable = true;
        rf_blend_sep& q = P->sep;
        q.nDim = nDim; q.Nmc = Nmc; q.mode = mode;
        for (int a = 0; a < 3; a++) {
            q.groupN[a] = b.groupN[a];
            q.nInt[a] = nInt[a];
            q.ax[a] = reinterpret_cast<const rf_blend_axis*>(d_blob + oAx[a]);
        }
        q.groupTot = gTot;
        q.sub = reinterpret_cast<const rf_blend_sub*>(d_blob + oSubT);
        P->subPlain = q.sub;
    }
    *out = P.release();
    return RF_OK;
}

extern "C" int rf_blend_plan_execute(rf_blend_plan* P, void* d_group, void* stream) {
    if (!P) return rf_set_error(nullptr, RF_E_ARG, "blend plan is NULL");
    if (!d_group) return rf_set_error(P->ctx, RF_E_ARG, "NULL group buffer");
    rf_context* ctx = P->ctx;
    int dev = -1;
    RF_HIP_TRY(ctx, hipGetDevice(&dev));
    if (dev != ctx->device) RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (P->fusedGroup && d_group != P->fusedGroup)
        return rf_set_error(ctx, RF_E_ARG, "this blend is fused into a grid batch writing another group buffer");
    if (P->fusedGroup) RF_HIP_TRY(ctx, rf_launch_blend_boxes(P->sep, P->d_boxes, P->nBoxes, P->boxPoints, d_group,
                                                             P->dtype, s));
    else if (P->separable) RF_HIP_TRY(ctx, rf_launch_blend_sep(P->sep, d_group, P->dtype, s));
    else RF_HIP_TRY(ctx, rf_launch_blend_dev(P->b, d_group, P->dtype, s));
    return RF_OK;
}

extern "C" int rf_blend_plan_destroy(rf_blend_plan* P) {
    if (!P) return RF_OK;
    (void)hipSetDevice(P->ctx->device);
    (void)hipFree(P->d_blob);          // the caller has no launch of this plan pending
    if (P->d_boxes) (void)hipFree(P->d_boxes);
    if (P->d_subFused) (void)hipFree(P->d_subFused);
    if (P->d_bands) (void)hipFree(P->d_bands);
    delete P;
    return RF_OK;
}

extern "C" int rf_blend_device(rf_context* ctx, int32_t nDim, const int32_t* groupN, int32_t Nmc,
                               int32_t nSub, const int32_t* origins, const int32_t* sizes,
                               const int32_t* nOvl, const int32_t* faceMasks, int32_t mode,
                               const void* const* d_fields, void* d_group, int32_t dtype,
                               void* stream) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!d_group) return rf_set_error(ctx, RF_E_ARG, "NULL blend argument");
    rf_blend_plan* P = nullptr;
    int rc = rf_blend_plan_create(ctx, nDim, groupN, Nmc, nSub, origins, sizes, nOvl, faceMasks, mode, d_fields,
                                  dtype, &P);
    if (rc) return rc;
    rc = rf_blend_plan_execute(P, d_group, stream);
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const hipError_t e = hipStreamSynchronize(s);   // tables freed below
    rf_blend_plan_destroy(P);
    if (rc) return rc;
    if (e != hipSuccess) return rf_set_error(ctx, RF_E_HIP, std::string("blend: ") + hipGetErrorString(e));
    return RF_OK;
}

// generateUnityPartition_Matrix (localization_RF.f90:705-857) as a host
// table: pu[p] = product over tapered faces of the face factor, multiplied
// face by face in the reference's order (lower faces of every axis, then
// upper faces, :755-777).
extern "C" int rf_unity_partition(int32_t nDim, const int32_t* xNStep, const int32_t* nOvl, int32_t faceMask,
                                  int32_t mode, double* pu) {
    if (nDim < 1 || nDim > 3 || !xNStep || !nOvl || !pu)
        return rf_set_error(nullptr, RF_E_ARG, "rf_unity_partition: bad arguments");
    if (mode != RF_REF_COMPAT && mode != RF_FIXED) return rf_set_error(nullptr, RF_E_ARG, "unknown mode");
    int n[3] = {1, 1, 1};
    int64_t tot = 1;
    for (int a = 0; a < nDim; a++) {
        if (xNStep[a] < 1) return rf_set_error(nullptr, RF_E_ARG, "rf_unity_partition: xNStep must be >= 1");
        for (int f = 0; f < 2; f++)
            if (((faceMask >> (2 * a + f)) & 1) && nOvl[a] > xNStep[a])
                return rf_set_error(nullptr, RF_E_ARG, "rf_unity_partition: overlap wider than the box");
        n[a] = xNStep[a];
        tot *= n[a];
    }
    std::vector<double> low[3], up[3];
    for (int a = 0; a < 3; a++) {
        low[a].assign((size_t)n[a], 1.0);
        up[a].assign((size_t)n[a], 1.0);
        if (a < nDim)
            axis_factors(n[a], nOvl[a], (faceMask >> (2 * a)) & 1, (faceMask >> (2 * a + 1)) & 1, mode,
                         low[a].data(), up[a].data());
    }
    for (int64_t p = 0; p < tot; p++) {
        const int i0 = (int)(p % n[0]), i1 = (int)((p / n[0]) % n[1]), i2 = (int)(p / ((int64_t)n[0] * n[1]));
        double w = 1.0;
        w = w * low[0][i0];
        w = w * low[1][i1];
        w = w * low[2][i2];
        w = w * up[0][i0];
        w = w * up[1][i1];
        w = w * up[2][i2];
        pu[p] = w;
    }
    return RF_OK;
}

// The blend of rf_blend_device on host buffers: fields[s] (sizes[s] product
// x Nmc, FP64) uploaded, blended on the device, group (prod(groupN) x Nmc)
// downloaded.  Synchronous.
extern "C" int rf_blend_host(rf_context* ctx, int32_t nDim, const int32_t* groupN, int32_t Nmc, int32_t nSub,
                             const int32_t* origins, const int32_t* sizes, const int32_t* nOvl,
                             const int32_t* faceMasks, int32_t mode, const double* const* fields, double* group) {
    if (!ctx) return rf_set_error(nullptr, RF_E_ARG, "ctx is NULL");
    if (!fields || !group || !sizes || !groupN || nSub < 1 || nDim < 1 || nDim > 3 || Nmc < 1)
        return rf_set_error(ctx, RF_E_ARG, "NULL blend argument");
    RF_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int64_t gTot = 1;
    for (int a = 0; a < nDim; a++) gTot *= groupN[a];
    std::vector<int64_t> off((size_t)nSub + 1, 0);
    for (int s = 0; s < nSub; s++) {
        int64_t t = 1;
        for (int a = 0; a < nDim; a++) t *= sizes[s * nDim + a];
        if (t < 1 || !fields[s]) return rf_set_error(ctx, RF_E_ARG, "bad subdomain field");
        off[(size_t)s + 1] = off[(size_t)s] + t * Nmc;
    }
    const size_t fieldBytes = (size_t)off[(size_t)nSub] * sizeof(double);
    const size_t groupBytes = (size_t)gTot * Nmc * sizeof(double);
    char* d = nullptr;
    if (hipMalloc((void**)&d, fieldBytes + groupBytes) != hipSuccess)
        return rf_set_error(ctx, RF_E_NOMEM, "blend buffers allocation failed");
    double* d_fields = reinterpret_cast<double*>(d);
    double* d_group = reinterpret_cast<double*>(d + fieldBytes);
    std::vector<const void*> ptrs((size_t)nSub);
    hipError_t e = hipSuccess;
    for (int s = 0; s < nSub && e == hipSuccess; s++) {
        ptrs[(size_t)s] = d_fields + off[(size_t)s];
        e = hipMemcpyAsync(d_fields + off[(size_t)s], fields[s],
                           (size_t)(off[(size_t)s + 1] - off[(size_t)s]) * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream);
    }
    int rc = e == hipSuccess ? rf_blend_device(ctx, nDim, groupN, Nmc, nSub, origins, sizes, nOvl, faceMasks, mode,
                                               ptrs.data(), d_group, RF_DTYPE_F64, nullptr)
                             : rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    if (rc == RF_OK) {
        e = hipMemcpyAsync(group, d_group, groupBytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = rf_set_error(ctx, RF_E_HIP, hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    return rc;
}

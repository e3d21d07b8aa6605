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
__global__ __launch_bounds__(kBlock) void k_marginal(int margi, double scale, double shift,
                                                     T* __restrict__ d, int64_t n) {
#pragma clang fp contract(off)
    constexpr int U = 4;                  // independent loads in flight per thread
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += U * stride) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = (e + u * stride < n) ? d[e + u * stride] : T(0);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (e + u * stride < n) {
                const double y = (double)v[u] * scale + shift;                  // :171
                d[e + u * stride] = (T)((margi == RF_FOM_LOGNORMAL) ? exp(y) : y);  // :175
            }
    }
}

// The same map over 16-byte vectors (4 floats or 2 doubles per load): nVec
// vectors from d, which must be 16-byte aligned; the tail goes to k_marginal.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_marginal_vec(int margi, double scale, double shift,
                                                         T* __restrict__ d, int64_t nVec) {
#pragma clang fp contract(off)
    constexpr int L = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(L)));
    vec* v = reinterpret_cast<vec*>(d);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nVec; e += 2 * stride) {
        vec a = v[e];
        vec b = (e + stride < nVec) ? v[e + stride] : a;
#pragma unroll
        for (int l = 0; l < L; l++) {
            const double ya = (double)a[l] * scale + shift;
            const double yb = (double)b[l] * scale + shift;
            a[l] = (T)((margi == RF_FOM_LOGNORMAL) ? exp(ya) : ya);
            b[l] = (T)((margi == RF_FOM_LOGNORMAL) ? exp(yb) : yb);
        }
        v[e] = a;
        if (e + stride < nVec) v[e + stride] = b;
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

// Partial sums of u and u^2 over the box lo..hi (0-based, inclusive) of
// realization 0 of a grid field (x fastest); grid (nPart).  The summed box is
// the window normalize_randField reads through its RF_2D / RF_3D views of
// column 1 (randomFieldND.f90:857-862).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_window_moments_partial(const T* __restrict__ d, rf_window w,
                                                                   double* __restrict__ part) {
    __shared__ double s1[kBlock], s2[kBlock];
    double a = 0.0, b = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < w.count; q += stride) {
        const int64_t i0 = q % w.ext[0];
        const int64_t r = q / w.ext[0];
        const int64_t i1 = r % w.ext[1];
        const int64_t i2 = r / w.ext[1];
        const int64_t off = (w.lo[0] + i0) + w.n[0] * ((w.lo[1] + i1) + w.n[1] * (w.lo[2] + i2));
        const double v = (double)d[off];
        a += v;
        b = fma(v, v, b);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int k = kBlock / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s1[threadIdx.x] += s1[threadIdx.x + k];
            s2[threadIdx.x] += s2[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[(int64_t)blockIdx.x * 2 + 0] = s1[0];
        part[(int64_t)blockIdx.x * 2 + 1] = s2[0];
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

// ------------------------------------------------- directional correlation
// rebuild_Sk_FFT (statistics_RF.f90:262-409) restated as the circular
// autocorrelation it computes: for lag m (grid y) the sum over all points of
// v(x) v(x + m e_a mod n_a), v = (u - avg)/std, in FP64; partials per block
// (grid x), reduced in fixed order by k_corr_final.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_corr_partial(const T* __restrict__ u, int64_t N, int64_t strideA,
                                                         int nA, double avg, double std,
                                                         double* __restrict__ part) {
    __shared__ double sh[kBlock];
    const int m = blockIdx.y;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
        const int64_t ia = (i / strideA) % nA;
        int64_t ja = ia + m;
        if (ja >= nA) ja -= nA;
        const int64_t j = i + (ja - ia) * strideA;
        const double vi = ((double)u[i] - avg) / std;
        const double vj = ((double)u[j] - avg) / std;
        acc = fma(vi, vj, acc);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)m * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_corr_final(const double* __restrict__ part, int nPart, int M, double invN,
                             double* __restrict__ R) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double s = 0.0;
    for (int p = 0; p < nPart; p++) s += part[(int64_t)m * nPart + p];
    R[m] = s * invN;
}

// ------------------------------------------------ interpolation to nodes
// interpolateToMesh (sample_RF.f90:872-1106) for every realization: one
// thread per node and chunk of kInterpR realizations (grid y).  1-based cell
// position c = (x - xMin)/xStep + 1 in FP64, cell floor(c) moved down one on
// the upper border (:1024), corners in the reference's order (:996-1013),
// weight prod(1 - |c - corner|), sum in corner order (:1072-1086).  Nodes
// outside the grid get NaN and raise *flag.
// The grid in memory may be a box of a larger grid: global indices lo..hi
// (1-based, hi = lo + nstep - 1), positions still measured from the larger
// grid's xMin, and hi is the border of the cell rule, as the reference reads
// the bounding box of the nodes (:936-975).
struct rf_interp_dev {
    int nDim, Nmc;
    int nstep[3];
    int lo[3], hi[3];
    double xMin[3], xStep[3];
    int64_t Ng, nNodes;
};
constexpr int kInterpR = 4;

template <typename TG, typename TO>
__global__ __launch_bounds__(kBlock) void k_interp(rf_interp_dev q, const TG* __restrict__ grid,
                                                   const double* __restrict__ x, TO* __restrict__ out,
                                                   int* __restrict__ flag) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= q.nNodes) return;
    const int r0 = blockIdx.y * kInterpR;
    double c[3] = {0.0, 0.0, 0.0};
    int64_t ci[3] = {1, 1, 1};
    bool inside = true;
    for (int a = 0; a < q.nDim; a++) {
        c[a] = (x[i * q.nDim + a] - q.xMin[a]) / q.xStep[a] + 1.0;
        ci[a] = (int64_t)floor(c[a]);
        if (ci[a] == q.hi[a]) ci[a] -= 1;
        if (ci[a] < q.lo[a] || ci[a] + 1 > q.hi[a]) inside = false;
    }
    if (!inside) {
        flag[0] = 1;
        for (int r = r0; r < r0 + kInterpR && r < q.Nmc; r++) out[(int64_t)r * q.nNodes + i] = (TO)NAN;
        return;
    }
    const int64_t s1 = q.nstep[0], s2 = (int64_t)q.nstep[0] * q.nstep[1];
    // corner offsets in the reference's order; 2-D uses the first four
    // entries of the 2-D table, 1-D (0), (1)
    const int o3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    const int o2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    const int nc = 1 << q.nDim;
    double w[8];
    int64_t idx[8];
    for (int j = 0; j < 8; j++) {
        w[j] = 0.0;
        idx[j] = 0;
        if (j >= nc) continue;
        double wj = 1.0;
        int64_t id = 0;
        for (int a = 0; a < q.nDim; a++) {
            const int off = q.nDim == 3 ? o3[j][a] : (q.nDim == 2 ? o2[j][a] : j);
            wj = wj * (1.0 - fabs(c[a] - (double)(ci[a] + off)));
            id += (ci[a] + off - q.lo[a]) * (a == 0 ? 1 : (a == 1 ? s1 : s2));
        }
        w[j] = wj;
        idx[j] = id;
    }
    // the 8 corner loads of a realization are unconditional (unused corners
    // read entry 0) so they issue together; only the first nc enter the sum
#pragma unroll
    for (int rr = 0; rr < kInterpR; rr++) {
        const int r = r0 + rr;
        if (r >= q.Nmc) break;
        const TG* g = grid + (int64_t)r * q.Ng;
        TG v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = g[idx[j]];
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (j < nc) acc = acc + (double)v[j] * w[j];
        out[(int64_t)r * q.nNodes + i] = (TO)acc;
    }
}

// ----------------------------------------------------------------- blend
// One thread per group point and chunk of kBlendR realizations (grid y), so
// the blocks in flight sweep the points of a few realizations at a time (few
// pages per wave) and the loads of a chunk are independent.  Sums run in FP64
// in ascending subdomain order (the reference's loop over fields,
// calls_RF.f90:237), each term v * PU then + (localization_RF.f90:228,253).
constexpr int kBlendR = 4;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_blend(rf_blend_dev b, T* __restrict__ group) {
#pragma clang fp contract(off)
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= b.groupTot) return;
    int gi[3];
    {
        int64_t rem = g;
        gi[0] = (int)(rem % b.groupN[0]);
        rem /= b.groupN[0];
        gi[1] = (int)(rem % b.groupN[1]);
        gi[2] = (int)(rem / b.groupN[1]);
    }
    int cell = b.seg[0][gi[0]];
    if (b.nDim > 1) cell += b.nSeg[0] * b.seg[1][gi[1]];
    if (b.nDim > 2) cell += b.nSeg[0] * b.nSeg[1] * b.seg[2][gi[2]];
    const int e0 = b.cellOff[cell];
    const int nh = b.cellOff[cell + 1] - e0;
    const T* F[kBlendMaxEntries];
    int64_t rs[kBlendMaxEntries], loc[kBlendMaxEntries];
    double w[kBlendMaxEntries];
#pragma unroll
    for (int h = 0; h < kBlendMaxEntries; h++) {
        F[h] = nullptr;
        rs[h] = 0;
        loc[h] = 0;
        w[h] = 0.0;
        if (h < nh) {
            const rf_blend_entry& E = b.entries[e0 + h];
            F[h] = static_cast<const T*>(E.field);
            rs[h] = E.subTot;
            loc[h] = E.base + gi[0] + (int64_t)gi[1] * E.stride1 + (int64_t)gi[2] * E.stride2;
            // PU product in the reference's order: lower faces of axes 1..d,
            // then upper faces (localization_RF.f90:755-851)
            // (all six factor loads issued together; axes beyond nDim read
            // entry 0 and multiply by an exact 1.0)
            const double* lf = b.lowF + (int64_t)E.s * 3 * b.maxSize;
            const double* uf = b.upF + (int64_t)E.s * 3 * b.maxSize;
            double fl[3], fu[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int pos = a < b.nDim ? gi[a] - E.o[a] : 0;
                fl[a] = lf[a * b.maxSize + pos];
                fu[a] = uf[a * b.maxSize + pos];
            }
            double pu = 1.0;
#pragma unroll
            for (int a = 0; a < 3; a++) pu = pu * (a < b.nDim ? fl[a] : 1.0);
#pragma unroll
            for (int a = 0; a < 3; a++) pu = pu * (a < b.nDim ? fu[a] : 1.0);
            w[h] = pu;
        }
    }
    double pu1 = 1.0;
    if (b.mode == RF_REF_COMPAT) {
        for (int a = 0; a < b.nDim; a++) pu1 = pu1 * b.pu1Low[a][gi[a]];
        for (int a = 0; a < b.nDim; a++) pu1 = pu1 * b.pu1Up[a][gi[a]];
    }
    {
        const int r0 = blockIdx.y * kBlendR;
        double acc[kBlendR];
#pragma unroll
        for (int j = 0; j < kBlendR; j++) acc[j] = 0.0;
        // the chunk's loads are unconditional (realization index clamped), so
        // they issue together instead of one branch and wait per load
#pragma unroll
        for (int h = 0; h < kBlendMaxEntries; h++) {
            if (h < nh) {
                T v[kBlendR];
#pragma unroll
                for (int j = 0; j < kBlendR; j++) v[j] = F[h][(int64_t)min(r0 + j, b.Nmc - 1) * rs[h] + loc[h]];
#pragma unroll
                for (int j = 0; j < kBlendR; j++) acc[j] = (double)v[j] * w[h] + acc[j];
            }
        }
#pragma unroll
        for (int j = 0; j < kBlendR; j++)
            if (r0 + j < b.Nmc) {
                double u = acc[j];
                if (b.mode == RF_REF_COMPAT && r0 + j == 0) u = u / pu1;   // calls_RF.f90:355
                group[(int64_t)(r0 + j) * b.groupTot + g] = (T)u;
            }
    }
}

// Separable form: one thread per group point, all realizations; the <= 2^d
// contributions come from the per-axis tables.  FP64 sums, each term v * PU
// then + (localization_RF.f90:228,253); the PU product is taken per axis
// (a reordering of the reference's lower-then-upper face product, equal up
// to rounding) and contributions are added in interval order.
//
// Every field load is unconditional (realization index clamped, absent
// contributions read a dummy address with stride 0 and are masked out of
// the sum): a load under a per-element condition becomes a branch that waits
// for its value before the next load issues, which leaves one load in
// flight per wave.  A wave whose points have at most two contributions
// takes a straight-line path chosen wave-uniformly (16 realizations of
// loads in flight for one contribution, 2 x 8 for two); larger counts (the
// corners where 3-D overlaps meet) take the general loop.
template <typename T>
__device__ __forceinline__ void blend_store(T* p, double u) {
    __builtin_nontemporal_store((T)u, p);   // written once, read by the next kernel
}

// Where field S holds its value at local position (p0, p1, p2) (s0, s1: its
// first two extents): a first-axis overlap band array of a fused batch, or
// the field itself; *st: the realization stride.
// Written with selects: an if-return form of this lookup was compiled (ROCm
// 7.2, gfx950) so that lanes leaving the high-band test for the field kept
// the band width as their stride.
template <typename T>
__device__ __forceinline__ const T* blend_locate(const rf_blend_sub& S, int64_t p0, int64_t p1, int64_t p2,
                                                 int64_t s0, int64_t s1, int64_t* st) {
    const int64_t row = p1 + p2 * s1;
    const int64_t wLo = S.wLo, wHi = S.wHi;
    const int64_t tLo = S.bandLoTot, tHi = S.bandHiTot, tF = S.subTot;
    const T* bLo = static_cast<const T*>(S.bandLo);
    const T* bHi = static_cast<const T*>(S.bandHi);
    const T* bF = static_cast<const T*>(S.field);
    const bool lo = bLo != nullptr && p0 < wLo;
    const bool hi = !lo && bHi != nullptr && p0 >= s0 - wHi;
    const int64_t offField = p0 + row * s0, offLo = row * wLo + p0, offHi = row * wHi + (p0 - (s0 - wHi));
    *st = lo ? tLo : (hi ? tHi : tF);
    return lo ? bLo + offLo : (hi ? bHi + offHi : bF + offField);
}

// One group point (gi0, gi1, gi2) of the separable blend, every realization.
template <typename T>
__device__ __forceinline__ void blend_sep_point(const rf_blend_sep& b, int gi0, int gi1, int gi2,
                                                T* __restrict__ group) {
#pragma clang fp contract(off)
    const int64_t g = gi0 + (int64_t)b.groupN[0] * (gi1 + (int64_t)b.groupN[1] * gi2);
    const rf_blend_axis A0 = b.ax[0][gi0];
    const rf_blend_axis A1 = b.ax[1][gi1];
    const rf_blend_axis A2 = b.ax[2][gi2];
    const double pu1 = (A0.pu1 * A1.pu1) * A2.pu1;
    const bool fix0 = b.mode == RF_REF_COMPAT;      // realization 0 divided by pu1 (calls_RF.f90:355)
    const int Nmc = b.Nmc;
    // the interval combinations covering this point (bit k: interval k & 1 of
    // axis 0, (k >> 1) & 1 of axis 1, k >> 2 of axis 2)
    uint32_t live = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int i0 = (k & 1) ? A0.iv[1] : A0.iv[0];
        const int i1 = ((k >> 1) & 1) ? A1.iv[1] : A1.iv[0];
        const int i2 = (k >> 2) ? A2.iv[1] : A2.iv[0];
        if (i0 >= 0 && i1 >= 0 && i2 >= 0) live |= 1u << k;
    }
    // the first two contributions in interval order; absent ones read the
    // axis-0 table entry with stride 0
    const T* dummy = reinterpret_cast<const T*>(b.ax[0]);
    const T* F0 = dummy; const T* F1 = dummy;
    int64_t st0 = 0, st1 = 0;
    double w0 = 0.0, w1 = 0.0;
    int n = 0;
    for (uint32_t lm = live; lm; lm &= lm - 1) {
        const int k = __builtin_ctz(lm);
        // selects, not indexing: the tables stay in registers
        const bool c0 = k & 1, c1 = (k >> 1) & 1, c2 = k >> 2;
        const int i0 = c0 ? A0.iv[1] : A0.iv[0];
        const int i1 = c1 ? A1.iv[1] : A1.iv[0];
        const int i2 = c2 ? A2.iv[1] : A2.iv[0];
        const rf_blend_sub S = b.sub[i0 + b.nInt[0] * (i1 + b.nInt[1] * i2)];
        if (!S.field) continue;
        if (n == 2) { n = 3; break; }
        const int64_t s0 = c0 ? A0.sz[1] : A0.sz[0], s1 = c1 ? A1.sz[1] : A1.sz[0];
        const int64_t p0 = c0 ? A0.pos[1] : A0.pos[0], p1 = c1 ? A1.pos[1] : A1.pos[0];
        const int64_t p2 = c2 ? A2.pos[1] : A2.pos[0];
        int64_t st;
        const T* F = blend_locate<T>(S, p0, p1, p2, s0, s1, &st);
        const double w = ((c0 ? A0.f[1] : A0.f[0]) * (c1 ? A1.f[1] : A1.f[0])) * (c2 ? A2.f[1] : A2.f[0]);
        if (n == 0) { F0 = F; st0 = st; w0 = w; }
        else { F1 = F; st1 = st; w1 = w; }
        n++;
    }
    if (b.skipSingle && n == 1) return;            // written by the fused last contraction
    if (!__any(n > 2)) {
        if (!__any(n > 1)) {
            constexpr int RB = 16;
            for (int r0 = 0; r0 < Nmc; r0 += RB) {
                T v[RB];
#pragma unroll
                for (int j = 0; j < RB; j++) v[j] = F0[(int64_t)min(r0 + j, Nmc - 1) * st0];
#pragma unroll
                for (int j = 0; j < RB; j++)
                    if (r0 + j < Nmc) {
                        double u = 0.0;
                        if (n == 1) u = (double)v[j] * w0 + u;
                        if (fix0 && r0 + j == 0) u = u / pu1;
                        blend_store(group + (int64_t)(r0 + j) * b.groupTot + g, u);
                    }
            }
        } else {
            constexpr int RB = 8;
            for (int r0 = 0; r0 < Nmc; r0 += RB) {
                T v0[RB], v1[RB];
#pragma unroll
                for (int j = 0; j < RB; j++) {
                    const int64_t r = min(r0 + j, Nmc - 1);
                    v0[j] = F0[r * st0];
                    v1[j] = F1[r * st1];
                }
#pragma unroll
                for (int j = 0; j < RB; j++)
                    if (r0 + j < Nmc) {
                        double u = 0.0;
                        if (n >= 1) u = (double)v0[j] * w0 + u;
                        if (n >= 2) u = (double)v1[j] * w1 + u;
                        if (fix0 && r0 + j == 0) u = u / pu1;
                        blend_store(group + (int64_t)(r0 + j) * b.groupTot + g, u);
                    }
            }
        }
        return;
    }
    // general loop: every contribution, 8 realizations per pass
    constexpr int RB = 8;
    for (int r0 = 0; r0 < Nmc; r0 += RB) {
        double acc[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) acc[j] = 0.0;
        for (uint32_t lm = live; lm; lm &= lm - 1) {
            const int k = __builtin_ctz(lm);
            const bool c0 = k & 1, c1 = (k >> 1) & 1, c2 = k >> 2;
            const int i0 = c0 ? A0.iv[1] : A0.iv[0];
            const int i1 = c1 ? A1.iv[1] : A1.iv[0];
            const int i2 = c2 ? A2.iv[1] : A2.iv[0];
            const rf_blend_sub S = b.sub[i0 + b.nInt[0] * (i1 + b.nInt[1] * i2)];
            if (!S.field) continue;
            const int64_t s0 = c0 ? A0.sz[1] : A0.sz[0], s1 = c1 ? A1.sz[1] : A1.sz[0];
            const int64_t p0 = c0 ? A0.pos[1] : A0.pos[0], p1 = c1 ? A1.pos[1] : A1.pos[0];
            const int64_t p2 = c2 ? A2.pos[1] : A2.pos[0];
            int64_t st;
            const T* F = blend_locate<T>(S, p0, p1, p2, s0, s1, &st);
            const double w = ((c0 ? A0.f[1] : A0.f[0]) * (c1 ? A1.f[1] : A1.f[0])) * (c2 ? A2.f[1] : A2.f[0]);
            T v[RB];
#pragma unroll
            for (int j = 0; j < RB; j++) v[j] = F[(int64_t)min(r0 + j, Nmc - 1) * st];
#pragma unroll
            for (int j = 0; j < RB; j++) acc[j] = (double)v[j] * w + acc[j];
        }
#pragma unroll
        for (int j = 0; j < RB; j++)
            if (r0 + j < Nmc) {
                double u = acc[j];
                if (fix0 && r0 + j == 0) u = u / pu1;
                blend_store(group + (int64_t)(r0 + j) * b.groupTot + g, u);
            }
    }
}

// 7 waves per SIMD (at most 72 VGPRs, no spills; unconstrained the
// allocation is 81 VGPRs, 5 waves)
// block (x segment, gi1, gi2): the axis-1 and axis-2 entries are
// block-uniform (scalar loads), one axis-0 entry per thread
template <typename T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_blend_sep(rf_blend_sep b, T* __restrict__ group) {
    const int gi0 = blockIdx.x * kBlock + threadIdx.x;
    if (gi0 >= b.groupN[0]) return;
    blend_sep_point<T>(b, gi0, blockIdx.y, blockIdx.z, group);
}

// The overlap bands only (the fused blend: the last contraction stored the
// points one subdomain covers): boxes of group indices, box k covering
// points first .. first + n0 n1 n2 - 1 of the launch, x fastest.
template <typename T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_blend_boxes(
        rf_blend_sep b, const rf_blend_box* __restrict__ boxes, int nBoxes, int64_t total, T* __restrict__ group) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    int k = 0;
    while (k + 1 < nBoxes && boxes[k + 1].first <= e) k++;
    const rf_blend_box& x = boxes[k];
    const int64_t q = e - x.first;
    const int i0 = (int)(q % x.n[0]);
    const int64_t r = q / x.n[0];
    blend_sep_point<T>(b, x.lo[0] + i0, x.lo[1] + (int)(r % x.n[1]), x.lo[2] + (int)(r / x.n[1]), group);
}

}  // namespace

hipError_t rf_launch_marginal(int margi, double scale, double shift, void* d, int64_t count,
                              int dtype, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const size_t esz = dtype == RF_DTYPE_F64 ? 8 : 4;
    int64_t head = 0;                        // elements done by the vector kernel
    if (((uintptr_t)d & 15) == 0) {
        const int64_t L = (int64_t)(16 / esz);
        const int64_t nVec = count / L;
        if (nVec > 0) {
            int64_t blocks = (nVec + 2 * kBlock - 1) / (2 * kBlock);
            if (blocks > 65536) blocks = 65536;
            if (dtype == RF_DTYPE_F64)
                hipLaunchKernelGGL(k_marginal_vec<double>, dim3((unsigned)blocks), dim3(kBlock), 0, s,
                                   margi, scale, shift, static_cast<double*>(d), nVec);
            else
                hipLaunchKernelGGL(k_marginal_vec<float>, dim3((unsigned)blocks), dim3(kBlock), 0, s,
                                   margi, scale, shift, static_cast<float*>(d), nVec);
            head = nVec * L;
        }
    }
    const int64_t rest = count - head;
    if (rest > 0) {
        int64_t blocks = (rest + 4 * kBlock - 1) / (4 * kBlock);
        if (blocks > 65536) blocks = 65536;
        if (dtype == RF_DTYPE_F64)
            hipLaunchKernelGGL(k_marginal<double>, dim3((unsigned)blocks), dim3(kBlock), 0, s, margi,
                               scale, shift, static_cast<double*>(d) + head, rest);
        else
            hipLaunchKernelGGL(k_marginal<float>, dim3((unsigned)blocks), dim3(kBlock), 0, s, margi,
                               scale, shift, static_cast<float*>(d) + head, rest);
    }
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

hipError_t rf_launch_window_moments(const void* d, int dtype, const rf_window& w, double* d_part, int nPart,
                                    double* d_sums, hipStream_t s) {
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_window_moments_partial<double>, dim3((unsigned)nPart), dim3(kBlock), 0, s,
                           static_cast<const double*>(d), w, d_part);
    else
        hipLaunchKernelGGL(k_window_moments_partial<float>, dim3((unsigned)nPart), dim3(kBlock), 0, s,
                           static_cast<const float*>(d), w, d_part);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, s, d_part, nPart, 1, d_sums);
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
    const dim3 grid((unsigned)((b.groupTot + kBlock - 1) / kBlock), (unsigned)((b.Nmc + kBlendR - 1) / kBlendR));
    if (grid.y > 65535) return hipErrorInvalidValue;
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_blend<double>, grid, dim3(kBlock), 0, s, b, static_cast<double*>(d_group));
    else
        hipLaunchKernelGGL(k_blend<float>, grid, dim3(kBlock), 0, s, b, static_cast<float*>(d_group));
    return hipGetLastError();
}

int rf_corr_parts(int64_t N) {
    int64_t p = (N + kBlock * 16 - 1) / (kBlock * 16);
    if (p < 1) p = 1;
    if (p > 512) p = 512;
    return (int)p;
}

hipError_t rf_launch_corr(const void* u, int dtype, int64_t N, int64_t strideA, int nA, int M, double avg,
                          double std, double* d_part, int nPart, double* d_R, hipStream_t s) {
    if (M <= 0) return hipSuccess;
    const dim3 grid((unsigned)nPart, (unsigned)M);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_corr_partial<double>, grid, dim3(kBlock), 0, s, static_cast<const double*>(u), N,
                           strideA, nA, avg, std, d_part);
    else
        hipLaunchKernelGGL(k_corr_partial<float>, grid, dim3(kBlock), 0, s, static_cast<const float*>(u), N,
                           strideA, nA, avg, std, d_part);
    hipLaunchKernelGGL(k_corr_final, dim3((unsigned)((M + 63) / 64)), dim3(64), 0, s, d_part, nPart, M,
                       1.0 / (double)N, d_R);
    return hipGetLastError();
}

hipError_t rf_launch_interp(int nDim, const int32_t* nstep, const int32_t* lo, const double* xMin,
                            const double* xStep, int Nmc, const void* grid, int gridDtype, const double* x,
                            int64_t nNodes, void* out, int outDtype, int* d_flag, hipStream_t s) {
    if (nNodes <= 0) return hipSuccess;
    rf_interp_dev q{};
    q.nDim = nDim;
    q.Nmc = Nmc;
    q.Ng = 1;
    for (int a = 0; a < 3; a++) {
        q.nstep[a] = a < nDim ? nstep[a] : 1;
        q.lo[a] = (a < nDim && lo) ? lo[a] : 1;
        q.hi[a] = q.lo[a] + q.nstep[a] - 1;
        q.xMin[a] = a < nDim ? xMin[a] : 0.0;
        q.xStep[a] = a < nDim ? xStep[a] : 1.0;
        q.Ng *= q.nstep[a];
    }
    q.nNodes = nNodes;
    const dim3 grid2((unsigned)((nNodes + kBlock - 1) / kBlock), (unsigned)((Nmc + kInterpR - 1) / kInterpR));
    if (grid2.y > 65535) return hipErrorInvalidValue;
    const bool g64 = gridDtype == RF_DTYPE_F64, o64 = outDtype == RF_DTYPE_F64;
    if (g64 && o64)
        hipLaunchKernelGGL((k_interp<double, double>), grid2, dim3(kBlock), 0, s, q, static_cast<const double*>(grid), x, static_cast<double*>(out), d_flag);
    else if (g64)
        hipLaunchKernelGGL((k_interp<double, float>), grid2, dim3(kBlock), 0, s, q, static_cast<const double*>(grid), x, static_cast<float*>(out), d_flag);
    else if (o64)
        hipLaunchKernelGGL((k_interp<float, double>), grid2, dim3(kBlock), 0, s, q, static_cast<const float*>(grid), x, static_cast<double*>(out), d_flag);
    else
        hipLaunchKernelGGL((k_interp<float, float>), grid2, dim3(kBlock), 0, s, q, static_cast<const float*>(grid), x, static_cast<float*>(out), d_flag);
    return hipGetLastError();
}

hipError_t rf_launch_blend_boxes(const rf_blend_sep& b, const rf_blend_box* d_boxes, int nBoxes, int64_t total,
                                 void* d_group, int dtype, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((total + kBlock - 1) / kBlock);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_blend_boxes<double>, dim3(grid), dim3(kBlock), 0, s, b, d_boxes, nBoxes, total,
                           static_cast<double*>(d_group));
    else
        hipLaunchKernelGGL(k_blend_boxes<float>, dim3(grid), dim3(kBlock), 0, s, b, d_boxes, nBoxes, total,
                           static_cast<float*>(d_group));
    return hipGetLastError();
}

hipError_t rf_launch_blend_sep(const rf_blend_sep& b, void* d_group, int dtype, hipStream_t s) {
    if (b.groupN[1] > 65535 || b.groupN[2] > 65535) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((b.groupN[0] + kBlock - 1) / kBlock), (unsigned)b.groupN[1], (unsigned)b.groupN[2]);
    if (dtype == RF_DTYPE_F64)
        hipLaunchKernelGGL(k_blend_sep<double>, grid, dim3(kBlock), 0, s, b, static_cast<double*>(d_group));
    else
        hipLaunchKernelGGL(k_blend_sep<float>, grid, dim3(kBlock), 0, s, b, static_cast<float*>(d_group));
    return hipGetLastError();
}

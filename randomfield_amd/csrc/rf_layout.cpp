// Localization geometry: the subdomain boxes, lattices and overlaps the
// reference derives from its input file before it generates one field per
// subdomain and blends them (one MPI rank per subdomain there, one batch
// entry here).
//
// Reference steps, one localization level (subdomains per axis = nFields):
//   round_basic_inputs   topography_RF.f90:19-58   xStep = corrL/pointsPerCorrL;
//                        overlap rounded to whole steps, at least 2 steps
//                        where an axis is split, 0 where it is not
//   set_global_extremes  topography_RF.f90:63-144  xMin rounded down, xMax up
//                        to whole steps; the non-overlapping length rounded
//                        up to a multiple of nFields steps; subdomain extent
//                        and stride; the global box grown to fit them
//   prepare_Localization localization_RF.f90:127-192  subdomain starts on the
//                        stride grid, field numbers with the LAST axis fastest
//                        (setGrid inverse); the group box and its point count
//   set_local_bounding_box topography_RF.f90:150-255  one rank per subdomain:
//                        the box is [start, start + extent], xRange = extent
//                        (the range that sizes the subdomain's wavevector
//                        lattice), points per axis 1 + nint(extent/xStep)
//   generateUnityPartition_Matrix localization_RF.f90:746  overlap points
//                        nint((overlap corrL - 2 xStep)/xStep) + 1
//
// Every length here is a whole number of steps once xStep is fixed, so the
// rules are evaluated in integer step counts; only the rounding of the
// caller's box and overlap to steps uses the reference's tolerance
// (areEqual with TOLERANCE = 1e-9, math_RF.f90:397-413,419-450).
#include <cmath>
#include <cstdint>

#include "rf_internal.h"

namespace {

constexpr double kTol = 1e-9;   // constants_RF.f90:8

// roundToMultiple(nmb, 1, up) on a quotient already divided by the step
int64_t round_steps(double q, bool up) {
    const double r = std::nearbyint(q);
    if (std::fabs(q - r) < kTol) return (int64_t)r;
    return (int64_t)(up ? std::ceil(q) : std::floor(q));
}

}  // namespace

extern "C" int rf_localization_layout(const rf_layout_params* in, rf_layout_info* info, int32_t* origins,
                                      int32_t* sizes, int32_t* faceMasks, double* xMinSub) {
    if (!in || !info) return rf_set_error(nullptr, RF_E_ARG, "rf_localization_layout: NULL argument");
    const int d = in->nDim;
    if (d < 1 || d > 3) return rf_set_error(nullptr, RF_E_ARG, "rf_localization_layout: nDim must be 1..3");
    *info = rf_layout_info{};
    info->nDim = d;
    int64_t nSub = 1;
    int64_t ext[3], stride[3], iMin[3];
    for (int a = 0; a < d; a++) {
        const int P = in->nFields[a];
        if (P < 1 || in->pointsPerCorrL[a] < 1 || !(in->corrL[a] > 0.0) || !(in->overlap[a] >= 0.0) ||
            !(in->xMaxGlob[a] >= in->xMinGlob[a]))
            return rf_set_error(nullptr, RF_E_ARG,
                                "rf_localization_layout: need nFields >= 1, pointsPerCorrL >= 1, corrL > 0, "
                                "overlap >= 0 and xMaxGlob >= xMinGlob per axis");
        nSub *= P;
        // round_basic_inputs
        const double xs = in->corrL[a] / (double)in->pointsPerCorrL[a];
        int64_t oa = 0;
        if (P > 1) {
            oa = (int64_t)std::nearbyint(in->overlap[a] * in->corrL[a] / xs);
            if (oa < 2) oa = 2;
        }
        info->xStep[a] = xs;
        info->overlap[a] = (double)oa * xs / in->corrL[a];
        // set_global_extremes, in steps
        iMin[a] = round_steps(in->xMinGlob[a] / xs, false);
        const int64_t iMax = round_steps(in->xMaxGlob[a] / xs, true);
        const int64_t ovlTot = P > 1 ? oa * (P - 1) + 2 * (oa - 1) : 0;
        int64_t core = (iMax - iMin[a]) - ovlTot;
        if (core < 0) core = 0;
        core = (core + P - 1) / P * P;                 // up to a multiple of nFields steps
        const int64_t per = core / P;
        ext[a] = P > 1 ? per + 2 * oa - 2 : per;        // subdomain extent (delta_Proc)
        stride[a] = per + oa;                           // stepProc
        if (ext[a] + 1 > INT32_MAX || (int64_t)(P - 1) * stride[a] + ext[a] + 1 > INT32_MAX)
            return rf_set_error(nullptr, RF_E_ARG, "rf_localization_layout: grid too large");
        info->xMinGlob[a] = (double)iMin[a] * xs;
        info->xMaxGlob[a] = info->xMinGlob[a] + (double)(ovlTot + core) * xs;
        info->procExtent[a] = (double)ext[a] * xs;
        info->stepProc[a] = (double)stride[a] * xs;
        info->xRangeSub[a] = info->procExtent[a];
        info->subN[a] = (int32_t)(ext[a] + 1);
        info->groupN[a] = (int32_t)((P - 1) * stride[a] + ext[a] + 1);
        // localization_RF.f90:746 (negative where the axis is not split)
        info->nOvlRaw[a] = (int32_t)std::lround((info->overlap[a] * in->corrL[a] - 2.0 * xs) / xs) + 1;
        info->nOvl[a] = info->nOvlRaw[a] > 0 ? info->nOvlRaw[a] : 0;
    }
    if (nSub > INT32_MAX) return rf_set_error(nullptr, RF_E_ARG, "rf_localization_layout: too many subdomains");
    info->nSub = (int32_t)nSub;
    // subdomain s is the reference's field s + 1: last axis fastest (setGrid inverse)
    for (int64_t s = 0; s < nSub; s++) {
        int64_t rem = s;
        int32_t c[3] = {0, 0, 0};
        for (int a = d - 1; a >= 0; a--) {
            c[a] = (int32_t)(rem % in->nFields[a]);
            rem /= in->nFields[a];
        }
        int32_t mask = 0;
        for (int a = 0; a < d; a++) {
            if (origins) origins[s * d + a] = (int32_t)(c[a] * stride[a]);
            if (sizes) sizes[s * d + a] = info->subN[a];
            if (xMinSub) xMinSub[s * d + a] = (double)(iMin[a] + c[a] * stride[a]) * info->xStep[a];
            if (c[a] > 0) mask |= 1 << (2 * a);
            if (c[a] < in->nFields[a] - 1) mask |= 1 << (2 * a + 1);
        }
        if (faceMasks) faceMasks[s] = mask;
    }
    return RF_OK;
}

// rf_blend.h -- device-side description of a tensor-product arrangement of
// overlapping subdomains, consumed by the gather-form blend kernel.
//
// Along each axis the group index range splits into segments over which the
// set of covering subdomain intervals is constant; a cell is a product of
// per-axis segments, and every group point of a cell reads the same <= 2^d
// subdomains.  The cell's entries are listed in ascending subdomain index,
// the order in which the reference adds fields into the group.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

struct rf_blend_entry {
    const void* field;            // subdomain field, [Nmc][subTot]
    int64_t subTot;               // points of the subdomain (realization stride)
    int64_t base;                 // -(o0 + o1 * n0 + o2 * n0 * n1)
    int64_t stride1, stride2;     // n0, n0 * n1
    int32_t s;                    // subdomain index (taper tables)
    int32_t o[3];                 // box origin in group indices
};

struct rf_blend_dev {
    int nDim, Nmc, mode;
    int groupN[3];
    int nSeg[3];
    int64_t groupTot;
    const int32_t* seg[3];        // [g] segment of group index g along axis a
    const int32_t* cellOff;       // [cell .. cell + 1]: entry range of a cell
    const rf_blend_entry* entries;
    const double* lowF;           // [s][3][maxSize] lower-face taper factor by position
    const double* upF;            // [s][3][maxSize] upper-face taper factor by position
    int maxSize;
    const double* pu1Low[3];      // [g] reverse taper of the group grid (REF_COMPAT)
    const double* pu1Up[3];
};

constexpr int kBlendMaxEntries = 8;

// Separable arrangement (every subdomain's tapered faces along axis a depend
// only on its interval along a, as in a tensor-product layout): the PU weight
// is a product of per-axis factors, so one table per axis, indexed by the
// group index, resolves both covering intervals, the positions inside them,
// their sizes and factors; a point needs one table load per axis and one
// subdomain lookup per contribution.
struct rf_blend_axis {
    int32_t iv[2];                // covering intervals (-1: none)
    int32_t pos[2];               // position of the group index inside each
    int32_t sz[2];                // interval lengths
    int32_t pad[2];
    double f[2];                  // taper factor (lower * upper face) at pos
    double pu1;                   // REF_COMPAT reverse taper of the group grid
};
struct rf_blend_sub {
    const void* field;            // NULL: no subdomain on this interval triple
    int64_t subTot;
    // fused with a grid batch: the first-axis overlap bands of the field,
    // [Nmc][n2][n1][wLo] and [Nmc][n2][n1][wHi] (NULL: none)
    const void* bandLo;
    const void* bandHi;
    int64_t bandLoTot, bandHiTot;   // realization strides
    int32_t wLo, wHi;
};
struct rf_blend_sep {
    int nDim, Nmc, mode;
    int groupN[3];
    int nInt[3];
    int64_t groupTot;
    const rf_blend_axis* ax[3];   // [groupN[a]] (one unit entry for a >= nDim)
    const rf_blend_sub* sub;      // [i0 + nInt0 * (i1 + nInt1 * i2)]
    int skipSingle;               // 1: points with one contribution are left
                                  // alone (the fused last contraction wrote them)
};
hipError_t rf_launch_blend_sep(const rf_blend_sep& b, void* d_group, int dtype, hipStream_t s);

// A box of group indices lo .. lo + n - 1 per axis; first = points of the
// boxes before it (the overlap bands the fused blend forms).
struct rf_blend_box {
    int lo[3], n[3];
    int64_t first;
};
hipError_t rf_launch_blend_boxes(const rf_blend_sep& b, const rf_blend_box* d_boxes, int nBoxes, int64_t total,
                                 void* d_group, int dtype, hipStream_t s);

struct rf_context;
// A blend prepared once (rf_blend_plan_create): device tables plus the host
// description a grid batch needs to fuse it into its last contraction.
struct rf_blend_plan {
    rf_context* ctx = nullptr;
    char* d_blob = nullptr;
    rf_blend_dev b{};
    bool separable = false;
    rf_blend_sep sep{};
    int dtype = 0;
    int nDim = 0, mode = 0;
    int groupN[3] = {1, 1, 1};
    std::vector<int32_t> origins, sizes;      // nSub x nDim
    std::vector<const void*> fields;          // nSub
    std::vector<int32_t> cover[3];            // [g][2]: intervals covering group index g
    void* fusedGroup = nullptr;               // the group a fused batch writes
    rf_blend_box* d_boxes = nullptr;          // fused: the overlap bands
    const rf_blend_sub* subPlain = nullptr;   // the sub table without band arrays
    rf_blend_sub* d_subFused = nullptr;       // fused: the sub table with them
    void* d_bands = nullptr;                  // fused: the band arrays of every field
    int nBoxes = 0;
    int64_t boxPoints = 0;
};

hipError_t rf_launch_blend_dev(const rf_blend_dev& b, void* d_group, int dtype, hipStream_t s);
int rf_moments_parts(int64_t N);
int rf_corr_parts(int64_t N);
hipError_t rf_launch_interp(int nDim, const int32_t* nstep, const int32_t* lo, const double* xMin,
                            const double* xStep, int Nmc, const void* grid, int gridDtype, const double* x,
                            int64_t nNodes, void* out, int outDtype, int* d_flag, hipStream_t s);
hipError_t rf_launch_corr(const void* u, int dtype, int64_t N, int64_t strideA, int nA, int M, double avg,
                          double std, double* d_part, int nPart, double* d_R, hipStream_t s);
hipError_t rf_launch_moments(const void* d, int64_t N, int Nmc, int dtype, double* d_part,
                             int nPart, double* d_sums, hipStream_t s);
// A box window lo .. lo + ext - 1 of an n[0] x n[1] x n[2] grid (x fastest).
struct rf_window {
    int64_t n[3], lo[3], ext[3];
    int64_t count;                // ext[0] * ext[1] * ext[2]
};
hipError_t rf_launch_window_moments(const void* d, int dtype, const rf_window& w, double* d_part, int nPart,
                                    double* d_sums, hipStream_t s);
hipError_t rf_launch_affine(void* d, int64_t N, int Nmc, int dtype, const double* d_shift,
                            const double* d_mul, hipStream_t s);

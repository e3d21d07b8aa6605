// rf_blend.h -- device-side description of a tensor-product arrangement of
// overlapping subdomains, consumed by the gather-form blend kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct rf_blend_dev {
    int nDim, Nmc, mode;
    int groupN[3];
    int nInt[3];                  // distinct subdomain intervals per axis
    int64_t groupTot;
    const int32_t* cover[3];      // [g][2]: intervals covering group index g (-1: none)
    const int32_t* intOrigin[3];  // [interval] first group index
    const int32_t* intSize[3];    // [interval] points
    const int32_t* subOf;         // [i0 + nInt0*(i1 + nInt1*i2)] -> subdomain (-1: none)
    const int64_t* subTot;        // [s] points of subdomain s
    const double* lowF;           // [s][3][maxSize] lower-face taper factor by position
    const double* upF;            // [s][3][maxSize] upper-face taper factor by position
    int maxSize;
    const double* pu1Low[3];      // [g] reverse taper of the group grid (REF_COMPAT)
    const double* pu1Up[3];
    const void* const* fields;    // device array of device pointers
};

hipError_t rf_launch_blend_dev(const rf_blend_dev& b, void* d_group, int dtype, hipStream_t s);
int rf_moments_parts(int64_t N);
hipError_t rf_launch_moments(const void* d, int64_t N, int Nmc, int dtype, double* d_part,
                             int nPart, double* d_sums, hipStream_t s);
hipError_t rf_launch_affine(void* d, int64_t N, int Nmc, int dtype, const double* d_shift,
                            const double* d_mul, hipStream_t s);

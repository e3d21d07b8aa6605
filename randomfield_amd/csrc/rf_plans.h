// rf_plans.h -- host objects shared by the batch and blend code: the blend
// plan (its host tables stay for rf_grid_batch_fuse_blend) and the grid batch.
#pragma once
#include <vector>

#include "rf_blend.h"
#include "rf_internal.h"

struct rf_blend_plan {
    rf_context* ctx = nullptr;
    char* d_blob = nullptr;
    rf_blend_dev b{};
    bool separable = false;
    rf_blend_sep sep{};
    int dtype = RF_DTYPE_F64;
    // host copies for fusing with a grid batch
    int nDim = 0, nSub = 0, mode = 0;
    int32_t groupN[3] = {1, 1, 1};
    std::vector<rf_blend_axis> axH[3];   // the separable per-axis tables
    std::vector<int32_t> org, siz;       // [s][nDim]
    std::vector<int32_t> subInt;         // [s][3] interval of subdomain s per axis
    std::vector<const void*> fields;     // [s]
    void* fusedGroup = nullptr;          // fused: the group buffer the batch writes
};

struct rf_grid_batch {
    rf_context* ctx = nullptr;
    char* d_mem = nullptr;              // scratch of every field, then descriptors
    rf_gemm_desc* d_desc[3] = {nullptr, nullptr, nullptr};
    int nDesc[3] = {0, 0, 0};
    int64_t tiles[3] = {0, 0, 0};
    int realOut[3] = {0, 0, 1};
    std::vector<rf_gemm_desc> desc[3];  // host copies of the descriptors
    std::vector<int32_t> nstep;         // [field][3] grid counts
    int outDtype = RF_DTYPE_F32;
    char* d_fuse = nullptr;             // fused blend: per-field tables (owned)
};

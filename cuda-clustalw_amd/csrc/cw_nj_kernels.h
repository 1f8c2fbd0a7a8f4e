// cw_nj_kernels.h -- device neighbour-joining over a DistMatrix held in device
// memory (NJTree::generateTree, tree/NJTree.cpp:161-751 of
// CPU/clustalw-2.0.11/src).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cwhip.h"

#define CW_NJ_MAX_SEQS CWHIP_NJ_DEVICE_MAX   // per-node arrays in LDS: 52 (n+1) bytes

struct NjArgs {
    const int32_t* pairs;     // staged pair list (si, sj), si < sj
    const int32_t* count;     // identity count per pair (negative: refused, distance 1.0)
    const int32_t* gapfree;   // residues per sequence (codes other than the two gap codes)
    int64_t npairs;
    double* D;                // (n+1)^2 DistMatrix, row-major, symmetric
    int32_t* desc;            // (n+1)^2 treeDesc
    double* left;             // n+2 leftBranch
    double* right;            // n+2 rightBranch
    int32_t* status;          // 0 ok; 1: a step found no pair below its bound
    int32_t n;
    int32_t pad_;
};

// D from the pair counts (when `pairs` is set; else D is already filled), then
// the NJ joins, all on `st`.
hipError_t cw_launch_nj(const NjArgs& a, hipStream_t st);
size_t cw_nj_lds_bytes(int n);

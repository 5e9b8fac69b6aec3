// Host entry points of the device M2L task loop and L2P (internal to the
// library).
#pragma once
#include <hip/hip_runtime.h>

namespace pp2p {

// M2L over the target CSR of a task list: row r adds the translations of
// sources src[row_off[r] .. row_off[r+1]) to L of cell row_tgt[r].
struct M2LArgs {
    const float4 *center;   // cell centres (w unused)
    const float *M;         // 20 multipole coefficients per cell
    const int *row_tgt, *row_off, *src;
    int nrows;
    float rs;               // split radius
    float *L;               // 20 local coefficients per cell, accumulated
};

// L2P: the force of leaf l's local expansion L (about center[l]) at each of
// its particles rng[l], added to the accumulator planes acc[d * npart + i].
struct L2PArgs {
    const int2 *rng;
    const float4 *center;
    const float *L;
    const double *x3;       // FP64 positions, AoS
    int nleaf;
    long long npart;
    double *acc;
};

hipError_t launch_m2l(const M2LArgs &a, hipStream_t st);
hipError_t launch_l2p(const L2PArgs &a, hipStream_t st);

}  // namespace pp2p

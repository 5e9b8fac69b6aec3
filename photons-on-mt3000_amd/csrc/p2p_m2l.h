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

// The device tree as the upward and downward passes read it.  Cells: leaf l
// is cell l, node k is cell nleaf + k.
struct FmmTree {
    int nleaf, nnode;
    const int2 *son;                       // node sons: >= 0 node, < 0 leaf -(id + 1)
    const float4 *node_center, *leaf_center;
    const int2 *rng;                       // leaf particle ranges
    const double *x3;                      // FP64 positions, AoS
};

hipError_t launch_p2m(const FmmTree &t, float mass, float *M, hipStream_t st);
// one level's nodes (idb[s].x = node id): M2M from their sons / L2L into them
hipError_t launch_m2m_level(const FmmTree &t, const int2 *idb, int nseg, float *M, hipStream_t st);
hipError_t launch_l2l_level(const FmmTree &t, const int2 *idb, int nseg, float *L, hipStream_t st);
hipError_t launch_cell_centers(const FmmTree &t, float4 *center, hipStream_t st);
hipError_t launch_zero_f32(float *p, long long n, hipStream_t st);
hipError_t launch_m2l(const M2LArgs &a, hipStream_t st);
hipError_t launch_l2p(const L2PArgs &a, hipStream_t st);

}  // namespace pp2p

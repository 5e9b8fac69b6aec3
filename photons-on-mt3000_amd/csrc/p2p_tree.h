// Host entry points of the device kd-tree build (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

namespace pp2p {

// Per-particle arrays of the build (n entries; sx and small n + 1).
struct TreeScratch {
    int *x[3];      // coordinate planes, permuted in place
    int *idx;       // original index of the particle at each position
    int *segof;     // segment of the current level holding the particle, -1 inside a leaf
    int *small;     // coordinate <= mean of its segment
    int *sx;        // exclusive scan of small
    int *tbl;       // partition partners
};

// One level of the tree as the placement pass reads and writes it.
struct TreeLevelView {
    int nseg;
    const int2 *kid;      // sons: >= 0 next-level segment, < 0 this level's leaf -(q + 1)
    const int2 *leaf;     // this level's leaves (ip, np)
    const float *split;
    const int2 *below;    // (nodes, leaves) below each segment
    int2 *idb;            // (node id, first leaf id) of each segment
    float4 *lo, *hi;      // box of each segment
};

// The finished tree: nodes by id, leaves by id.
struct TreeOut {
    int2 *son;
    float4 *node_center, *node_width;
    float4 *leaf_center, *leaf_width;
    int2 *leaf_rng;
    int *leaf_active;
};

// x planes, idx = 0..n-1 and segof = 0 (one root segment) from n x 3 int32.
hipError_t tree_init(const int *posi, long long n, int *const x[3], int *idx, int *segof, hipStream_t st);
// One level: split values, partition (particles permuted in place), left
// son sizes np0 and pairing bounds kk.  rocPRIM convention: temp == nullptr
// only queries *temp_bytes.
hipError_t tree_split(const TreeScratch &w, const int2 *seg, int nseg, int dir, double int2pos, long long n,
                      double *mean, float *split, int *np0, int *kk, void *temp, size_t *temp_bytes, hipStream_t st);
// Son flags (2 nseg + 1 entries each) and their exclusive scans: noff[2 nseg]
// = next level's segments, loff[2 nseg] = this level's leaves.
hipError_t tree_children(const int2 *seg, int nseg, const int *np0, int maxleaf, int *nodef, int *leaff, int *noff,
                         int *loff, void *temp, size_t *temp_bytes, hipStream_t st);
// Next level's segments, this level's leaves and sons; particles move to the
// segment of their son.
hipError_t tree_next(const TreeScratch &w, const int2 *seg, int nseg, const int *np0, const int *nodef,
                     const int *noff, const int *loff, int2 *next, int2 *leaf, int2 *kid, long long n, hipStream_t st);
hipError_t tree_below(const int2 *kid, int nseg, const int2 *below_next, int2 *below, hipStream_t st);
hipError_t tree_place(const TreeLevelView &lv, const TreeLevelView &nx, int dir, const TreeOut &out, hipStream_t st);
// Target marking of the leaves and the FP64 positions in tree order.
hipError_t tree_finish(const TreeScratch &w, const TreeOut &out, int nleaf, const int *tag, double int2pos,
                       long long n, double *x3, hipStream_t st);

}  // namespace pp2p

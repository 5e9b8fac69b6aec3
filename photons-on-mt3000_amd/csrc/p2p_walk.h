// Host entry points of the device dual-tree walk (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

namespace pp2p {

// The kd-tree as the walk reads it.  son[k] = the two children of node k
// (>= 0: node id, < 0: leaf -(id + 1)); centres and widths of the node and
// leaf boxes (w unused); leaf_active >= 0 marks a leaf holding a local
// particle (a target).
struct WalkTree {
    const int2 *son;
    const float4 *node_center, *node_width;
    const float4 *leaf_center, *leaf_width;
    const int *leaf_active;
    float open_angle, cutoff;
};

// Forward pass, round of n frames f: code[i] = children (0, 2, 4) | task << 3
// | well-separated << 4,
// off[0..n] = exclusive scan of the children (off[n] = next round's size),
// bs[b] = well-separated frames in block b (walk_blocks(n) blocks),
// *emitted = frames of the round that emit a task, *separated = its
// well-separated frames.
// rocPRIM convention: temp == nullptr only queries *temp_bytes.
hipError_t walk_code(const WalkTree &t, const int2 *f, long long n, unsigned char *code, unsigned *off, int *bs,
                     unsigned long long *emitted, unsigned long long *separated, void *temp, size_t *temp_bytes,
                     hipStream_t st);
// Blocks of a round of n frames (the length of walk_code's bs).
long long walk_blocks(long long n);
// The children of round f at their offsets in next.
hipError_t walk_children(const WalkTree &t, const int2 *f, long long n, const unsigned *off, int2 *next,
                         hipStream_t st);
// Backward pass: cnt[i] = tasks at and below frame i (xn: the next round's x,
// null for the last round), x[0..n] = exclusive scan of cnt.  32-bit: the
// caller checks that the list length (the emitted count) fits.
hipError_t walk_below(const unsigned char *code, const unsigned *off, const int *xn, long long n, int *cnt, int *x,
                      void *temp, size_t *temp_bytes, hipStream_t st);
// Position pass: pos[i] = first task position of frame i; fills pos_next for
// the next round and writes the tasks of the emitting frames.
hipError_t walk_emit(const int2 *f, const unsigned char *code, const unsigned *off, const int *xn, long long n,
                     const int *pos, int *pos_next, int *task_t, int *task_s, hipStream_t st);

// The round's well-separated frames (code bit 4), in frame order, written as
// cells (leaf id, or nleaf + node id) to ct/cs: bso = exclusive scan of
// walk_code's bs (walk_blocks(n) entries).  rocPRIM convention for temp.
hipError_t walk_m2l(const int2 *f, const unsigned char *code, long long n, const int *bs, int *bso, int nleaf,
                    int *ct, int *cs, void *temp, size_t *temp_bytes, hipStream_t st);

}  // namespace pp2p

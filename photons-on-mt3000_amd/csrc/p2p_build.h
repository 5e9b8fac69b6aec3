// Host entry points of the build kernels (internal to the library).
#pragma once
#include "p2p_device.h"

namespace pp2p {

// A staged source particle costs a wave about as much as this many pair
// evaluations (its frame shift, cull and compaction, the window walk and the
// dependent loads of a source block).  Measured on the 1 M uneven leaf-pair
// list: 8 -> 0.66 ms, 16 -> 0.62, 24 -> 0.614, 32 -> 0.62, 64 -> 0.63; the
// clustered lists move by less than 0.5%.
constexpr long long kStagePairs = 24;

struct RowArgs {
    const int *row_tgt;   // target leaf of each CSR row
    const int *row_off;   // nrows + 1 offsets into src
    const int *src;       // source leaves
    const int2 *rng;      // leaf (ip, np)
    const unsigned *nsx;  // src_prefix: ntask + 1 entries
    const int *act_cnt;   // per leaf active count (null: all particles)
    const int *act_off;   // per leaf offset into ilist
    int nrows;
    float stage_ops;      // > 0: most i-particles per item (default 64 * ipl_max)
    int ipl_max;          // most i-particles per lane, 1..3
};

hipError_t launch_leaf_frame(const double *x3, const int2 *rng, int nleaf, double inv_rc, bool frame_abs,
                             double4 *ref, float4 *pos, hipStream_t st);
hipError_t launch_i32_to_f64(const int *posi, long long n3, double int2pos, bool round_f32, double *x3,
                             hipStream_t st);
hipError_t launch_round_f32(double *x3, long long n3, hipStream_t st);
hipError_t launch_fill_f64(double *p, long long n, double v, hipStream_t st);
hipError_t launch_f64_to_f32(const double *a, long long n, float *b, hipStream_t st);
hipError_t launch_fill_i32(int *p, long long n, int v, hipStream_t st);

// The functions below follow the rocPRIM convention: temp == nullptr only
// queries the temporary storage size into *temp_bytes.
hipError_t active_lists(const unsigned char *act, int level, const int2 *rng, int nleaf,
                        int *cnt, int *off, int *ilist, void *temp, size_t *temp_bytes, hipStream_t st);
hipError_t csr_sort(const int *key_in, const int *val_in, int *key_out, int *val_out, int n, int key_bits,
                    void *temp, size_t *temp_bytes, hipStream_t st);
hipError_t csr_rows(const int *key_sorted, int n, int *flag, int *rowidx, int *row_tgt, int *row_off,
                    void *temp, size_t *temp_bytes, hipStream_t st);
// olist[ip .. ip + np) = the particles of each leaf in the order the force
// kernel takes them (along the leaf box's longest axis; leaves of more than
// 256 particles unchanged); ocnt / ooff = the leaves' np / ip (nleaf entries).
hipError_t leaf_order(const float4 *pos, const int2 *rng, int nleaf, int *olist, int *ocnt, int *ooff, hipStream_t st);
// nsx[k] = sum of the source sizes rng[src[j]].y over j < k, k = 0..n, modulo
// 2^32 (differences within one CSR row are exact).  rocPRIM convention.
hipError_t src_prefix(const int *src, const int2 *rng, long long n, unsigned *nsx, void *temp, size_t *temp_bytes,
                      hipStream_t st);
// cost[r] = the pairs of row r; total[0] = their sum (the list's pair
// count), total[1] = the source particles staged over all rows (the sum of
// the rows' n_s).  rocPRIM convention.
hipError_t rows_cost(const RowArgs &ra, unsigned long long *cost, unsigned long long *total,
                     void *temp, size_t *temp_bytes, hipStream_t st);
hipError_t items_count(const RowArgs &ra, long long split, int *count, int *off,
                       void *temp, size_t *temp_bytes, hipStream_t st);
hipError_t items_emit(const RowArgs &ra, long long split, const int *off, Item *items, unsigned *cost,
                      hipStream_t st);
hipError_t items_sort(const unsigned *cost, unsigned *cost_sorted, int *idx, int *idx_sorted, const Item *items,
                      Item *items_sorted, int n, void *temp, size_t *temp_bytes, hipStream_t st);

}  // namespace pp2p

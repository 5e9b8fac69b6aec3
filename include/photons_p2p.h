/*
 * photons_p2p.h -- native C-ABI of the MI355X (gfx950) short-range P2P
 * gravity library libphotons_p2p.so.
 *
 * The library evaluates the PhotoNs PM-split short-range pair force
 *
 *     a_i += m * sum_j dx_ij * g(u) / r_eff^3,   u = r_eff / (2 rs),
 *     g(u) = erfc(u) + (2/sqrt(pi)) u exp(-u^2),  zero for u > ucut,
 *
 * over leaf-pair interaction lists (target leaf, source leaf), the work that
 * the reference performs in P2P_kernel / P2P_core
 * (src/gpu_kernel.impl.h:50-157, :364-437) behind task_compute_p2p_gpu
 * (src/gpu_entry.cu:32-209).  The reference-compatible entry points
 * (task_compute_p2p_gpu and friends) are declared in p2p_hip.h and live in
 * the drop-in library libp2p_hip.so, which is a thin layer over this API.
 *
 * Conventions: every function returns a pp2p_status (0 = success) unless
 * stated otherwise; pointers are host pointers unless the name ends in
 * _dev; leaf ids are 0-based; a context is used by one host thread at a
 * time.  All device work of a context runs on one HIP stream, in call order.
 */
#ifndef PHOTONS_P2P_H
#define PHOTONS_P2P_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PP2P_OK = 0,
    PP2P_ERR_ARG = 1,       /* invalid argument (size, range, null pointer) */
    PP2P_ERR_DEVICE = 2,    /* HIP runtime error */
    PP2P_ERR_ALLOC = 3,     /* device or host allocation failed */
    PP2P_ERR_STATE = 4,     /* call out of protocol order */
    PP2P_ERR_NOGPU = 5      /* no HIP device visible */
} pp2p_status;

/* softening: how r_eff is formed from r */
enum {
    PP2P_SOFT_REF = 0,      /* r_eff = max(r, eps)       src/gpu_kernel.impl.h:66-81   */
    PP2P_SOFT_PLUMMER = 1   /* r_eff = sqrt(r^2 + eps^2) src/gpu_kernel.impl.h:274-277 */
};

/* law: how g(u) is evaluated on the device */
enum {
    PP2P_LAW_TABLE = 0,     /* F(s) = g/r^3 as a cubic on 16 or 32 segments per binary
                               octave of s = r^2/rc^2, LDS table, no rsqrt (default) */
    PP2P_LAW_POLY = 1,      /* rsqrt, exp2 and a degree-12 polynomial, no table      */
    PP2P_LAW_LINEAR = 2     /* F(s) = a + b s on 512 segments per binary octave of s,
                               8-byte LDS entries, one workgroup of 16 waves per CU;
                               needs eps / rc >= 2^-12.5 */
};

/* frame: how positions reach the FP32 pair arithmetic */
enum {
    PP2P_FRAME_LEAF = 0,    /* leaf-relative FP32, FP64 per-leaf-pair offsets with
                               optional minimum-image wrap, FP64 fold per source
                               leaf (default) */
    PP2P_FRAME_ABS32 = 1    /* FP32 absolute coordinates and a single FP32
                               accumulator per particle, as the reference kernel
                               (src/gpu_entry.cu:90-96, gpu_kernel.impl.h:405-435);
                               for precision comparisons only */
};

/* execution form of pp2p_compute */
enum {
    PP2P_EXEC_QUEUE = 0,    /* one persistent launch over the cost-sorted queue */
    PP2P_EXEC_GRID = 1,     /* one launch, one wave per work item, list order  */
    PP2P_EXEC_PER_TARGET = 2/* one launch per target leaf (the plain form)     */
};

typedef struct {
    int    softening;   /* PP2P_SOFT_* */
    int    law;         /* PP2P_LAW_* */
    int    frame;       /* PP2P_FRAME_* */
    int    reserved;
    double eps;         /* softening length (SoftenScale) */
    double rs;          /* split radius (splitRadius) */
    double ucut;        /* cutoff on u; 2.8125 in the reference (gpu_kernel.impl.h:86) */
    double mass;        /* particle mass (MASSPART) */
    double box;         /* periodic box length for the minimum-image offsets; 0 = open */
} pp2p_params;

/* kernel tuning: work-item shape; filled by the autotuner or by hand */
typedef struct {
    int waves_per_cu;     /* persistent waves per CU (0 = occupancy maximum) */
    int split_pairs;      /* queue items of about this many pair-interactions
                             (0 = automatic: ~8 items per persistent wave) */
    int segs_per_octave;  /* table law: 16 or 32 segments per octave of s (0 = 32) */
    int no_cull;          /* 1 disables the per-item j-culling (for measurement) */
    int chunk_model;      /* largest i-chunk of a target per work item: <= 0 the
                             kernel maximum of 256 (64 lanes x 4 per lane), > 0
                             at most this many (balanced chunks) */
    int ipl_max;          /* most i-particles per lane an item may use, 1..4
                             (0 = 4); fewer trades idle lanes for registers */
    int reserved[2];
} pp2p_tuning;

typedef struct pp2p_ctx pp2p_ctx;

/* version string of the library */
const char *pp2p_version(void);
/* number of visible HIP devices (0 when none); never fails */
int pp2p_device_count(void);
/* human-readable text for the last error raised on any context of this thread */
const char *pp2p_last_error(void);

pp2p_ctx *pp2p_create(int device, int *status);
void pp2p_destroy(pp2p_ctx *ctx);

int pp2p_set_params(pp2p_ctx *ctx, const pp2p_params *params);
int pp2p_set_tuning(pp2p_ctx *ctx, const pp2p_tuning *tuning);
int pp2p_get_tuning(pp2p_ctx *ctx, pp2p_tuning *tuning);

/*
 * Particles and leaves.  Leaf l owns particles [leaf_ip[l], leaf_ip[l] +
 * leaf_np[l]) (any size, empty included; the ranges of non-empty leaves must
 * be disjoint, as the kd-tree leaves are: PP2P_ERR_ARG otherwise; they need
 * not cover every particle nor come in order).  Positions are FP64 absolute
 * coordinates, npart x 3 (AoS).  Uploads, builds the per-leaf FP64
 * reference points and the leaf-relative FP32 positions on the device and
 * zeroes the acceleration accumulator.
 */
int pp2p_set_particles_f64(pp2p_ctx *ctx, long long npart, const double *pos,
                           int nleaf, const int *leaf_ip, const int *leaf_np);
/*
 * Same for integer grid positions as the reference stores them (Body.posi
 * under INTXYZ, inc/typedef.h:29-46): three int32 at byte offset 0 of each
 * record of `stride` bytes; position = int2pos * posi (src/gpu_entry.cu:94-96).
 */
int pp2p_set_particles_i32(pp2p_ctx *ctx, long long npart, const void *posi,
                           size_t stride, double int2pos,
                           int nleaf, const int *leaf_ip, const int *leaf_np);
/* Per-particle levels (uint8) and the level that receives forces
 * (ACT_OPT semantics, src/gpu_kernel.impl.h:392-396); act == NULL clears. */
int pp2p_set_active(pp2p_ctx *ctx, const unsigned char *act, int active_level);

/*
 * Interaction list: ntask (target, source) leaf pairs.  pp2p_set_tasks
 * replaces the list of the context; pp2p_add_tasks appends (flushes of a
 * streamed list, LEN_TASK in src/fmm.c:500-515).  The list is not modified.
 */
int pp2p_set_tasks(pp2p_ctx *ctx, long long ntask, const int *task_t, const int *task_s);
int pp2p_add_tasks(pp2p_ctx *ctx, long long ntask, const int *task_t, const int *task_s);

/*
 * Interaction list built on the device from a kd-tree: the symmetric dual
 * walk of src/fmm.c:656-792 (acceptance src/fmm.c:411-479, targets marked as
 * src/fmm.c:1226-1236), run level by level.  It replaces the context's list
 * with the walk's (target, source) leaf pairs, 0-based, in the order the
 * reference's recursive walk emits them, so the list equals what the host
 * walk would pass to pp2p_set_tasks.  Tree: nnode nodes, node k with sons
 * node_son[2k], node_son[2k+1] (>= 0: node id, which must be > k; < 0: leaf
 * -(id + 1)), root = node 0 (or the single leaf when nnode == 0); centres and
 * widths of the node and leaf boxes as nnode x 3 / nleaf x 3 FP32
 * (src/fmm.c:273-327); leaf_active[l] >= 0 marks a target leaf.  nleaf is
 * the context's (pp2p_set_particles_*).  *ntask (nullable) receives the list
 * length.  Synchronous.
 */
int pp2p_walk_tasks(pp2p_ctx *ctx, int nnode, const int *node_son, const float *node_center,
                    const float *node_width, const float *leaf_center, const float *leaf_width,
                    const int *leaf_active, float open_angle, float cutoff_radius, long long *ntask);
/*
 * The kd-tree on the device (build_localtree, src/fmm.c:168-327, and the
 * target marking of src/fmm.c:1226-1236), node for node, leaf for leaf and in
 * the particle order the reference's sequential build gives: mean splits
 * cycling x, y, z from the root box bdl..bdr, leaves of at most maxleaf
 * particles.  posi: n x 3 int32 grid positions (position = int2pos * posi);
 * tag: 1 local particle, 0 ghost copy.  Sets the context's particles in tree
 * order and its leaves (as pp2p_set_particles_i32 would with the tree's
 * permutation) and keeps the tree for pp2p_walk_tree.  *nleaf / *nnode
 * (nullable) receive the sizes.  Synchronous.
 */
int pp2p_build_tree(pp2p_ctx *ctx, long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                    const float *bdl, const float *bdr, int *nleaf, int *nnode);
/* pp2p_walk_tasks over the tree of the last pp2p_build_tree. */
int pp2p_walk_tree(pp2p_ctx *ctx, float open_angle, float cutoff_radius, long long *ntask);
/* Copies out the tree of the last pp2p_build_tree (each pointer nullable):
 * perm[k] = input index of the particle at tree position k (npart); leaf
 * ranges, target marks (0 / -1), centres and widths (nleaf x 3); node sons
 * (nnode x 2, >= 0 node, < 0 leaf -(id + 1)), centres and widths (nnode x 3).
 * Synchronous. */
int pp2p_get_tree(pp2p_ctx *ctx, int *perm, int *leaf_ip, int *leaf_np, int *leaf_active, float *leaf_center,
                  float *leaf_width, int *node_son, float *node_center, float *node_width);
/* Copies the context's current list out (each pointer nullable); *ntask
 * (nullable) receives its length.  Synchronous. */
int pp2p_get_tasks(pp2p_ctx *ctx, long long *ntask, int *task_t, int *task_s);

/*
 * The short-range M2L task loop (task_compute_m2l, src/fmm.c:1188-1217) with
 * the erfc-modified kernel of the long/short split (m2l, src/operator.c:
 * 331-474, up to octupole): for each task k, L[task_t[k]] += the local
 * expansion at center[task_t[k]] of the multipole M[task_s[k]] about
 * center[task_s[k]], for split radius rs.  ncell cells (nodes and leaves in
 * one index space); center ncell x 3, M and L ncell x 20 in the reference's
 * coefficient layout (inc/operator.h, NMULTI = 20), FP32.  L is read and
 * written (it accumulates, as the reference's toL += ...).  Separations with
 * r^2 < 1 add nothing, as in the reference (its abs() of an int).
 * *kernel_ms (nullable) receives the kernel's device time.  Synchronous;
 * independent of the context's particles and P2P list.
 */
int pp2p_m2l(pp2p_ctx *ctx, int ncell, const float *center, const float *M, long long ntask, const int *task_t,
             const int *task_s, float rs, float *L, double *kernel_ms);
/* L2P (l2p, src/operator.c:203-262): the force of each leaf's local
 * expansion L (nleaf x 20, about center, nleaf x 3) at the leaf's particles,
 * added to the accumulator (pp2p_get_acc_*), for the context's particles and
 * leaves.  Synchronous. */
int pp2p_l2p(pp2p_ctx *ctx, const float *L, const float *center);

/*
 * The multipole part of the short-range force on the tree of the last
 * pp2p_build_tree, all on the device: P2M of the leaves (src/operator.c:
 * 14-98, particle mass from the parameters), M2M up the tree (walk_m2m,
 * :100-198), the M2L loop over the nm2l well-separated cell pairs (as
 * pp2p_m2l; cells: leaf id, or nleaf + node id, the numbering of the host
 * builder's phost_get_m2l; nm2l = -1 and null arrays: the pairs of the last
 * pp2p_walk_tree), L2L down the tree (walk_l2l, :477-600) and L2P
 * (:203-262), whose forces are added to the accumulator.  P2M, M2M and L2L
 * use the reference's FP32 arithmetic operation for operation.
 * *device_ms (nullable) receives the device time of the passes.  Synchronous.
 */
int pp2p_fmm_short(pp2p_ctx *ctx, long long nm2l, const int *m2l_t, const int *m2l_s, double *device_ms);
/* The well-separated cell pairs of the last device walk (pp2p_walk_tasks or
 * pp2p_walk_tree), round by round, as cells (leaf id, or nleaf + node id);
 * pp2p_fmm_short(ctx, -1, NULL, NULL, ...) uses them directly.  Each pointer
 * nullable.  Synchronous. */
int pp2p_get_m2l_tasks(pp2p_ctx *ctx, long long *n, int *m2l_t, int *m2l_s);
/* The multipoles M and local expansions L (cells x 20 each, nullable) of the
 * last pp2p_fmm_short.  Synchronous. */
int pp2p_get_expansions(pp2p_ctx *ctx, float *M, float *L);

/* Long-range particle-mesh accelerations on one GPU holding the whole
 * periodic mesh: the force part of partmesh_kick (src/partmesh.c:1476-2954)
 * with its convolution (src/conv.f90:168-352).  posi: npart x 3 int32 grid
 * positions (Body.posi of INTXYZ), x = (float)(posi * int2pos); tag: NULL
 * (every particle local) or npart ints, local iff tag == 1 as the reference
 * tests part[n].tag.  box = BOXSIZE, nside = NSIDEMESH (even, 8..1024),
 * split_radius = splitRadius, mass = MASSPART.  Cell-centred CIC deposit,
 * conversion to density (x (nside / box)^3), FFT, Green's function
 * exp(-k^2 rs^2) / k^2 with the CIC deconvolution,
 * inverse FFT, 4-point gradient at the 8 CIC cells.  acc: 3 x npart FP32
 * (SoA, the reference's acc[3][NPART]); zero for non-local particles.  The
 * mesh, the transforms and the gradient are FP64 (the reference: FP32 mesh).
 * *device_ms: deposit to interpolation, on the device.  Synchronous. */
int pp2p_pm_forces(pp2p_ctx *ctx, long long npart, const int *posi, const int *tag, double int2pos,
                   float box, int nside, float split_radius, float mass, float *acc, double *device_ms);
/* The same particle mesh on the particles of the last pp2p_build_tree (its
 * grid positions and tags stay on the device; int2pos is the build's): the
 * local particles' accelerations are ADDED, in FP64, to the accumulator that
 * the P2P force and the multipole pass fill, so that one accumulator holds the
 * whole force of the split.  Synchronous (it waits for its closing timing
 * event); *device_ms as for pp2p_pm_forces.  Has no single counterpart in the
 * reference, whose partmesh_kick reads part[] on the host. */
int pp2p_pm_add_tree(pp2p_ctx *ctx, float box, int nside, float split_radius, float mass, double *device_ms);
/* The last particle-mesh call's mesh: which = 0 the deposited mass per cell
 * (before the conversion to density, x (nside / box)^3, that the convolution applies),
 * 1 the potential; nside^3 doubles, C order [x][y][z]. */
int pp2p_pm_get_mesh(pp2p_ctx *ctx, int which, double *out);

/* Builds the target CSR and the cost-sorted work queue of the current list
 * on the device.  pp2p_compute calls it when the list changed. */
int pp2p_prepare(pp2p_ctx *ctx);

/* Enqueues one evaluation of the current list; accelerations accumulate
 * (+=) into the context's accumulator.  Asynchronous. */
int pp2p_compute(pp2p_ctx *ctx, int exec);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pp2p_compute(ctx, PP2P_EXEC_QUEUE) over a list given
 * as device arrays, accumulating (+=) into a device array, with the list's
 * CSR, work items, sort storage and queue head in the library's static
 * device scratch array (64 MiB per GPU, held until the process ends)
 * instead of the context's buffers.  It calls no hipMalloc, hipFree or
 * allocating library routine.  The particles, law and active mask are the
 * context's (set beforehand).  task_t, task_s: ntask device ints in
 * [0, nleaf); acc: 3 planes of npart device doubles (x, y, z).  A list
 * whose CSR and items need more than the scratch is refused with
 * PP2P_ERR_ARG (about 1.5 M leaf pairs of leaves of up to 64 particles).
 * The call synchronises the context stream while it builds the list, then
 * enqueues the force launch; acc is complete after pp2p_sync(ctx).  Every
 * call on one GPU shares that GPU's scratch, which keeps the last call's
 * contents: calls that overlap on one GPU (other contexts or threads)
 * race. */
int pp2p_compute_noalloc(pp2p_ctx *ctx, long long ntask, const int *task_t, const int *task_s, double *acc);

/* Zeroes the accumulator (asynchronous). */
int pp2p_zero_acc(pp2p_ctx *ctx);

/* Captures nsteps x (zero + compute(exec)) into a HIP graph and replays it
 * nreplay times; the accumulator holds the last step's result. */
int pp2p_run_graph(pp2p_ctx *ctx, int exec, int nsteps, int nreplay);

/* Waits for the context's stream. */
int pp2p_sync(pp2p_ctx *ctx);

/* Accelerations (synchronous copy-out).  f64: npart x 3 AoS; f32: three
 * planes of npart (x | y | z) as dev_part_acc in the reference. */
int pp2p_get_acc_f64(pp2p_ctx *ctx, double *acc);
int pp2p_get_acc_f32_soa(pp2p_ctx *ctx, float *acc);

/* Statistics of the prepared list: number of target rows, work items,
 * sum over tasks of n_t * n_s (src/gpu_entry.cu:50-61; active particles
 * only when an active mask is set). */
int pp2p_list_stats(pp2p_ctx *ctx, long long *ntask, long long *nrows,
                    long long *nitems, long long *npairs);

/* The target CSR of the current list, as the device built it (the reference's
 * thrust sort_by_key + reduce_by_key + exclusive_scan, src/gpu_entry.cu:
 * 132-138): key_sorted and src are task_t and task_s stably sorted by target
 * (ntask each), row_tgt the distinct targets in increasing order (nrows),
 * row_off the exclusive prefix of their task counts with the total last
 * (nrows + 1).  Each pointer nullable; *nrows (nullable) receives the row
 * count.  Builds the CSR when the list changed.  Synchronous. */
int pp2p_get_csr(pp2p_ctx *ctx, long long *nrows, int *key_sorted, int *src, int *row_tgt, int *row_off);

/* Number of listed pair-interactions with r_eff inside the cutoff (u <= ucut),
 * counted on the device by a dedicated pass over the prepared list; the
 * accumulator is untouched.  Synchronous. */
int pp2p_count_inside(pp2p_ctx *ctx, long long *n_inside);

/* Runs nsteps x (zero accumulator + compute(exec)) back to back on the
 * context stream and returns their total device time (HIP events around the
 * region) and, in force_ms (nullable), the summed device time of the force
 * launches alone (events around them in every step).  Synchronous. */
int pp2p_time_steps(pp2p_ctx *ctx, int exec, int nsteps, double *total_ms, double *force_ms);

/* Host evaluation of the device force law, operation for operation as the
 * pair function of params->law computes it in FP32 (PP2P_LAW_TABLE with
 * segs_per_octave segments, or PP2P_LAW_LINEAR): for n values of
 * s = r_eff^2 / rc^2 (already softened, 0 < s), F[k] = g(u) / s^{3/2} in
 * kernel units (rc = 1), or 0 at and beyond the cutoff.  Also returns the
 * table size in entries.  For host-side tests of the tables; no device use. */
int pp2p_law_eval(const pp2p_params *params, int segs_per_octave, long long n, const float *s,
                  float *F, int *table_len);

/* Device time in milliseconds of the force kernel launches of the last
 * pp2p_compute (HIP events on the context stream), and of the last
 * pp2p_prepare. */
double pp2p_last_kernel_ms(pp2p_ctx *ctx);
double pp2p_last_prepare_ms(pp2p_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

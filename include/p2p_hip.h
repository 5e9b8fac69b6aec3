/*
 * p2p_hip.h -- reference-compatible entry points of libp2p_hip.so, the
 * drop-in replacement of the PhotoNs device library (Makefile:149 builds the
 * reference's lib/libp2p_hip.so from src/gpu_entry.cpp; the host links it
 * with -lp2p_hip, Makefile:19,138).
 *
 * Every function keeps the reference's name, signature and call protocol
 * (src/gravity.c:441-447, src/fmm.c:1243-1292):
 *   hostmalloc_task -> task_compute_p2p_gpu(..., 0) x flushes ->
 *   task_compute_p2p_gpu(..., 1) -> get_last_kntime -> hostmalloc_out ->
 *   devMemcpyDH_updateAcc(packlist, NPACK, -1) -> hostfree_out.
 * They read the host program's globals (part, leaf, first_leaf, last_leaf,
 * NPART, packlist, NPACK, SoftenScale, splitRadius, MASSPART, INT2POS, acc,
 * PROC_RANK, PROC_SIZE, task_t, task_s), which the host program defines
 * (src/global.c in the reference).
 *
 * Deviations from the reference: errors are recorded and reported by
 * p2p_hip_status() instead of exit(); the caller's task arrays are not
 * rebased in place; the launch for a step is issued at the last_time call
 * (one persistent launch over all flushes); get_last_kntime returns the
 * device time of that step's force launch.
 *
 * The three record types are those of inc/typedef.h under the reference's
 * default build (INTXYZ, FP32 cpuType, no INTRA_LB); when the reference's
 * typedef.h has been included first (include guard TYPEDEF_H) its
 * definitions are used instead.
 */
#ifndef P2P_HIP_H
#define P2P_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef TYPEDEF_H
/* inc/typedef.h:29-46 (INTXYZ) */
typedef struct {
    int tag;                /* 1 local particle, 0 ghost copy */
    int act;                /* time-step level */
    int posi[3];            /* integer grid position, position = INT2POS * posi */
    float vel[3];
    unsigned long id;
} Body;
/* inc/typedef.h:82-91 (NMULTI 20) */
typedef struct {
    int active;
    int npart;
    int ipart;
    float width[3];
    float center[3];
    float M[20];
    float L[20];
    float acc_pm[3];
} Pack;
/* inc/typedef.h:109-119 */
typedef struct {
    int rankid;
    int ipart;
    int npart;
    int ileaf;
    int nleaf;
    int ntask;
    int *task_t;
    int *task_s;
} P2PPack;
#endif

/* src/gpu_entry.cu:32 -- one flush of (target, source) leaf pairs with
 * absolute leaf ids in [first_leaf, last_leaf); last_time = 1 ends the step. */
void *task_compute_p2p_gpu(int ntask, int *ltask_t, int *ltask_s, int last_time);
/* src/gpu_entry.cu:211 -- all packs of a step at once (INTRA_LB) */
void *p2ppack_compute_gpu(P2PPack *packs, int npack);
/* src/gpu_active.impl.h:1 -- active-level variant; on last_time it also
 * adds the results of particles with act == active_level into acc[] */
void *task_compute_p2p_act_gpu(int ntask, int *ltask_t, int *ltask_s, int last_time, int active_level);
/* src/gpu_active.impl.h:155 */
void *p2ppack_compute_act_gpu(P2PPack *packs, int npack, int active_level);

/* src/gpu_func.impl.h:184-197: host task buffers task_t / task_s */
void hostmalloc_task(int len_task);
void hostrealloc_task(int len_task);
void hostfree_task(void);
/* src/gpu_func.impl.h:259-275: host staging buffers */
void hostmalloc_in(int npart, int nleaf);
void hostfree_in(void);
void hostmalloc_out(int npart);
void hostfree_out(void);
/* src/gpu_func.impl.h:229-257, :396-400: release device bookkeeping / buffers */
void dinfofree(void);
void devfree(void);

/* src/gpu_func.impl.h:644: acc[d][j] += staged result of pack particle j
 * when active_level == -1, or (active_level != 0 and part[j].act == active_level) */
void updateAcc_fromHostBuff(P2PPack pack, int active_level);
/* src/gpu_func.impl.h:655: copy results out and apply updateAcc_fromHostBuff
 * to every pack with ntask > 0 */
void devMemcpyDH_updateAcc(P2PPack *packs, int npack, int active_level);
/* src/gpu_func.impl.h:133 */
void sync_with_dev(void);
/* src/gpu_func.impl.h:715: seconds of the step's force launch */
double get_last_kntime(int proc_rank);

/* Additions without a reference counterpart: status of the last call (0 =
 * success) and its message, for callers that check instead of exiting. */
int p2p_hip_status(void);
const char *p2p_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif

/* Host-side interaction-list producer (kd-tree + dual walk), C-ABI of
 * libphotons_host.so.  Restates src/fmm.c:168-327, :411-479, :656-792 and the
 * target marking of :1226-1236; see kdtree.cpp. */
#ifndef PHOTONS_HOST_H
#define PHOTONS_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

enum { PHOST_NPART = 0, PHOST_NLEAF = 1, PHOST_NNODE = 2, PHOST_NTASK = 3, PHOST_NM2L = 4 };

typedef struct phost_lists phost_lists;

/* posi: n x 3 int32 grid positions (position = int2pos * posi); tag: 1 local
 * particle, 0 ghost copy.  bdl/bdr: root box (particle_combine bounds).
 * Returns NULL on invalid input. */
phost_lists *phost_build(long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                         float open_angle, float cutoff_radius, const float bdl[3], const float bdr[3],
                         int nthreads);
/* The same in two steps: the tree alone (boxes and target marking included,
 * no tasks), then the walk over it (replaces the tasks; 0 on success). */
phost_lists *phost_build_tree(long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                              const float bdl[3], const float bdr[3], int nthreads);
int phost_walk(phost_lists *h, float open_angle, float cutoff_radius, int nthreads);
/* The walk of phost_walk with its list handed out while it runs, in the
 * protocol of turn2compute_p2p (src/fmm.c:500-515): the frames' tasks, in
 * the order of phost_get_tasks, fill buf_t / buf_s (len_task entries each,
 * leaf ids + leaf_base), and flush(n, buf_t, buf_s, 0) is called from the
 * calling thread each time they are full, then flush(rest, buf_t, buf_s, 1)
 * once at the end; the buffers may be refilled as soon as flush returns.
 * flush has the signature of task_compute_p2p_gpu (libp2p_hip.so), which can
 * be passed directly.  overlap = 0 walks every frame before the first flush
 * (for comparison).  The lists stay available as after phost_walk.  0 on
 * success. */
typedef void *(*phost_flush_fn)(int ntask, int *task_t, int *task_s, int last_time);
int phost_walk_stream(phost_lists *h, float open_angle, float cutoff_radius, int nthreads, int len_task,
                      int leaf_base, int *buf_t, int *buf_s, phost_flush_fn flush, int overlap);
long long phost_count(const phost_lists *h, int what);
/* perm[k] = original index of the particle at tree position k */
void phost_get_perm(const phost_lists *h, int *perm);
/* any output pointer may be NULL; center/width are nleaf x 3 */
void phost_get_leaves(const phost_lists *h, int *ip, int *np, int *active, float *center, float *width);
/* any output pointer may be NULL; son is nnode x 2 (>= 0 node id, < 0 leaf
 * -(id + 1)), center/width nnode x 3 */
void phost_get_nodes(const phost_lists *h, int *son, float *center, float *width);
void phost_get_tasks(const phost_lists *h, int *task_t, int *task_s);
/* The well-separated cell pairs the walk met (acceptance 1: the M2L tasks of
 * walk_task_m2l, src/fmm.c:812-900), in walk order; cells are numbered leaf
 * id, or nleaf + node id. */
void phost_get_m2l(const phost_lists *h, int *task_t, int *task_s);
void phost_free(phost_lists *h);

#ifdef __cplusplus
}
#endif
#endif

/*
 * Driver of the reference's own list producer -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/ref.mk, in this container only, together with the
 * reference sources src/fmm.c, src/operator.c, src/memory.c, src/tasklog.c
 * and src/utility.cc compiled where they lie under the reference tree, into
 * oracle/_ref/libreflist.so.  It runs the reference's build_localtree
 * (src/fmm.c:330-406) and task_prepare_p2p / walk_task_p2p
 * (src/fmm.c:1000-1005, :656-792) on given particles and captures the leaf
 * pairs the walk hands to task_compute_p2p_gpu (src/fmm.c:500-515), so that
 * the product's list builder can be checked against them.
 *
 * This file defines the host-program state those reference functions read
 * (the reference defines it in src/global.c, which needs MPI) and the three
 * task-buffer/launch entry points the walk calls.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "photoNs.h"

/* ---- host-program state read by src/fmm.c, src/operator.c, src/memory.c */
Body *part;
Pack *leaf;
Node *btree;
int NPART, MAXLEAF, MAXNLEAF, MAXNNODE;
int first_leaf, last_leaf, first_node, last_node;
int LEN_TASK, PROC_RANK, PROC_RANK_SUDOM;
double INT2POS;
cpuType open_angle, cutoffRadius, splitRadius, MASSPART;
cpuType *acc[3];
int *task_t, *task_s;
double dtime_task, dtime_prep, dtime_m2l, dtime_p2p, dtime_adptv_task;
int idxP2P, idxM2L, m2l_count, walk_m2l_count;
char PathSnapshot[256];

int print_msg(void) { return 0; }

/* ---- task capture (the calls src/fmm.c makes into the P2P library) */
static int *cap_t, *cap_s;
static long long cap_n, cap_cap;

static void cap_append(int n, const int *tt, const int *ts)
{
    if (cap_n + n > cap_cap) {
        long long want = (cap_n + n) * 2 + 1024;
        cap_t = (int *)realloc(cap_t, sizeof(int) * want);
        cap_s = (int *)realloc(cap_s, sizeof(int) * want);
        cap_cap = want;
    }
    memcpy(cap_t + cap_n, tt, sizeof(int) * n);
    memcpy(cap_s + cap_n, ts, sizeof(int) * n);
    cap_n += n;
}

void hostmalloc_task(int len)
{
    task_t = (int *)malloc(sizeof(int) * len);
    task_s = (int *)malloc(sizeof(int) * len);
    for (int k = 0; k < len; ++k)
        task_t[k] = task_s[k] = -1;
}

void hostfree_task(void)
{
    free(task_t);
    free(task_s);
    task_t = task_s = NULL;
}

void *task_compute_p2p_gpu(int ntask, int *tt, int *ts, int last_time)
{
    (void)last_time;
    cap_append(ntask, tt, ts);
    for (int k = 0; k < ntask; ++k)
        tt[k] = ts[k] = -1;
    return NULL;
}

void build_localtree(int direct, cpuType bdleft[3], cpuType bdright[3]);
void task_prepare_p2p(void);

/*
 * posi: n x 3 int32; tag: 1 local / 0 ghost.  Outputs are malloc'ed and
 * owned by the caller (free with reflist_free): perm[n] (tree order ->
 * input index), leaf_ip/leaf_np[nleaf], task_t/task_s[ntask] (0-based leaf
 * ids).  Returns 0 on success.
 */
int reflist_build(long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                  float open_ang, float cutoff, const float bdl[3], const float bdr[3],
                  int **perm, int *nleaf, int **leaf_ip, int **leaf_np, long long *ntask,
                  int **out_t, int **out_s)
{
    NPART = (int)n;
    part = (Body *)calloc((size_t)n, sizeof(Body));
    for (long long i = 0; i < n; ++i) {
        part[i].tag = tag[i];
        part[i].act = 0;
        part[i].posi[0] = posi[3 * i];
        part[i].posi[1] = posi[3 * i + 1];
        part[i].posi[2] = posi[3 * i + 2];
        part[i].id = (unsigned long)i;
    }
    MAXLEAF = maxleaf;
    INT2POS = int2pos;
    open_angle = open_ang;
    cutoffRadius = cutoff;
    cpuType l[3] = {bdl[0], bdl[1], bdl[2]}, r[3] = {bdr[0], bdr[1], bdr[2]};
    build_localtree(0, l, r);

    /* target marking of fmm_task (src/fmm.c:1226-1236) */
    for (int k = first_leaf; k < last_leaf; ++k) {
        leaf[k].active = -1;
        for (int i = 0; i < leaf[k].npart; ++i)
            if (part[i + leaf[k].ipart].tag == 1) {
                leaf[k].active = 0;
                break;
            }
    }
    LEN_TASK = 1 << 22;
    cap_n = 0;
    hostmalloc_task(LEN_TASK);
    idxP2P = 0;
    task_prepare_p2p();
    int rest = 0;
    while (rest < LEN_TASK && task_t[rest] != -1)
        ++rest;
    cap_append(rest, task_t, task_s);
    hostfree_task();

    *nleaf = last_leaf - first_leaf;
    *perm = (int *)malloc(sizeof(int) * n);
    for (long long i = 0; i < n; ++i)
        (*perm)[i] = (int)part[i].id;
    *leaf_ip = (int *)malloc(sizeof(int) * (*nleaf + 1));
    *leaf_np = (int *)malloc(sizeof(int) * (*nleaf + 1));
    for (int k = 0; k < *nleaf; ++k) {
        (*leaf_ip)[k] = leaf[k + first_leaf].ipart;
        (*leaf_np)[k] = leaf[k + first_leaf].npart;
    }
    for (long long k = 0; k < cap_n; ++k) {
        cap_t[k] -= first_leaf;
        cap_s[k] -= first_leaf;
    }
    *ntask = cap_n;
    *out_t = cap_t;
    *out_s = cap_s;
    cap_t = cap_s = NULL;
    cap_n = cap_cap = 0;
    free(part);
    part = NULL;
    return 0;
}

void reflist_free(void *p) { free(p); }

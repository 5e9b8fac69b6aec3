/*
 * The reference's own force step linked against the drop-in library --
 * TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/ref.mk, in this container only, into oracle/_ref/link_host:
 * the reference's src/fmm.c, src/operator.c, src/memory.c, src/tasklog.c and
 * src/utility.cc, compiled where they lie under the reference tree, plus this
 * file, linked with -lp2p_hip against photons-on-mt3000_amd/lib/libp2p_hip.so.
 * Unlike ref_harness.c it defines no P2P entry point: every call the
 * reference makes into its device library (hostmalloc_task, the flushes of
 * task_compute_p2p_gpu from turn2compute_p2p and fmm_task, hostfree_task,
 * src/fmm.c:500-515, :1221-1300) resolves to the drop-in, unchanged.  This
 * file only defines the host-program state (src/global.c in the reference,
 * which needs MPI), sets up the pack list as p2ppack_init does
 * (src/gravity.c:189-192), runs fmm_prepare + fmm_task and collects the
 * results as src/gravity.c:441-447 does.
 *
 * Usage: link_host IN.bin OUT.bin
 *   IN.bin:  int64 n, int32 maxleaf, int32 len_task, float32 open_angle,
 *            cutoff_radius, soften_scale, split_radius, mass, float64
 *            int2pos, float32 bdl[3], bdr[3], int32 posi[n][3], int32 tag[n]
 *   OUT.bin: int32 status (p2p_hip_status), int32 nleaf, float64 kernel
 *            seconds (get_last_kntime), int32 perm[n] (tree order -> input
 *            index), int32 leaf_ip[nleaf], leaf_np[nleaf], float32 acc[3][n]
 *            (the P2P accelerations added by devMemcpyDH_updateAcc)
 * Exit status 0 when the step ran to its end (whatever the library's
 * status), 2 on bad usage or input.
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
/* ---- and by the device library (src/gpu_func.impl.h, src/gpu_entry.cu) */
int PROC_SIZE, NEW_RANK, NPACK, MAXNPACK;
P2PPack *packlist;
cpuType SoftenScale;

int print_msg(void) { return 0; }

/* the drop-in's entry points, as the reference's host code calls them */
void hostmalloc_out(int npart);
void hostfree_out(void);
void devMemcpyDH_updateAcc(P2PPack *packs, int npack, int active_level);
double get_last_kntime(int proc_rank);
void devfree(void);
int p2p_hip_status(void);

void fmm_prepare(int direct, cpuType bdl[3], cpuType bdr[3]);
void fmm_task(void);

static int rd(FILE *f, void *p, size_t sz, size_t n)
{
    return fread(p, sz, n, f) == n ? 0 : -1;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: link_host IN.bin OUT.bin\n");
        return 2;
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in)
        return 2;
    long long n;
    int maxleaf, len_task;
    float fp[5], bdl[3], bdr[3];
    double int2pos;
    if (rd(in, &n, 8, 1) || rd(in, &maxleaf, 4, 1) || rd(in, &len_task, 4, 1) || rd(in, fp, 4, 5) ||
        rd(in, &int2pos, 8, 1) || rd(in, bdl, 4, 3) || rd(in, bdr, 4, 3) || n < 1 || n > (1ll << 30) ||
        maxleaf < 1 || len_task < 1) {
        fclose(in);
        return 2;
    }
    int *posi = malloc(sizeof(int) * 3 * (size_t)n), *tag = malloc(sizeof(int) * (size_t)n);
    if (!posi || !tag || rd(in, posi, 4, 3 * (size_t)n) || rd(in, tag, 4, (size_t)n)) {
        fclose(in);
        return 2;
    }
    fclose(in);

    NPART = (int)n;
    part = calloc((size_t)n, sizeof(Body));
    for (long long i = 0; i < n; ++i) {
        part[i].tag = tag[i];
        part[i].act = 0;
        part[i].posi[0] = posi[3 * i];
        part[i].posi[1] = posi[3 * i + 1];
        part[i].posi[2] = posi[3 * i + 2];
        part[i].id = (unsigned long)i;
    }
    MAXLEAF = maxleaf;
    LEN_TASK = len_task;
    open_angle = fp[0];
    cutoffRadius = fp[1];
    SoftenScale = fp[2];
    splitRadius = fp[3];
    MASSPART = fp[4];
    INT2POS = int2pos;
    PROC_RANK = 0;
    PROC_SIZE = 1;
    for (int d = 0; d < 3; ++d)
        acc[d] = calloc((size_t)n, sizeof(cpuType));
    /* p2ppack_init, src/gravity.c:189-192 (packLevel 0) */
    NPACK = 0;
    MAXNPACK = 2;
    packlist = malloc(sizeof(P2PPack) * MAXNPACK);

    cpuType l[3] = {bdl[0], bdl[1], bdl[2]}, r[3] = {bdr[0], bdr[1], bdr[2]};
    fmm_prepare(0, l, r);
    fmm_task();
    /* src/gravity.c:441-447 */
    const double kt = get_last_kntime(PROC_RANK);
    hostmalloc_out(NPART);
    devMemcpyDH_updateAcc(packlist, NPACK, -1);
    hostfree_out();
    const int status = p2p_hip_status();

    FILE *out = fopen(argv[2], "wb");
    if (!out)
        return 2;
    const int nleaf = last_leaf - first_leaf;
    fwrite(&status, 4, 1, out);
    fwrite(&nleaf, 4, 1, out);
    fwrite(&kt, 8, 1, out);
    for (long long i = 0; i < n; ++i) {
        const int id = (int)part[i].id;
        fwrite(&id, 4, 1, out);
    }
    for (int k = 0; k < nleaf; ++k)
        fwrite(&leaf[first_leaf + k].ipart, 4, 1, out);
    for (int k = 0; k < nleaf; ++k)
        fwrite(&leaf[first_leaf + k].npart, 4, 1, out);
    for (int d = 0; d < 3; ++d)
        fwrite(acc[d], sizeof(cpuType), (size_t)n, out);
    fclose(out);
    devfree();
    fprintf(stderr, "link_host: N=%d leaves=%d status %d kernel %.3f ms\n", NPART, nleaf, status, kt * 1e3);
    return 0;
}

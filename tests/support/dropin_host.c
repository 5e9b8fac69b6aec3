/* A C host of the drop-in library, compiled against include/p2p_hip.h and
 * linked with -lp2p_hip the way the reference host links its own build
 * (Makefile:19,138).  It fills the host-program globals (host_globals.c),
 * builds a small clustered problem on the reference's integer grid, and runs
 * one step through the reference's call protocol: hostmalloc_task, flushes of
 * LEN_TASK tasks with absolute leaf ids and last_time = 0, the final flush
 * with last_time = 1 (src/fmm.c:500-515, :1243-1292), then get_last_kntime,
 * hostmalloc_out, devMemcpyDH_updateAcc, hostfree_out (src/gravity.c:441-447)
 * and the teardown (src/photoNs.c:383-384).
 *
 * Usage: dropin_host OUT.bin.  OUT.bin receives, little-endian:
 *   int32 N, int32 nleaf, int32 ntask, float64 int2pos, float32 eps, rs, mass,
 *   int32 posi[N][3], int32 leaf_ip[nleaf], leaf_np[nleaf],
 *   int32 task_t[ntask], task_s[ntask] (0-based leaf ids),
 *   float32 acc[3][N].
 * Exit status: 0 on success, else the library's p2p_hip_status(). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/p2p_hip.h"

extern Body *part;
extern Pack *leaf;
extern int first_leaf, last_leaf, NPART, NPACK, MAXNPACK, MAXLEAF;
extern P2PPack *packlist;
extern float SoftenScale, MASSPART, splitRadius;
extern double INT2POS;
extern float *acc[3];
extern int *task_t, *task_s;

enum { NSIDE = 6, LEN_TASK = 1000 };

static unsigned long long rng_state = 88172645463325252ull;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int check(const char *what)
{
    const int st = p2p_hip_status();
    if (st != 0)
        fprintf(stderr, "dropin_host: %s: status %d: %s\n", what, st, p2p_hip_error());
    return st;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: dropin_host OUT.bin\n");
        return 2;
    }
    /* leaves on a NSIDE^3 lattice of spacing 2, 1 to 40 particles each, in a
     * box of side 2 NSIDE (rs = 1: cutoff radius 2 rs 2.8125 = 5.625) */
    const double box = 2.0 * NSIDE;
    const int nleaf = NSIDE * NSIDE * NSIDE;
    int *lip = malloc(sizeof(int) * nleaf), *lnp = malloc(sizeof(int) * nleaf);
    int n = 0;
    for (int l = 0; l < nleaf; ++l) {
        lip[l] = n;
        lnp[l] = 1 + (int)(urand() * 40.0);
        n += lnp[l];
    }
    INT2POS = (box + 1.1 * 4.5) / (double)(1 << 30);
    Body *b = calloc((size_t)n, sizeof(Body));
    for (int l = 0; l < nleaf; ++l) {
        const double c[3] = {2.0 * (l % NSIDE) + 1.0, 2.0 * ((l / NSIDE) % NSIDE) + 1.0, 2.0 * (l / (NSIDE * NSIDE)) + 1.0};
        for (int k = 0; k < lnp[l]; ++k)
            for (int d = 0; d < 3; ++d)
                b[lip[l] + k].posi[d] = (int)lround((c[d] + (urand() - 0.5) * 1.8) / INT2POS);
    }
    /* leaf pairs whose centres are within 6 of each other, both ways (no wrap) */
    int ntask = 0, cap = nleaf * nleaf;
    int *tt = malloc(sizeof(int) * cap), *ts = malloc(sizeof(int) * cap);
    for (int a = 0; a < nleaf; ++a)
        for (int c = 0; c < nleaf; ++c) {
            const int dx = a % NSIDE - c % NSIDE, dy = (a / NSIDE) % NSIDE - (c / NSIDE) % NSIDE,
                      dz = a / (NSIDE * NSIDE) - c / (NSIDE * NSIDE);
            if (4 * (dx * dx + dy * dy + dz * dz) <= 36) {
                tt[ntask] = a;
                ts[ntask] = c;
                ++ntask;
            }
        }

    /* host-program state, as src/global.c and src/fmm.c:393 set it up */
    Pack *leaves = calloc((size_t)nleaf, sizeof(Pack));
    for (int l = 0; l < nleaf; ++l) {
        leaves[l].ipart = lip[l];
        leaves[l].npart = lnp[l];
    }
    part = b;
    NPART = n;
    first_leaf = n;
    last_leaf = n + nleaf;
    leaf = leaves - first_leaf;
    NPACK = 0;
    MAXNPACK = 1;
    packlist = calloc(1, sizeof(P2PPack));
    SoftenScale = 0.05f;
    splitRadius = 1.0f;
    MASSPART = 0.5f;
    for (int d = 0; d < 3; ++d)
        acc[d] = calloc((size_t)n, sizeof(float));

    /* one step through the reference protocol */
    hostmalloc_task(LEN_TASK);
    int k = 0;
    while (ntask - k > LEN_TASK) {
        for (int j = 0; j < LEN_TASK; ++j) {
            task_t[j] = tt[k + j] + first_leaf;
            task_s[j] = ts[k + j] + first_leaf;
        }
        task_compute_p2p_gpu(LEN_TASK, task_t, task_s, 0);
        if (check("task_compute_p2p_gpu"))
            return p2p_hip_status();
        k += LEN_TASK;
    }
    for (int j = 0; j < ntask - k; ++j) {
        task_t[j] = tt[k + j] + first_leaf;
        task_s[j] = ts[k + j] + first_leaf;
    }
    task_compute_p2p_gpu(ntask - k, task_t, task_s, 1);
    if (check("task_compute_p2p_gpu(last_time)"))
        return p2p_hip_status();
    hostfree_task();
    const double kt = get_last_kntime(0);
    hostmalloc_out(NPART);
    devMemcpyDH_updateAcc(packlist, NPACK, -1);
    hostfree_out();
    if (check("devMemcpyDH_updateAcc"))
        return p2p_hip_status();
    dinfofree();
    devfree();

    FILE *f = fopen(argv[1], "wb");
    if (!f)
        return 3;
    const float law[3] = {SoftenScale, splitRadius, MASSPART};
    fwrite(&n, 4, 1, f);
    fwrite(&nleaf, 4, 1, f);
    fwrite(&ntask, 4, 1, f);
    fwrite(&INT2POS, 8, 1, f);
    fwrite(law, 4, 3, f);
    for (int i = 0; i < n; ++i)
        fwrite(b[i].posi, 4, 3, f);
    fwrite(lip, 4, (size_t)nleaf, f);
    fwrite(lnp, 4, (size_t)nleaf, f);
    fwrite(tt, 4, (size_t)ntask, f);
    fwrite(ts, 4, (size_t)ntask, f);
    for (int d = 0; d < 3; ++d)
        fwrite(acc[d], 4, (size_t)n, f);
    fclose(f);
    printf("dropin_host: N=%d leaves=%d tasks=%d flushes=%d kernel %.3f ms\n", n, nleaf, ntask,
           ntask / LEN_TASK + 1, kt * 1e3);
    return 0;
}

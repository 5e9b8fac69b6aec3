/*
 * CPU oracle for the PhotoNs short-range P2P path.  TEST INFRASTRUCTURE ONLY:
 * the product library never links or calls this file.  See oracle.h for the
 * contents and the reference lines each function restates.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, so that every
 * FP32 operation of the reference restatement rounds as written).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NTAB 512                 /* NTABINT, src/gpu_kernel.impl.h:1 */
static const double TWO_OVER_SQRTPI = 1.1283791670955126;   /* 2 * PIisq */

double oracle_g(double u)
{
    return erfc(u) + TWO_OVER_SQRTPI * u * exp(-u * u);
}

/* Table prologue of P2P_kernel (src/gpu_kernel.impl.h:374-383):
 * x = i * 3/512, tab0 = erfc(x) + coeff x e^{-x^2}, tab1 = -2 coeff e^{-x^2},
 * coeff = 2 PIisq (src/gpu_entry.cu:80).  The reference evaluates these with
 * device erfc/exp in FP32; here host libm in FP64, rounded once to FP32. */
void oracle_ref_tables(float *tab0, float *tab1)
{
    for (int i = 0; i < NTAB; ++i) {
        double x = (double)i * (3.0 / NTAB);
        tab0[i] = (float)oracle_g(x);
        tab1[i] = (float)(-2.0 * TWO_OVER_SQRTPI * exp(-x * x));
    }
}

/* One pair of P2P_core (REF softening, src/gpu_kernel.impl.h:50-157) or
 * P2P_core_1 (Plummer, :260-362) with SPI (table) + 4th order Taylor +
 * GPUCUTOFF, all in FP32.  `mass` is MASSPART, multiplied per pair as in
 * P2P_core (:92).  rrs = 1/splitRadius. */
void oracle_pair_ref32(int softening, float eps, float mass, float rrs,
                       const float *tab0, const float *tab1,
                       const float xi[3], const float xj[3], float acc[3])
{
    const float half = 0.5f;
    float dx = xj[0] - xi[0];
    float dy = xj[1] - xi[1];
    float dz = xj[2] - xi[2];
    float dr, rdr;
    if (softening == ORACLE_SOFT_PLUMMER) {
        float dr2 = dx * dx + dy * dy + dz * dz + eps * eps;
        rdr = 1.0f / sqrtf(dr2);
        dr = rdr * dr2;
    } else {
        float dr2 = dx * dx + dy * dy + dz * dz;
        if (dr2 > eps * eps) {
            rdr = 1.0f / sqrtf(dr2);       /* hardware rsq on the device */
            dr = rdr * dr2;
        } else {
            dr = eps;
            rdr = (float)(1.0 / (double)eps);
        }
    }
    float drs = half * dr * rrs;
    if (drs > 2.8125f)                     /* GPUCUTOFF */
        return;
    float ir3 = mass * (rdr * rdr * rdr);

    const float widinv = (float)(NTAB / 3.0);
    const float width = (float)(3.0 / NTAB);
    /* the truncated local constants of P2P_core (:103-107) */
    const float one_3 = 0.333333f, five_3 = 1.66666667f, two_3 = 0.6666667f;
    const float one = 1.0f, three_2 = 1.5f;

    int idx = (int)(drs * widinv);
    float x0 = idx * width;
    float e = drs - x0;
    if (e > half * width) {                /* nearest node */
        idx++;
        x0 = idx * width;
        e = drs - x0;
    }
    if (idx >= NTAB || idx < 0)
        return;
    float t0 = tab0[idx], t1 = tab1[idx];

    float tdx = e;
    float x2 = x0 * x0;
    float a1 = x2 * tdx;
    tdx *= e;
    float a2 = x0 * (one - x2) * tdx;
    tdx *= e;
    float a3 = (one_3 - five_3 * x2 + two_3 * x2 * x2) * tdx;
    tdx *= e;
    float a4 = x0 * (three_2 * x2 - one - one_3 * x2 * x2) * tdx;
    ir3 *= (t0 + (a1 + a2 + a3 + a4) * t1);

    acc[0] += dx * ir3;
    acc[1] += dy * ir3;
    acc[2] += dz * ir3;
}

static inline int particle_active(const unsigned char *act, int level, long i)
{
    return act == NULL || act[i] == (unsigned char)level;
}

/* FP64 pair law shared by the direct sum and the pair counter.  Returns the
 * scalar f with a_i += m f dx, and sets *inside when u_eff <= ucut. */
static inline double pair_f64(const oracle_params *p, double r2, int *inside)
{
    double reff;
    if (p->softening == ORACLE_SOFT_PLUMMER)
        reff = sqrt(r2 + p->eps * p->eps);
    else
        reff = r2 > p->eps * p->eps ? sqrt(r2) : p->eps;
    double u = 0.5 * reff / p->rs;
    if (u > p->ucut) {
        *inside = 0;
        return 0.0;
    }
    *inside = 1;
    return oracle_g(u) / (reff * reff * reff);
}

static inline double min_image(double d, double box)
{
    if (box > 0.0)
        d -= box * nearbyint(d / box);
    return d;
}

void oracle_p2p_f64(const oracle_params *p, const double *pos,
                    const int *leaf_ip, const int *leaf_np,
                    int ntgt, const int *tgt_leaf, const long long *tgt_off,
                    const int *src_leaf, const unsigned char *act,
                    int active_level, double *acc, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int t = 0; t < ntgt; ++t) {
        int tl = tgt_leaf[t];
        int ip = leaf_ip[tl], np = leaf_np[tl];
        for (int a = 0; a < np; ++a) {
            long i = (long)ip + a;
            if (!particle_active(act, active_level, i))
                continue;
            double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (long long k = tgt_off[t]; k < tgt_off[t + 1]; ++k) {
                int sl = src_leaf[k];
                int jp = leaf_ip[sl], jn = leaf_np[sl];
                for (int b = 0; b < jn; ++b) {
                    long j = (long)jp + b;
                    double dx = min_image(pos[3 * j] - xi, p->box);
                    double dy = min_image(pos[3 * j + 1] - yi, p->box);
                    double dz = min_image(pos[3 * j + 2] - zi, p->box);
                    int inside;
                    double f = pair_f64(p, dx * dx + dy * dy + dz * dz, &inside);
                    ax += dx * f;
                    ay += dy * f;
                    az += dz * f;
                }
            }
            acc[3 * i] += p->mass * ax;
            acc[3 * i + 1] += p->mass * ay;
            acc[3 * i + 2] += p->mass * az;
        }
    }
}

void oracle_count_pairs(const oracle_params *p, const double *pos,
                        const int *leaf_ip, const int *leaf_np,
                        int ntgt, const int *tgt_leaf, const long long *tgt_off,
                        const int *src_leaf, const unsigned char *act,
                        int active_level, long long *n_list, long long *n_in,
                        int nthreads)
{
    long long tot = 0, in = 0;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot, in) num_threads(nthreads)
#endif
    for (int t = 0; t < ntgt; ++t) {
        int tl = tgt_leaf[t];
        int ip = leaf_ip[tl], np = leaf_np[tl];
        for (int a = 0; a < np; ++a) {
            long i = (long)ip + a;
            if (!particle_active(act, active_level, i))
                continue;
            for (long long k = tgt_off[t]; k < tgt_off[t + 1]; ++k) {
                int sl = src_leaf[k];
                int jp = leaf_ip[sl], jn = leaf_np[sl];
                tot += jn;
                for (int b = 0; b < jn; ++b) {
                    long j = (long)jp + b;
                    double dx = min_image(pos[3 * j] - pos[3 * i], p->box);
                    double dy = min_image(pos[3 * j + 1] - pos[3 * i + 1], p->box);
                    double dz = min_image(pos[3 * j + 2] - pos[3 * i + 2], p->box);
                    int inside;
                    (void)pair_f64(p, dx * dx + dy * dy + dz * dz, &inside);
                    in += inside;
                }
            }
        }
    }
    *n_list = tot;
    *n_in = in;
}

/* P2P_kernel loop order (src/gpu_kernel.impl.h:388-436): target leaf ->
 * i-particle -> source leaf (CSR order) -> j-particle; FP32 accumulator per
 * i, added once into the output (:431-435). */
void oracle_p2p_ref32(const oracle_params *p, const float *pos_soa, int npart,
                      const int *leaf_ip, const int *leaf_np,
                      int ntgt, const int *tgt_leaf, const long long *tgt_off,
                      const int *src_leaf, const unsigned char *act,
                      int active_level, float *acc_soa, int nthreads)
{
    float tab0[NTAB], tab1[NTAB];
    oracle_ref_tables(tab0, tab1);
    const float eps = (float)p->eps;
    const float mass = (float)p->mass;
    const float rrs = (float)(1.0 / p->rs);
    const float *X = pos_soa, *Y = pos_soa + npart, *Z = pos_soa + 2 * (long)npart;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int t = 0; t < ntgt; ++t) {
        int tl = tgt_leaf[t];
        int ip = leaf_ip[tl], np = leaf_np[tl];
        for (int a = 0; a < np; ++a) {
            long i = (long)ip + a;
            if (!particle_active(act, active_level, i))
                continue;
            float xi[3] = {X[i], Y[i], Z[i]};
            float ac[3] = {0.0f, 0.0f, 0.0f};
            for (long long k = tgt_off[t]; k < tgt_off[t + 1]; ++k) {
                int sl = src_leaf[k];
                int jp = leaf_ip[sl], jn = leaf_np[sl];
                for (int b = 0; b < jn; ++b) {
                    long j = (long)jp + b;
                    float xj[3] = {X[j], Y[j], Z[j]};
                    oracle_pair_ref32(p->softening, eps, mass, rrs, tab0, tab1, xi, xj, ac);
                }
            }
            acc_soa[i] += ac[0];
            acc_soa[i + npart] += ac[1];
            acc_soa[i + 2 * (long)npart] += ac[2];
        }
    }
}

int oracle_build_csr(long long ntask, const int *task_t, const int *task_s,
                     int nleaf, int *tgt_leaf, long long *tgt_off, int *src_leaf)
{
    long long *count = (long long *)calloc((size_t)nleaf + 1, sizeof(long long));
    if (!count)
        return -1;
    for (long long k = 0; k < ntask; ++k) {
        if (task_t[k] < 0 || task_t[k] >= nleaf || task_s[k] < 0 || task_s[k] >= nleaf) {
            free(count);
            return -1;
        }
        count[task_t[k] + 1]++;
    }
    /* counting sort keeps the input order within a target: a stable sort */
    for (int l = 0; l < nleaf; ++l)
        count[l + 1] += count[l];
    long long *cursor = (long long *)malloc(sizeof(long long) * ((size_t)nleaf + 1));
    if (!cursor) {
        free(count);
        return -1;
    }
    memcpy(cursor, count, sizeof(long long) * ((size_t)nleaf + 1));
    for (long long k = 0; k < ntask; ++k)
        src_leaf[cursor[task_t[k]]++] = task_s[k];
    int ntgt = 0;
    for (int l = 0; l < nleaf; ++l) {
        if (count[l + 1] > count[l]) {
            tgt_leaf[ntgt] = l;
            tgt_off[ntgt] = count[l];
            ntgt++;
        }
    }
    tgt_off[ntgt] = ntask;
    free(cursor);
    free(count);
    return ntgt;
}

/*
 * CPU oracle for the PhotoNs short-range particle-to-particle (P2P) path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; it is the checker used by tests/, by __graft_entry__.smoke() and
 * by the cpu_baseline leg of bench.py.
 *
 * Contents
 *   oracle_g            closed form g(u) = erfc(u) + (2/sqrt(pi)) u exp(-u^2)
 *   oracle_ref_tables   the 512-node tables of the reference kernel prologue
 *                       (src/gpu_kernel.impl.h:374-383), evaluated with host
 *                       libm in FP64 and rounded to FP32
 *   oracle_pair_ref32   one pair of the reference pair functions P2P_core
 *                       (src/gpu_kernel.impl.h:50-157, REF softening) and
 *                       P2P_core_1 (:260-362, Plummer softening) in FP32
 *   oracle_p2p_f64      FP64 direct sum over a target-CSR interaction list
 *                       (same lists, same force law; exact g, no table)
 *   oracle_p2p_ref32    FP32 restatement of the reference leaf-pair kernel
 *                       P2P_kernel (src/gpu_kernel.impl.h:364-437) driven by
 *                       the same CSR; OpenMP over target leaves.  This is the
 *                       "reference CPU P2P" baseline: the reference ships no
 *                       CPU P2P (src/fmm.c:484-491 is a stub).
 *   oracle_build_csr    stable sort of (task_t, task_s) by target and run
 *                       length encoding, as thrust::sort_by_key +
 *                       reduce_by_key + exclusive_scan do in
 *                       src/gpu_entry.cu:132-138.
 *
 * Parity pins (tests/test_oracle.py): closed-form values of g, the identity
 * g'(x) = x^2 tab1(x), Taylor coefficients against finite differences of g,
 * brute-force pair sums on tiny inputs, momentum conservation on symmetric
 * lists, translation invariance, Newtonian limit r << rs, cutoff edges,
 * self-pair zero.  The CSR build is pinned against an independent Python
 * stable sort.
 */
#ifndef PHOTONS_P2P_ORACLE_H
#define PHOTONS_P2P_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* softening modes */
#define ORACLE_SOFT_REF     0   /* r_eff = max(r, eps)      src/gpu_kernel.impl.h:66-81 */
#define ORACLE_SOFT_PLUMMER 1   /* r_eff = sqrt(r^2+eps^2)  src/gpu_kernel.impl.h:274-277 */

typedef struct {
    int    softening;   /* ORACLE_SOFT_* */
    int    pad_;
    double eps;         /* SoftenScale (src/setparam.c:23) */
    double rs;          /* splitRadius (src/setparam.c:29) */
    double ucut;        /* cutoff on u = r_eff/(2 rs); 2.8125 in the reference */
    double mass;        /* MASSPART */
    double box;         /* periodic box for minimum image; 0 = open */
} oracle_params;

double oracle_g(double u);

/* tab0[i] = g(x_i), tab1[i] = -2 (2/sqrt(pi)) exp(-x_i^2), x_i = 3 i / 512 */
void oracle_ref_tables(float *tab0, float *tab1);

/* acc (3 floats, in/out) += the FP32 reference pair contribution of j on i */
void oracle_pair_ref32(int softening, float eps, float mass, float rrs,
                       const float *tab0, const float *tab1,
                       const float xi[3], const float xj[3], float acc[3]);

/*
 * pos: N x 3 doubles (AoS).  leaf_ip/leaf_np: particle range of each leaf
 * (leaf ids are 0-based).  CSR: tgt_leaf[ntgt], tgt_off[ntgt+1], src_leaf[].
 * act (nullable): per particle level; with act != NULL only particles with
 * act[i] == active_level receive forces (ACT_OPT semantics,
 * src/gpu_kernel.impl.h:392-396).  acc: N x 3 doubles, accumulated (+=).
 * nthreads <= 0: OpenMP default.
 */
void oracle_p2p_f64(const oracle_params *p, const double *pos,
                    const int *leaf_ip, const int *leaf_np,
                    int ntgt, const int *tgt_leaf, const long long *tgt_off,
                    const int *src_leaf, const unsigned char *act,
                    int active_level, double *acc, int nthreads);

/*
 * FP32 restatement of P2P_kernel.  pos_soa: 3 planes of npart floats
 * (x | y | z) as h_part_pos in src/gpu_entry.cu:85-105.  acc_soa: 3 planes,
 * accumulated (+=) as dev_part_acc (src/gpu_kernel.impl.h:431-435).
 */
void oracle_p2p_ref32(const oracle_params *p, const float *pos_soa, int npart,
                      const int *leaf_ip, const int *leaf_np,
                      int ntgt, const int *tgt_leaf, const long long *tgt_off,
                      const int *src_leaf, const unsigned char *act,
                      int active_level, float *acc_soa, int nthreads);

/*
 * Stable CSR of a task list.  Outputs: tgt_leaf[nleaf] (first ntgt used),
 * tgt_off[nleaf+1], src_leaf[ntask].  Returns ntgt, or -1 on a task id
 * outside [0, nleaf).
 */
int oracle_build_csr(long long ntask, const int *task_t, const int *task_s,
                     int nleaf, int *tgt_leaf, long long *tgt_off, int *src_leaf);

/* sum over tasks of n_t * n_s (src/gpu_entry.cu:50-61) and the number of
 * listed pairs with u_eff <= ucut (FP64), for FLOP accounting. */
void oracle_count_pairs(const oracle_params *p, const double *pos,
                        const int *leaf_ip, const int *leaf_np,
                        int ntgt, const int *tgt_leaf, const long long *tgt_off,
                        const int *src_leaf, const unsigned char *act,
                        int active_level, long long *n_list, long long *n_in,
                        int nthreads);

/*
 * Short-range multipole operators (oracle_multipole.c): the reference's M2L
 * (src/operator.c:331-474, LONGSHORT, up to octupole) and L2P (:203-262) in
 * its FP32 arithmetic, and an independent FP64 form of M2L, L2P and P2M.
 * Coefficient layout of inc/operator.h, ORACLE_NMULTI = 20.
 */
#define ORACLE_NMULTI 20
void oracle_m2l_ref32(float dx, float dy, float dz, const float *M, float rs, float *L);
void oracle_l2p_ref32(int npart, const int *posi, double int2pos, const float *center, const float *L,
                      float *acc_x, float *acc_y, float *acc_z);
void oracle_radial_f64(double r, double rs, double *f);
void oracle_m2l_f64(double dx, double dy, double dz, const double *M, double rs, double *L);
void oracle_l2p_f64(double dx, double dy, double dz, const double *L, double *F);
void oracle_p2m_f64(int n, const double *pos, double mass, const double *center, double *M);
void oracle_m2l_tasks_f64(int ncell, const double *center, const double *M, long long ntask, const int *task_t,
                          const int *task_s, double rs, double *L);
void oracle_p2m_ref32(int npart, const int *posi, double int2pos, const float *center, float mass, float *M);
void oracle_m2m_ref32(float dx, float dy, float dz, const float *M, float *tM);
void oracle_l2l_ref32(float dx, float dy, float dz, const float *L, float *toL);
void oracle_m2m_f64(double sx, double sy, double sz, const double *M, double *tM);
void oracle_l2l_f64(double sx, double sy, double sz, const double *L, double *toL);
void oracle_m2l_tasks_ref32(int ncell, const float *center, const float *M, long long ntask, const int *task_t,
                            const int *task_s, float rs, float *L);

#ifdef __cplusplus
}
#endif
#endif

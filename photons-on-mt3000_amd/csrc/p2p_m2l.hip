// The short-range multipole part of the force on the device: the M2L task
// loop (task_compute_m2l, src/fmm.c:1188-1217) with the erfc-modified
// kernel of the long/short split (m2l, src/operator.c:331-474, LONGSHORT,
// up to octupole) and the L2P evaluation of the leaves' local expansions
// (l2p, src/operator.c:203-262).
//
// M2L: the task list is sorted by target cell (the CSR of the P2P path), and
// one thread owns a target: it adds the translations of all its sources in
// list order, each evaluated in FP32 as the reference does (derivative
// tensors and contractions), summed in FP64 registers and added to the
// target's L once.  The radial factors f0..f3 are evaluated in FP64 and
// rounded, as the reference's mixed double/float expressions are close to.
// Separations with r^2 < 1 add nothing: the reference's test is abs() of the
// int-converted r^2.  Coefficient layout of inc/operator.h (20 per cell).
#include <cstring>

#include "p2p_m2l.h"

#pragma clang fp contract(off)

namespace pp2p {
namespace {

constexpr int kM2LBlock = 256;
enum { X = 1, Y, Z, XX, XY, XZ, YY, YZ, ZZ, XXX, XXY, XXZ, XYY, XYZ, XZZ, YYY, YYZ, YZZ, ZZZ };

// Adds the M2L translation of multipole M over (dx, dy, dz) = target centre
// - source centre to acc.
__device__ __forceinline__ void m2l_add(float dx, float dy, float dz, const float *M, double rs, double *acc)
{
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 >= 1.0f))
        return;
    // radial factors ((1/r) d/dr)^k of erfc(r / 2 rs) / r
    const double r = sqrt((double)r2);
    const double ir = 1.0 / r, ir2 = ir * ir;
    const double E = erfc(0.5 * r / rs);
    const double q = exp(-0.25 * r2 / (rs * rs)) * (0.5641895835477563 / rs);   // 1/sqrt(pi) / rs
    const double is2 = 1.0 / (rs * rs);
    const float f0 = (float)(E * ir);
    const float f1 = (float)(-(E * ir + q) * ir2);
    const float f2 = (float)((3.0 * E * ir + 3.0 * q) * ir2 * ir2 + 0.5 * q * is2 * ir2);
    const float f3 = (float)(-(15.0 * E * ir + 15.0 * q) * ir2 * ir2 * ir2 - 2.5 * q * is2 * ir2 * ir2 -
                             0.25 * q * is2 * is2 * ir2);

    float D[20];
    D[0] = f0;
    D[X] = f1 * dx;
    D[Y] = f1 * dy;
    D[Z] = f1 * dz;
    D[XX] = f2 * dx * dx + f1;
    D[XY] = f2 * dx * dy;
    D[XZ] = f2 * dx * dz;
    D[YY] = f2 * dy * dy + f1;
    D[YZ] = f2 * dy * dz;
    D[ZZ] = f2 * dz * dz + f1;
    D[XXX] = f3 * dx * dx * dx + f2 * dx * 3.0f;
    D[XXY] = f3 * dx * dx * dy + f2 * dy;
    D[XXZ] = f3 * dx * dx * dz + f2 * dz;
    D[XYY] = f3 * dx * dy * dy + f2 * dx;
    D[XYZ] = f3 * dx * dy * dz;
    D[XZZ] = f3 * dx * dz * dz + f2 * dx;
    D[YYY] = f3 * dy * dy * dy + f2 * dy * 3.0f;
    D[YYZ] = f3 * dy * dy * dz + f2 * dz;
    D[YZZ] = f3 * dy * dz * dz + f2 * dy;
    D[ZZZ] = f3 * dz * dz * dz + f2 * dz * 3.0f;

    // L[b] += sum over a, |a| + |b| <= 3, of M[a] D[a + b]
    acc[0] += M[0] * D[0] + M[X] * D[X] + M[Y] * D[Y] + M[Z] * D[Z] + M[XX] * D[XX] + M[XY] * D[XY] +
              M[XZ] * D[XZ] + M[YY] * D[YY] + M[YZ] * D[YZ] + M[ZZ] * D[ZZ] + M[XXX] * D[XXX] + M[XXY] * D[XXY] +
              M[XXZ] * D[XXZ] + M[XYY] * D[XYY] + M[XYZ] * D[XYZ] + M[YYY] * D[YYY] + M[YYZ] * D[YYZ] +
              M[YZZ] * D[YZZ] + M[XZZ] * D[XZZ] + M[ZZZ] * D[ZZZ];
    acc[X] += M[0] * D[X] + M[X] * D[XX] + M[Y] * D[XY] + M[Z] * D[XZ] + M[XX] * D[XXX] + M[XY] * D[XXY] +
              M[XZ] * D[XXZ] + M[YY] * D[XYY] + M[YZ] * D[XYZ] + M[ZZ] * D[XZZ];
    acc[Y] += M[0] * D[Y] + M[X] * D[XY] + M[Y] * D[YY] + M[Z] * D[YZ] + M[YY] * D[YYY] + M[YZ] * D[YYZ] +
              M[XY] * D[XYY] + M[ZZ] * D[YZZ] + M[XZ] * D[XYZ] + M[XX] * D[XXY];
    acc[Z] += M[0] * D[Z] + M[X] * D[XZ] + M[Y] * D[YZ] + M[Z] * D[ZZ] + M[ZZ] * D[ZZZ] + M[XZ] * D[XZZ] +
              M[YZ] * D[YZZ] + M[XX] * D[XXZ] + M[XY] * D[XYZ] + M[YY] * D[YYZ];
    acc[XX] += M[0] * D[XX] + M[X] * D[XXX] + M[Y] * D[XXY] + M[Z] * D[XXZ];
    acc[XY] += M[0] * D[XY] + M[X] * D[XXY] + M[Y] * D[XYY] + M[Z] * D[XYZ];
    acc[XZ] += M[0] * D[XZ] + M[X] * D[XXZ] + M[Y] * D[XYZ] + M[Z] * D[XZZ];
    acc[YY] += M[0] * D[YY] + M[X] * D[XYY] + M[Y] * D[YYY] + M[Z] * D[YYZ];
    acc[YZ] += M[0] * D[YZ] + M[X] * D[XYZ] + M[Y] * D[YYZ] + M[Z] * D[YZZ];
    acc[ZZ] += M[0] * D[ZZ] + M[X] * D[XZZ] + M[Y] * D[YZZ] + M[Z] * D[ZZZ];
#pragma unroll
    for (int k = XXX; k <= ZZZ; ++k)
        acc[k] += M[0] * D[k];
}

__global__ __launch_bounds__(kM2LBlock) void k_m2l_rows(M2LArgs a)
{
    const int row = blockIdx.x * kM2LBlock + threadIdx.x;
    if (row >= a.nrows)
        return;
    const int t = a.row_tgt[row];
    const float4 ct = a.center[t];
    double acc[20];
#pragma unroll
    for (int k = 0; k < 20; ++k)
        acc[k] = 0.0;
    for (int j = a.row_off[row]; j < a.row_off[row + 1]; ++j) {
        const int s = a.src[j];
        const float4 cs = a.center[s];
        float M[20];
        const float4 *m4 = reinterpret_cast<const float4 *>(a.M + 20 * (size_t)s);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const float4 v = m4[k];
            M[4 * k] = v.x;
            M[4 * k + 1] = v.y;
            M[4 * k + 2] = v.z;
            M[4 * k + 3] = v.w;
        }
        m2l_add(ct.x - cs.x, ct.y - cs.y, ct.z - cs.z, M, (double)a.rs, acc);
    }
    float *L = a.L + 20 * (size_t)t;
#pragma unroll
    for (int k = 0; k < 20; ++k)
        L[k] = (float)((double)L[k] + acc[k]);
}

// L2P: one wave per leaf, a lane per particle; the force of the leaf's local
// expansion at the particle, in FP32 as the reference, added to the FP64
// accumulator planes.
__global__ __launch_bounds__(kM2LBlock) void k_l2p(L2PArgs a)
{
    const int leaf = blockIdx.x * (kM2LBlock / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (leaf >= a.nleaf)
        return;
    const int2 r = a.rng[leaf];
    const float4 c = a.center[leaf];
    const float *L = a.L + 20 * (size_t)leaf;
    for (int k = lane; k < r.y; k += 64) {
        const long long i = (long long)r.x + k;
        const float dx = (float)a.x3[3 * i] - c.x;
        const float dy = (float)a.x3[3 * i + 1] - c.y;
        const float dz = (float)a.x3[3 * i + 2] - c.z;
        float fx = L[X], fy = L[Y], fz = L[Z];
        fx += L[XX] * dx + L[XY] * dy + L[XZ] * dz;
        fy += L[XY] * dx + L[YY] * dy + L[YZ] * dz;
        fz += L[XZ] * dx + L[YZ] * dy + L[ZZ] * dz;
        fx += L[XXX] * dx * dx * 0.5f + L[XXY] * dx * dy + L[XXZ] * dx * dz + L[XYZ] * dy * dz +
              L[XYY] * dy * dy * 0.5f + L[XZZ] * dz * dz * 0.5f;
        fy += L[YYY] * dy * dy * 0.5f + L[YYZ] * dy * dz + L[XYY] * dx * dy + L[XYZ] * dx * dz +
              L[YZZ] * dz * dz * 0.5f + L[XXY] * dx * dx * 0.5f;
        fz += L[ZZZ] * dz * dz * 0.5f + L[XZZ] * dx * dz + L[YZZ] * dy * dz + L[XYZ] * dx * dy +
              L[XXZ] * dx * dx * 0.5f + L[YYZ] * dy * dy * 0.5f;
        atomicAdd(&a.acc[i], (double)fx);
        atomicAdd(&a.acc[a.npart + i], (double)fy);
        atomicAdd(&a.acc[2 * a.npart + i], (double)fz);
    }
}

// ----------------------------------------------------------- upward/downward
// P2M (src/operator.c:14-98), M2M (:100-165, walk_m2m :171-198) and L2L
// (:477-580, walk_l2l :582-600) in the reference's FP32 arithmetic, operation
// for operation (no contraction), so that the multipoles and the shifted
// local expansions equal the reference's.  Cells: leaf l is cell l, node k
// is cell nleaf + k.

__device__ void p2m_cell(const double *x3, int2 r, float4 c, float mass, float *M)
{
    for (int n = 0; n < 20; ++n)
        M[n] = 0.0f;
    for (int k = 0; k < r.y; ++k) {
        const long long i = (long long)r.x + k;
        const float d0 = (float)x3[3 * i] - c.x;
        const float d1 = (float)x3[3 * i + 1] - c.y;
        const float d2 = (float)x3[3 * i + 2] - c.z;
        const float m0 = mass, m1 = -m0, m2 = m0, m3 = -m0;
        M[0] += m0;
        M[X] += m1 * d0;
        M[Y] += m1 * d1;
        M[Z] += m1 * d2;
        M[XX] += m2 * d0 * d0 / 2;
        M[XY] += m2 * d0 * d1;
        M[XZ] += m2 * d0 * d2;
        M[YY] += m2 * d1 * d1 / 2;
        M[YZ] += m2 * d1 * d2;
        M[ZZ] += m2 * d2 * d2 / 2;
        M[XXX] += m3 * d0 * d0 * d0 / 6;
        M[XXY] += m3 * d1 * d0 * d0 / 2;
        M[XXZ] += m3 * d2 * d0 * d0 / 2;
        M[XYY] += m3 * d1 * d1 * d0 / 2;
        M[XYZ] += m3 * d2 * d1 * d0;
        M[XZZ] += m3 * d2 * d2 * d0 / 2;
        M[YYY] += m3 * d1 * d1 * d1 / 6;
        M[YYZ] += m3 * d1 * d2 * d1 / 2;
        M[YZZ] += m3 * d2 * d2 * d1 / 2;
        M[ZZZ] += m3 * d2 * d2 * d2 / 6;
    }
}

__device__ void m2m_add(float dx, float dy, float dz, const float *M, float *tM)
{
    float t[20];
    t[0] = M[0];
    t[X] = M[X] + M[0] * dx;
    t[Y] = M[Y] + M[0] * dy;
    t[Z] = M[Z] + M[0] * dz;
    t[XX] = M[XX] + M[X] * dx + M[0] * dx * dx / 2;
    t[XY] = M[XY] + M[X] * dy + M[Y] * dx + M[0] * dx * dy;
    t[XZ] = M[XZ] + M[X] * dz + M[Z] * dx + M[0] * dx * dz;
    t[YY] = M[YY] + M[Y] * dy + M[0] * dy * dy / 2;
    t[YZ] = M[YZ] + M[Y] * dz + M[Z] * dy + M[0] * dy * dz;
    t[ZZ] = M[ZZ] + M[Z] * dz + M[0] * dz * dz / 2;
    t[XXX] = M[XXX] + M[XX] * dx + M[X] * dx * dx / 2 + M[0] * dx * dx * dx / 6;
    t[YYY] = M[YYY] + M[YY] * dy + M[Y] * dy * dy / 2 + M[0] * dy * dy * dy / 6;
    t[ZZZ] = M[ZZZ] + M[ZZ] * dz + M[Z] * dz * dz / 2 + M[0] * dz * dz * dz / 6;
    t[XXY] = M[XXY] + M[XX] * dy + M[XY] * dx + M[X] * dx * dy + M[Y] * dx * dx / 2 + M[0] * dx * dx * dy / 2;
    t[XXZ] = M[XXZ] + M[XX] * dz + M[XZ] * dx + M[X] * dx * dz + M[Z] * dx * dx / 2 + M[0] * dx * dx * dz / 2;
    t[XYY] = M[XYY] + M[YY] * dx + M[XY] * dy + M[Y] * dx * dy + M[X] * dy * dy / 2 + M[0] * dx * dy * dy / 2;
    t[XZZ] = M[XZZ] + M[ZZ] * dx + M[XZ] * dz + M[Z] * dx * dz + M[X] * dz * dz / 2 + M[0] * dx * dz * dz / 2;
    t[YZZ] = M[YZZ] + M[ZZ] * dy + M[YZ] * dz + M[Z] * dy * dz + M[Y] * dz * dz / 2 + M[0] * dy * dz * dz / 2;
    t[YYZ] = M[YYZ] + M[YY] * dz + M[YZ] * dy + M[Y] * dy * dz + M[Z] * dy * dy / 2 + M[0] * dy * dy * dz / 2;
    t[XYZ] = M[XYZ] + M[XY] * dz + M[YZ] * dx + M[XZ] * dy + M[X] * dy * dz + M[Y] * dx * dz + M[Z] * dy * dx +
             M[0] * dx * dy * dz;
    for (int n = 0; n < 20; ++n)
        tM[n] += t[n];
}

__device__ void l2l_add(float dx, float dy, float dz, const float *L, float *toL)
{
    toL[0] += L[0] + L[X] * dx + L[Y] * dy + L[Z] * dz;
    toL[X] += L[X];
    toL[Y] += L[Y];
    toL[Z] += L[Z];
    toL[0] += L[XX] * dx * dx / 2 + L[XY] * dx * dy + L[XZ] * dx * dz + L[YZ] * dy * dz + L[YY] * dy * dy / 2 +
              L[ZZ] * dz * dz / 2;
    toL[X] += L[XX] * dx + L[XY] * dy + L[XZ] * dz;
    toL[Y] += L[XY] * dx + L[YY] * dy + L[YZ] * dz;
    toL[Z] += L[XZ] * dx + L[YZ] * dy + L[ZZ] * dz;
    for (int n = XX; n <= ZZ; ++n)
        toL[n] += L[n];
    toL[0] += L[XXX] * dx * dx * dx / 6 + L[XXY] * dx * dx * dy / 2 + L[XXZ] * dx * dx * dz / 2 +
              L[XYY] * dx * dy * dy / 2 + L[XYZ] * dx * dy * dz + L[XZZ] * dx * dz * dz / 2 +
              L[YYY] * dy * dy * dy / 6 + L[YZZ] * dy * dz * dz / 2 + L[YYZ] * dy * dy * dz / 2 +
              L[ZZZ] * dz * dz * dz / 6;
    // the reference's 0.5 literals: these sums are FP64, rounded once
    toL[X] = (float)((double)toL[X] + ((double)(L[XXX] * dx * dx) * 0.5 + (double)(L[XXY] * dx * dy) +
                                       (double)(L[XXZ] * dx * dz) + (double)(L[XYY] * dy * dy) * 0.5 +
                                       (double)(L[XYZ] * dy * dz) + (double)(L[XZZ] * dz * dz) * 0.5));
    toL[Y] = (float)((double)toL[Y] + ((double)(L[XXY] * dx * dx) * 0.5 + (double)(L[XYY] * dx * dy) +
                                       (double)(L[XYZ] * dx * dz) + (double)(L[YYY] * dy * dy) * 0.5 +
                                       (double)(L[YYZ] * dy * dz) + (double)(L[YZZ] * dz * dz) * 0.5));
    toL[Z] = (float)((double)toL[Z] + ((double)(L[XXZ] * dx * dx) * 0.5 + (double)(L[XYZ] * dx * dy) +
                                       (double)(L[XZZ] * dx * dz) + (double)(L[YYZ] * dy * dy) * 0.5 +
                                       (double)(L[YZZ] * dy * dz) + (double)(L[ZZZ] * dz * dz) * 0.5));
    toL[XX] += L[XXX] * dx + L[XXY] * dy + L[XXZ] * dz;
    toL[XY] += L[XXY] * dx + L[XYY] * dy + L[XYZ] * dz;
    toL[XZ] += L[XXZ] * dx + L[XYZ] * dy + L[XZZ] * dz;
    toL[YY] += L[XYY] * dx + L[YYY] * dy + L[YYZ] * dz;
    toL[YZ] += L[XYZ] * dx + L[YYZ] * dy + L[YZZ] * dz;
    toL[ZZ] += L[XZZ] * dx + L[YZZ] * dy + L[ZZZ] * dz;
    for (int n = XXX; n <= ZZZ; ++n)
        toL[n] += L[n];
}

__global__ __launch_bounds__(kM2LBlock) void k_p2m(FmmTree t, float mass, float *M)
{
    const int l = blockIdx.x * kM2LBlock + threadIdx.x;
    if (l >= t.nleaf)
        return;
    float m[20];
    p2m_cell(t.x3, t.rng[l], t.leaf_center[l], mass, m);
    for (int n = 0; n < 20; ++n)
        M[20 * (size_t)l + n] = m[n];
}

__device__ __forceinline__ float4 cell_center(const FmmTree &t, int son)
{
    return son >= 0 ? t.node_center[son] : t.leaf_center[-son - 1];
}

__device__ __forceinline__ size_t cell_of(const FmmTree &t, int son)
{
    return son >= 0 ? (size_t)t.nleaf + son : (size_t)(-son - 1);
}

// M2M of one level's nodes: M[node] = sum over its two sons, son 0 first
__global__ __launch_bounds__(kM2LBlock) void k_m2m_level(FmmTree t, const int2 *idb, int nseg, float *M)
{
    const int s = blockIdx.x * kM2LBlock + threadIdx.x;
    if (s >= nseg)
        return;
    const int k = idb[s].x;
    const float4 c = t.node_center[k];
    const int2 son = t.son[k];
    float acc[20];
    for (int n = 0; n < 20; ++n)
        acc[n] = 0.0f;
    for (int h = 0; h < 2; ++h) {
        const int sn = h ? son.y : son.x;
        const float4 cs = cell_center(t, sn);
        float m[20];
        const float *src = M + 20 * cell_of(t, sn);
        for (int n = 0; n < 20; ++n)
            m[n] = src[n];
        m2m_add(c.x - cs.x, c.y - cs.y, c.z - cs.z, m, acc);
    }
    float *dst = M + 20 * ((size_t)t.nleaf + k);
    for (int n = 0; n < 20; ++n)
        dst[n] = acc[n];
}

// L2L of one level's nodes into their sons
__global__ __launch_bounds__(kM2LBlock) void k_l2l_level(FmmTree t, const int2 *idb, int nseg, float *L)
{
    const int s = blockIdx.x * kM2LBlock + threadIdx.x;
    if (s >= nseg)
        return;
    const int k = idb[s].x;
    const float4 c = t.node_center[k];
    const int2 son = t.son[k];
    float lp[20];
    const float *src = L + 20 * ((size_t)t.nleaf + k);
    for (int n = 0; n < 20; ++n)
        lp[n] = src[n];
    for (int h = 0; h < 2; ++h) {
        const int sn = h ? son.y : son.x;
        const float4 cs = cell_center(t, sn);
        float *dst = L + 20 * cell_of(t, sn);
        float ls[20];
        for (int n = 0; n < 20; ++n)
            ls[n] = dst[n];
        l2l_add(cs.x - c.x, cs.y - c.y, cs.z - c.z, lp, ls);
        for (int n = 0; n < 20; ++n)
            dst[n] = ls[n];
    }
}

__global__ __launch_bounds__(kM2LBlock) void k_cell_centers(FmmTree t, float4 *center)
{
    const int c = blockIdx.x * kM2LBlock + threadIdx.x;
    if (c >= t.nleaf + t.nnode)
        return;
    center[c] = c < t.nleaf ? t.leaf_center[c] : t.node_center[c - t.nleaf];
}

__global__ __launch_bounds__(kM2LBlock) void k_fill_zero(float *p, long long n)
{
    const long long i = (long long)blockIdx.x * kM2LBlock + threadIdx.x;
    if (i < n)
        p[i] = 0.0f;
}

}  // namespace

hipError_t launch_m2l(const M2LArgs &a, hipStream_t st)
{
    if (a.nrows > 0)
        k_m2l_rows<<<(unsigned)((a.nrows + kM2LBlock - 1) / kM2LBlock), kM2LBlock, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_l2p(const L2PArgs &a, hipStream_t st)
{
    if (a.nleaf > 0)
        k_l2p<<<(unsigned)((a.nleaf + kM2LBlock / 64 - 1) / (kM2LBlock / 64)), kM2LBlock, 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace pp2p

namespace pp2p {

static unsigned blocks(long long n) { return (unsigned)((n + kM2LBlock - 1) / kM2LBlock); }

hipError_t launch_p2m(const FmmTree &t, float mass, float *M, hipStream_t st)
{
    if (t.nleaf > 0)
        k_p2m<<<blocks(t.nleaf), kM2LBlock, 0, st>>>(t, mass, M);
    return hipGetLastError();
}

hipError_t launch_m2m_level(const FmmTree &t, const int2 *idb, int nseg, float *M, hipStream_t st)
{
    if (nseg > 0)
        k_m2m_level<<<blocks(nseg), kM2LBlock, 0, st>>>(t, idb, nseg, M);
    return hipGetLastError();
}

hipError_t launch_l2l_level(const FmmTree &t, const int2 *idb, int nseg, float *L, hipStream_t st)
{
    if (nseg > 0)
        k_l2l_level<<<blocks(nseg), kM2LBlock, 0, st>>>(t, idb, nseg, L);
    return hipGetLastError();
}

hipError_t launch_cell_centers(const FmmTree &t, float4 *center, hipStream_t st)
{
    k_cell_centers<<<blocks((long long)t.nleaf + t.nnode), kM2LBlock, 0, st>>>(t, center);
    return hipGetLastError();
}

hipError_t launch_zero_f32(float *p, long long n, hipStream_t st)
{
    if (n > 0)
        k_fill_zero<<<blocks(n), kM2LBlock, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace pp2p

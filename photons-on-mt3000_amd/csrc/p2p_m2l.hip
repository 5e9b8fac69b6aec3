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

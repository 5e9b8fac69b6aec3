// Particle-mesh long-range force (partmesh_kick, src/partmesh.c:1476-2954, and
// its convolution, src/conv.f90:168-352) with the whole periodic mesh on one GPU.
//
// * CIC weights in the reference's FP32 arithmetic: posf = (float)(posi *
//   INT2POS); i = (int)(posf * norm); w = (posf - (i + 0.5) delta) norm with the
//   neighbour on the side of w (src/partmesh.c:1569-1612).
// * The deposit adds the FP32 products MASSPART * wx * wy * wz into an FP64
//   mesh with atomics, first summed per chunk of particles in LDS (the
//   reference sums in FP32 in particle order).
// * The Green's function multiplies the FP64 r2c spectrum in place; it
//   includes the reference's conversion of mass per cell to density.
// * The acceleration: 4-point central differences of the potential at the 8
//   CIC cells, weighted by the CIC weights, in FP64; stored as FP32 like the
//   reference's acc arrays.
// Cell indices wrap modulo N where the reference reads its ghost layers.
#include "p2p_pm.h"

#pragma clang fp contract(off)

namespace pp2p {
namespace {

constexpr int kPmBlock = 256;
// conv.f90's M_PI = 3.1415926, a REAL(cputype) constant
constexpr double kPiRef = (double)3.1415926f;

struct Axis {
    int i, ii;
    float w, wn;     // weight of the neighbour ii, of the own cell i
};

__device__ __forceinline__ Axis cic_axis(float posf, float norm, float delta)
{
    Axis a;
    a.i = (int)(posf * norm);
    float w = (float)(((double)posf - ((double)a.i + 0.5) * (double)delta) * (double)norm);
    if (w > 0.0f) {
        a.ii = a.i + 1;
    } else {
        w = -w;
        a.ii = a.i - 1;
    }
    a.w = w;
    a.wn = (float)(1.0 - (double)w);
    return a;
}

__device__ __forceinline__ int wrap(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ void particle_axes(const PmArgs &a, long long p, Axis ax[3])
{
    const float norm = (float)a.nside / a.box;
    const float delta = (float)(1.0 / (double)norm);
    const int *pc[3] = {a.px, a.py, a.pz};
    for (int d = 0; d < 3; ++d) {
        const float posf = (float)((double)pc[d][p * a.stride] * a.int2pos);
        ax[d] = cic_axis(posf, norm, delta);
    }
}

// One block deposits a chunk of kPmChunk consecutive particles.  Particles
// that arrive in spatial order (tree order) cover few cells per chunk: the
// chunk's cells (unwrapped indices, both CIC neighbours) are bounded first,
// and when the box fits in kPmTile cells the chunk sums into LDS and adds each
// touched cell to the mesh once; otherwise every corner goes to the mesh.
constexpr int kPmChunk = 1024;
constexpr int kPmTile = 4096;

__device__ __forceinline__ bool is_local(const PmArgs &a, long long p)
{
    return !a.tag || a.tag[a.tag_idx ? a.tag_idx[p] : p] == 1;
}

__device__ __forceinline__ bool local_particle(const PmArgs &a, long long p)
{
    return p < a.n && is_local(a, p);
}

__global__ __launch_bounds__(kPmBlock) void k_pm_deposit(PmArgs a, double *__restrict__ rho)
{
    __shared__ double tile[kPmTile];
    __shared__ int lo[3], hi[3];
    const int N = a.nside;
    const long long p0 = (long long)blockIdx.x * kPmChunk;
    if (threadIdx.x < 3) {
        lo[threadIdx.x] = 0x7fffffff;
        hi[threadIdx.x] = -0x7fffffff;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kPmChunk; k += kPmBlock) {
        const long long p = p0 + k;
        if (!local_particle(a, p))
            continue;
        Axis ax[3];
        particle_axes(a, p, ax);
        for (int d = 0; d < 3; ++d) {
            atomicMin(&lo[d], min(ax[d].i, ax[d].ii));
            atomicMax(&hi[d], max(ax[d].i, ax[d].ii));
        }
    }
    __syncthreads();
    if (hi[0] < lo[0])
        return;                                   // no local particle in the chunk
    const int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
    const bool fits = (long long)sx * sy * sz <= kPmTile;
    if (fits) {
        for (int c = threadIdx.x; c < kPmTile; c += kPmBlock)
            tile[c] = 0.0;
        __syncthreads();
    }
    for (int k = threadIdx.x; k < kPmChunk; k += kPmBlock) {
        const long long p = p0 + k;
        if (!local_particle(a, p))
            continue;
        Axis ax[3];
        particle_axes(a, p, ax);
        const int cx[2] = {ax[0].i, ax[0].ii}, cy[2] = {ax[1].i, ax[1].ii}, cz[2] = {ax[2].i, ax[2].ii};
        const float wx[2] = {ax[0].wn, ax[0].w}, wy[2] = {ax[1].wn, ax[1].w}, wz[2] = {ax[2].wn, ax[2].w};
        for (int c = 0; c < 8; ++c) {
            const int x = c & 1, y = (c >> 1) & 1, z = c >> 2;
            const float m = ((a.mass * wx[x]) * wy[y]) * wz[z];
            if (fits)
                atomicAdd(&tile[((cx[x] - lo[0]) * sy + (cy[y] - lo[1])) * sz + (cz[z] - lo[2])], (double)m);
            else
                atomicAdd(&rho[((long long)wrap(cx[x], N) * N + wrap(cy[y], N)) * N + wrap(cz[z], N)], (double)m);
        }
    }
    if (!fits)
        return;
    __syncthreads();
    const int ncell = sx * sy * sz;
    for (int c = threadIdx.x; c < ncell; c += kPmBlock) {
        const double v = tile[c];
        if (v == 0.0)
            continue;
        const int x = c / (sy * sz), y = (c / sz) % sy, z = c % sz;
        atomicAdd(&rho[((long long)wrap(lo[0] + x, N) * N + wrap(lo[1] + y, N)) * N + wrap(lo[2] + z, N)], v);
    }
}

__device__ __forceinline__ double sinc_ref(int n, int N)
{
    if (n == 0)
        return 1.0;
    const double x = kPiRef * (double)n / (double)N;
    return sin(x) / x;
}

// spectrum [N][N][N/2+1]: x (l) slowest, the halved axis z (n) fastest
__global__ __launch_bounds__(kPmBlock) void k_pm_green(int N, double ismth2, double pref, double2 *__restrict__ spec)
{
    const int nh = N / 2 + 1;
    const long long e = (long long)blockIdx.x * kPmBlock + threadIdx.x;
    const long long total = (long long)N * N * nh;
    if (e >= total)
        return;
    const int n = (int)(e % nh);
    const int m0 = (int)((e / nh) % N);
    const int l0 = (int)(e / ((long long)nh * N));
    const int l = l0 > N / 2 ? l0 - N : l0;
    const int m = m0 > N / 2 ? m0 - N : m0;
    double g = 0.0;
    if (l != 0 || m != 0 || n != 0) {
        const double k2 = (double)(l * l + m * m + n * n);
        const double ff = 1.0 / (sinc_ref(l, N) * sinc_ref(m, N) * sinc_ref(n, N));
        const double ff2 = ff * ff;
        g = exp(-k2 * ismth2) / k2 * (ff2 * ff2) * pref;
    }
    double2 v = spec[e];
    v.x *= g;
    v.y *= g;
    spec[e] = v;
}

// grad[d][cell] = f1 invx (p[+1] - p[-1]) - f2 invx (p[+2] - p[-2]) along d:
// the reference evaluates this difference at each of a particle's 8 CIC
// cells (src/partmesh.c:2694-2922); it depends on the cell alone
__global__ __launch_bounds__(kPmBlock) void k_pm_gradient(const double *__restrict__ pot, int N, double invx,
                                                          double *__restrict__ grad)
{
    const long long ncell = (long long)N * N * N;
    const long long e = (long long)blockIdx.x * kPmBlock + threadIdx.x;
    if (e >= ncell)
        return;
    const int q[3] = {(int)(e / ((long long)N * N)), (int)((e / N) % N), (int)(e % N)};
    const double f1 = (double)(float)(4.0 / 3.0), f2 = (double)(float)(1.0 / 6.0);
    for (int d = 0; d < 3; ++d) {
        auto at = [&](int o) {
            int r[3] = {q[0], q[1], q[2]};
            r[d] = wrap(r[d] + o, N);
            return pot[((long long)r[0] * N + r[1]) * N + r[2]];
        };
        grad[d * ncell + e] = f1 * invx * (at(1) - at(-1)) - f2 * invx * (at(2) - at(-2));
    }
}

// acc[d][p] = sum over the 8 CIC cells of wx wy wz grad[d][cell], in the
// reference's corner order
__global__ __launch_bounds__(kPmBlock) void k_pm_interp(PmArgs a, const double *__restrict__ grad,
                                                        float *__restrict__ acc, double *__restrict__ acc64)
{
    const long long p = (long long)blockIdx.x * kPmBlock + threadIdx.x;
    if (p >= a.n)
        return;
    if (!is_local(a, p)) {
        if (acc) {
            acc[p] = 0.0f;
            acc[a.n + p] = 0.0f;
            acc[2 * a.n + p] = 0.0f;
        }
        return;
    }
    Axis ax[3];
    particle_axes(a, p, ax);
    const int N = a.nside;
    const long long ncell = (long long)N * N * N;
    const int cx[2] = {wrap(ax[0].i, N), wrap(ax[0].ii, N)};
    const int cy[2] = {wrap(ax[1].i, N), wrap(ax[1].ii, N)};
    const int cz[2] = {wrap(ax[2].i, N), wrap(ax[2].ii, N)};
    const double wx[2] = {ax[0].wn, ax[0].w}, wy[2] = {ax[1].wn, ax[1].w}, wz[2] = {ax[2].wn, ax[2].w};
    double s[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 8; ++c) {
        const int x = c & 1, y = (c >> 1) & 1, z = c >> 2;
        const long long cell = ((long long)cx[x] * N + cy[y]) * N + cz[z];
        const double w = wx[x] * wy[y] * wz[z];
        for (int d = 0; d < 3; ++d)
            s[d] += w * grad[d * ncell + cell];
    }
    if (acc) {
        for (int d = 0; d < 3; ++d)
            acc[d * a.n + p] = (float)s[d];
    } else {
        for (int d = 0; d < 3; ++d)
            acc64[d * a.n + p] += s[d];
    }
}

__global__ void k_pm_zero(double *p, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0.0;
}

inline unsigned blocks_of(long long n) { return (unsigned)((n + kPmBlock - 1) / kPmBlock); }

}  // namespace

hipError_t pm_deposit(const PmArgs &a, double *rho, hipStream_t st)
{
    if (a.n > 0)
        k_pm_deposit<<<(unsigned)((a.n + kPmChunk - 1) / kPmChunk), kPmBlock, 0, st>>>(a, rho);
    return hipGetLastError();
}

hipError_t pm_green(int nside, float split_radius, float box, double2 *spec, hipStream_t st)
{
    // src/conv.f90:258-262: ismth2 = (2 pi smooth / box)^2, pref = box^2 / (pi N^3);
    // the mesh holds mass per cell, which the reference turns into a density
    // first: pmesh *= (NSIDEMESH / BOXSIZE)^3 in FP32 (src/partmesh.c:1797-1807),
    // folded into G here (the transform is linear)
    const double smooth = (double)split_radius, bx = (double)box;
    double ismth2 = 2.0 * kPiRef * smooth / bx;
    ismth2 = ismth2 * ismth2;
    const float rn1 = (float)nside / box;
    const float renormal = (rn1 * rn1) * rn1;
    const double pref = bx * bx / (kPiRef * (double)nside * (double)nside * (double)nside) * (double)renormal;
    const long long total = (long long)nside * nside * (nside / 2 + 1);
    k_pm_green<<<blocks_of(total), kPmBlock, 0, st>>>(nside, ismth2, pref, spec);
    return hipGetLastError();
}

hipError_t pm_interp(const PmArgs &a, const double *pot, double *grad, float *acc32, double *acc64,
                     hipStream_t st)
{
    const int N = a.nside;
    const double invx = (double)(float)(0.5 * (double)N / (double)a.box);   // src/partmesh.c:2596, FP32
    k_pm_gradient<<<blocks_of((long long)N * N * N), kPmBlock, 0, st>>>(pot, N, invx, grad);
    if (a.n > 0)
        k_pm_interp<<<blocks_of(a.n), kPmBlock, 0, st>>>(a, grad, acc32, acc64);
    return hipGetLastError();
}

hipError_t pm_zero(double *p, long long n, hipStream_t st)
{
    if (n > 0) {
        const long long b = (n + kPmBlock - 1) / kPmBlock;
        k_pm_zero<<<(unsigned)(b < 65536 ? b : 65536), kPmBlock, 0, st>>>(p, n);
    }
    return hipGetLastError();
}

}  // namespace pp2p

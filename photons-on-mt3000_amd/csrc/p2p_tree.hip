// The kd-tree on the device: build_localtree's mean-split tree
// (src/fmm.c:168-270), its boxes (center_kdtree, src/fmm.c:273-327) and the
// target marking (src/fmm.c:1226-1236), built level by level instead of
// recursively, with the reference's numbering and particle order.
//
// Level l holds the nodes at depth l as particle ranges (segments), left to
// right; all split along coordinate l % 3.  Per level:
//   mean    one thread per segment sums the sampled coordinates in FP64 in
//           the reference's order (stride len / 100000 above 10^5
//           particles), so the split is the reference's to the bit
//   flags   small = coordinate <= mean; an exclusive scan counts them
//   swaps   the reference's in-place partition (src/fmm.c:199-219) scans
//           from the left for particles above the mean and swaps each with
//           the next particle at or below it from the right.  Its outcome
//           in closed form: with k particles at or below the mean, the i-th
//           particle above it in [0, k) is swapped with the i-th particle at
//           or below it counted from the end of [k, len); the left child
//           gets k particles, or len - 1 when every particle is at or below
//           the mean (the scan then stops one short).  Two particles are
//           ordered by a plain compare (src/fmm.c:180-187).
//   children  sons with more than maxleaf particles are the next level's
//           segments, the others are leaves (scans keep both in order)
// Afterwards the subtree sizes are summed bottom-up and the depth-first ids
// (nodes in creation order, leaves in creation order) and boxes handed
// top-down.  FP64 sums and FP32 boxes are rounded per operation (no
// contraction), as the host code computes them.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "p2p_tree.h"

#pragma clang fp contract(off)

namespace pp2p {
namespace {

constexpr int kTreeBlock = 256;

inline unsigned grid_of(long long n) { return (unsigned)((n + kTreeBlock - 1) / kTreeBlock); }

__device__ __forceinline__ long long gid() { return (long long)blockIdx.x * kTreeBlock + threadIdx.x; }

__global__ __launch_bounds__(kTreeBlock) void k_tree_init(const int *__restrict__ posi, long long n,
                                                        int *__restrict__ x0, int *__restrict__ x1,
                                                        int *__restrict__ x2, int *__restrict__ idx,
                                                        int *__restrict__ segof)
{
    const long long i = gid();
    if (i >= n)
        return;
    x0[i] = posi[3 * i];
    x1[i] = posi[3 * i + 1];
    x2[i] = posi[3 * i + 2];
    idx[i] = (int)i;
    segof[i] = 0;
}

// Split value of each segment (src/fmm.c:168-197): the sampled FP64 mean for
// more than two particles, the midpoint of two, 0 for fewer.  One wave per
// segment: the lanes load 512 samples at a time into the wave's LDS slice,
// and every lane adds them to its sum one after the other in sample order
// (the same chain in every lane, each value a broadcast LDS read), so the sum
// is the reference's sequential one.  Padding the last block with +0.0
// leaves the sum unchanged: it is never -0.0 (it starts at +0.0 and an exact
// cancellation rounds to +0.0).
constexpr int kMeanBatch = 512;

__global__ __launch_bounds__(kTreeBlock) void k_tree_mean(const int2 *__restrict__ seg, int nseg,
                                                        const int *__restrict__ x, double int2pos,
                                                        double *__restrict__ mean, float *__restrict__ split)
{
    __shared__ double buf[kTreeBlock / 64][kMeanBatch];
    const int s = blockIdx.x * (kTreeBlock / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (s >= nseg)
        return;
    double *wb = buf[threadIdx.x / 64];
    const int a = seg[s].x, len = seg[s].y;
    double m = 0.0;
    float sp = 0.0f;
    if (len > 2) {
        const long long mul = len > 100000 ? len / 100000 : 1;
        const long long ns = (len + mul - 1) / mul;
        constexpr int kU = kMeanBatch / 64;
        for (long long b = 0; b < ns; b += kMeanBatch) {
            // indices past the end are clamped for the load and masked after
            int r[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const long long k = b + u * 64 + lane;
                r[u] = x[a + (k < ns ? k : ns - 1) * mul];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                wb[u * 64 + lane] = b + u * 64 + lane < ns ? int2pos * (double)r[u] : 0.0;
            __builtin_amdgcn_wave_barrier();
            const int cnt = ns - b < kMeanBatch ? (int)(ns - b) : kMeanBatch;     // the same in every lane
            int j = 0;
            for (; j + 16 <= cnt; j += 16) {
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    m += wb[j + q];
            }
            for (; j < cnt; ++j)
                m += wb[j];
            __builtin_amdgcn_wave_barrier();
        }
        m /= (double)len;
        m *= (double)mul;
        sp = (float)m;
    } else if (len == 2) {
        sp = (float)(0.5 * (int2pos * (double)(x[a] + x[a + 1])));
    }
    if (lane == 0) {
        mean[s] = m;
        split[s] = sp;
    }
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_flags(const int *__restrict__ x, const int *__restrict__ segof,
                                                         const int2 *__restrict__ seg,
                                                         const double *__restrict__ mean, double int2pos,
                                                         long long n, int *__restrict__ small)
{
    const long long i = gid();
    if (i >= n)
        return;
    const int s = segof[i];
    int f = 0;
    if (s >= 0 && seg[s].y > 2)
        f = !(int2pos * (double)x[i] > mean[s]);
    small[i] = f;
}

__global__ void k_tree_scan_end(int *__restrict__ sx, const int *__restrict__ v, long long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        sx[n] = sx[n - 1] + v[n - 1];
}

// Left child size np0 and the pairing bound kk (-1: no swaps) per segment.
__global__ __launch_bounds__(kTreeBlock) void k_tree_count(const int2 *__restrict__ seg, int nseg,
                                                         const int *__restrict__ sx, int *__restrict__ np0,
                                                         int *__restrict__ kk)
{
    const long long s = gid();
    if (s >= nseg)
        return;
    const int a = seg[s].x, len = seg[s].y;
    int left = len == 2 ? 1 : 0, k = -1;
    if (len > 2) {
        const int sm = sx[a + len] - sx[a];
        left = sm == len ? len - 1 : sm;
        k = sm == len ? -1 : sm;
    }
    np0[s] = left;
    kk[s] = k;
}

// tbl[a + r] = position of the r-th particle at or below the mean counted
// from the end of the right part [a + k, a + len).
__global__ __launch_bounds__(kTreeBlock) void k_tree_table(const int *__restrict__ segof,
                                                         const int2 *__restrict__ seg, const int *__restrict__ kk,
                                                         const int *__restrict__ small, const int *__restrict__ sx,
                                                         long long n, int *__restrict__ tbl)
{
    const long long i = gid();
    if (i >= n)
        return;
    const int s = segof[i];
    if (s < 0 || !small[i])
        return;
    const int k = kk[s];
    const int a = seg[s].x, len = seg[s].y;
    if (k < 0 || i < a + k)
        return;
    tbl[a + (sx[a + len] - sx[i + 1])] = (int)i;
}

__device__ __forceinline__ void swap_part(int *x0, int *x1, int *x2, int *idx, long long i, long long j)
{
    int t = x0[i]; x0[i] = x0[j]; x0[j] = t;
    t = x1[i]; x1[i] = x1[j]; x1[j] = t;
    t = x2[i]; x2[i] = x2[j]; x2[j] = t;
    t = idx[i]; idx[i] = idx[j]; idx[j] = t;
}

// The partition's swaps, in place: each swapped pair belongs to one thread.
__global__ __launch_bounds__(kTreeBlock) void k_tree_swap(int *__restrict__ x0, int *__restrict__ x1,
                                                        int *__restrict__ x2, int *__restrict__ idx, int dir,
                                                        const int *__restrict__ segof,
                                                        const int2 *__restrict__ seg, const int *__restrict__ kk,
                                                        const int *__restrict__ small, const int *__restrict__ sx,
                                                        const int *__restrict__ tbl, long long n)
{
    const long long i = gid();
    if (i >= n)
        return;
    const int s = segof[i];
    if (s < 0)
        return;
    const int a = seg[s].x, len = seg[s].y;
    if (len == 2) {
        if (i == a) {
            const int *x = dir == 0 ? x0 : dir == 1 ? x1 : x2;
            if (x[a] > x[a + 1])
                swap_part(x0, x1, x2, idx, a, a + 1);
        }
        return;
    }
    const int k = kk[s];
    if (k < 0 || i >= a + k || small[i])
        return;
    const int rb = (int)(i - a) - (sx[i] - sx[a]);
    swap_part(x0, x1, x2, idx, i, tbl[a + rb]);
}

// nodef/leaff[2s + c]: son c of segment s is a node / a leaf.
__global__ __launch_bounds__(kTreeBlock) void k_tree_kidflags(const int2 *__restrict__ seg, int nseg,
                                                            const int *__restrict__ np0, int maxleaf,
                                                            int *__restrict__ nodef, int *__restrict__ leaff)
{
    const long long s = gid();
    if (s >= nseg)
        return;
    const int np[2] = {np0[s], seg[s].y - np0[s]};
    for (int c = 0; c < 2; ++c) {
        nodef[2 * s + c] = np[c] > maxleaf;
        leaff[2 * s + c] = np[c] <= maxleaf;
    }
    if (s == nseg - 1) {
        // a trailing zero: the scans' last entries are the totals
        nodef[2 * s + 2] = 0;
        leaff[2 * s + 2] = 0;
    }
}

// The next level's segments, this level's leaves, and kid[s] = the two sons
// (>= 0: next-level segment, < 0: this level's leaf -(q + 1)).
__global__ __launch_bounds__(kTreeBlock) void k_tree_next(const int2 *__restrict__ seg, int nseg,
                                                        const int *__restrict__ np0, const int *__restrict__ nodef,
                                                        const int *__restrict__ noff, const int *__restrict__ loff,
                                                        int2 *__restrict__ next, int2 *__restrict__ leaf,
                                                        int2 *__restrict__ kid)
{
    const long long s = gid();
    if (s >= nseg)
        return;
    const int a = seg[s].x, len = seg[s].y, l0 = np0[s];
    int r[2];
    for (int c = 0; c < 2; ++c) {
        const int2 rg = c ? make_int2(a + l0, len - l0) : make_int2(a, l0);
        if (nodef[2 * s + c]) {
            next[noff[2 * s + c]] = rg;
            r[c] = noff[2 * s + c];
        } else {
            leaf[loff[2 * s + c]] = rg;
            r[c] = -loff[2 * s + c] - 1;
        }
    }
    kid[s] = make_int2(r[0], r[1]);
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_segof(int *__restrict__ segof, const int2 *__restrict__ seg,
                                                         const int *__restrict__ np0, const int2 *__restrict__ kid,
                                                         long long n)
{
    const long long i = gid();
    if (i >= n)
        return;
    const int s = segof[i];
    if (s < 0)
        return;
    const int k = i >= seg[s].x + np0[s] ? kid[s].y : kid[s].x;
    segof[i] = k >= 0 ? k : -1;
}

// below[s] = (nodes, leaves) in the subtree of segment s.
__global__ __launch_bounds__(kTreeBlock) void k_tree_below(const int2 *__restrict__ kid, int nseg,
                                                         const int2 *__restrict__ below_next,
                                                         int2 *__restrict__ below)
{
    const long long s = gid();
    if (s >= nseg)
        return;
    const int k[2] = {kid[s].x, kid[s].y};
    int nb = 1, lb = 0;
    for (int c = 0; c < 2; ++c) {
        if (k[c] >= 0) {
            nb += below_next[k[c]].x;
            lb += below_next[k[c]].y;
        } else {
            lb += 1;
        }
    }
    below[s] = make_int2(nb, lb);
}

__device__ __forceinline__ void set_dim(float4 &v, int dir, float x)
{
    if (dir == 0)
        v.x = x;
    else if (dir == 1)
        v.y = x;
    else
        v.z = x;
}

// Depth-first ids and boxes: node s of this level has id idb[s].x, its
// subtree's leaves start at idb[s].y, its box is [lo, hi].  Writes the node,
// its leaf sons, and the ids and boxes of its node sons.
__global__ __launch_bounds__(kTreeBlock) void k_tree_place(TreeLevelView lv, TreeLevelView nx, int dir,
                                                         TreeOut out)
{
    const long long s = gid();
    if (s >= lv.nseg)
        return;
    const int I = lv.idb[s].x, B = lv.idb[s].y;
    const float4 lo = lv.lo[s], hi = lv.hi[s];
    const float w[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    const float cc[3] = {(float)(0.5 * (double)(hi.x + lo.x)), (float)(0.5 * (double)(hi.y + lo.y)),
                         (float)(0.5 * (double)(hi.z + lo.z))};
    out.node_width[I] = make_float4(w[0], w[1], w[2], 0.f);
    out.node_center[I] = make_float4(cc[0], cc[1], cc[2], 0.f);
    const int k[2] = {lv.kid[s].x, lv.kid[s].y};
    const float sp = lv.split[s];
    int id[2], base[2];
    id[0] = I + 1;
    base[0] = B;
    if (k[0] >= 0) {
        id[1] = I + 1 + nx.below[k[0]].x;
        base[1] = B + nx.below[k[0]].y;
    } else {
        id[1] = I + 1;
        base[1] = B + 1;
    }
    const float lod = dir == 0 ? lo.x : dir == 1 ? lo.y : lo.z;
    const float hid = dir == 0 ? hi.x : dir == 1 ? hi.y : hi.z;
    int son[2];
    for (int c = 0; c < 2; ++c) {
        if (k[c] >= 0) {
            float4 clo = lo, chi = hi;
            set_dim(c == 0 ? chi : clo, dir, sp);
            nx.idb[k[c]] = make_int2(id[c], base[c]);
            nx.lo[k[c]] = clo;
            nx.hi[k[c]] = chi;
            son[c] = id[c];
        } else {
            const int l = base[c];
            const int2 rg = lv.leaf[-k[c] - 1];
            float4 lw = make_float4(w[0], w[1], w[2], 0.f), lc = make_float4(cc[0], cc[1], cc[2], 0.f);
            if (c == 0) {
                set_dim(lw, dir, sp - lod);
                set_dim(lc, dir, (float)(0.5 * (double)(lod + sp)));
            } else {
                set_dim(lw, dir, hid - sp);
                set_dim(lc, dir, (float)(0.5 * (double)(hid + sp)));
            }
            out.leaf_rng[l] = rg;
            out.leaf_width[l] = lw;
            out.leaf_center[l] = lc;
            son[c] = -l - 1;
        }
    }
    out.son[I] = make_int2(son[0], son[1]);
}

// Target marking (src/fmm.c:1226-1236): 0 when the leaf holds a particle
// with tag 1, else -1.
__global__ __launch_bounds__(kTreeBlock) void k_tree_active(const int2 *__restrict__ rng, int nleaf,
                                                          const int *__restrict__ idx, const int *__restrict__ tag,
                                                          int *__restrict__ active)
{
    const long long l = gid();
    if (l >= nleaf)
        return;
    const int2 r = rng[l];
    int a = -1;
    for (int k = 0; k < r.y; ++k)
        if (tag[idx[r.x + k]] == 1) {
            a = 0;
            break;
        }
    active[l] = a;
}

// Positions in tree order: x3 = int2pos * posi (FP64, AoS).
__global__ __launch_bounds__(kTreeBlock) void k_tree_x3(const int *__restrict__ x0, const int *__restrict__ x1,
                                                      const int *__restrict__ x2, double int2pos, long long n,
                                                      double *__restrict__ x3)
{
    const long long i = gid();
    if (i >= n)
        return;
    x3[3 * i] = int2pos * (double)x0[i];
    x3[3 * i + 1] = int2pos * (double)x1[i];
    x3[3 * i + 2] = int2pos * (double)x2[i];
}

}  // namespace

hipError_t tree_init(const int *posi, long long n, int *const x[3], int *idx, int *segof, hipStream_t st)
{
    k_tree_init<<<grid_of(n), kTreeBlock, 0, st>>>(posi, n, x[0], x[1], x[2], idx, segof);
    return hipGetLastError();
}

hipError_t tree_split(const TreeScratch &w, const int2 *seg, int nseg, int dir, double int2pos, long long n,
                      double *mean, float *split, int *np0, int *kk, void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (!temp)
        return rocprim::exclusive_scan(nullptr, *temp_bytes, w.small, w.sx, 0, (size_t)n, rocprim::plus<int>(), st);
    const int *x = w.x[dir];
    k_tree_mean<<<(unsigned)((nseg + kTreeBlock / 64 - 1) / (kTreeBlock / 64)), kTreeBlock, 0, st>>>(
        seg, nseg, x, int2pos, mean, split);
    k_tree_flags<<<grid_of(n), kTreeBlock, 0, st>>>(x, w.segof, seg, mean, int2pos, n, w.small);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, w.small, w.sx, 0, (size_t)n, rocprim::plus<int>(), st);
    if (e != hipSuccess)
        return e;
    k_tree_scan_end<<<1, 64, 0, st>>>(w.sx, w.small, n);
    k_tree_count<<<grid_of(nseg), kTreeBlock, 0, st>>>(seg, nseg, w.sx, np0, kk);
    k_tree_table<<<grid_of(n), kTreeBlock, 0, st>>>(w.segof, seg, kk, w.small, w.sx, n, w.tbl);
    k_tree_swap<<<grid_of(n), kTreeBlock, 0, st>>>(w.x[0], w.x[1], w.x[2], w.idx, dir, w.segof, seg, kk, w.small,
                                                   w.sx, w.tbl, n);
    return hipGetLastError();
}

hipError_t tree_children(const int2 *seg, int nseg, const int *np0, int maxleaf, int *nodef, int *leaff, int *noff,
                         int *loff, void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (!temp) {
        size_t a = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, a, nodef, noff, 0, (size_t)2 * nseg + 1,
                                               rocprim::plus<int>(), st);
        *temp_bytes = a;
        return e;
    }
    k_tree_kidflags<<<grid_of(nseg), kTreeBlock, 0, st>>>(seg, nseg, np0, maxleaf, nodef, leaff);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, nodef, noff, 0, (size_t)2 * nseg + 1,
                                           rocprim::plus<int>(), st);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(temp, *temp_bytes, leaff, loff, 0, (size_t)2 * nseg + 1, rocprim::plus<int>(),
                                    st);
    return e;
}

hipError_t tree_next(const TreeScratch &w, const int2 *seg, int nseg, const int *np0, const int *nodef,
                     const int *noff, const int *loff, int2 *next, int2 *leaf, int2 *kid, long long n, hipStream_t st)
{
    k_tree_next<<<grid_of(nseg), kTreeBlock, 0, st>>>(seg, nseg, np0, nodef, noff, loff, next, leaf, kid);
    k_tree_segof<<<grid_of(n), kTreeBlock, 0, st>>>(w.segof, seg, np0, kid, n);
    return hipGetLastError();
}

hipError_t tree_below(const int2 *kid, int nseg, const int2 *below_next, int2 *below, hipStream_t st)
{
    k_tree_below<<<grid_of(nseg), kTreeBlock, 0, st>>>(kid, nseg, below_next, below);
    return hipGetLastError();
}

hipError_t tree_place(const TreeLevelView &lv, const TreeLevelView &nx, int dir, const TreeOut &out, hipStream_t st)
{
    k_tree_place<<<grid_of(lv.nseg), kTreeBlock, 0, st>>>(lv, nx, dir, out);
    return hipGetLastError();
}

hipError_t tree_finish(const TreeScratch &w, const TreeOut &out, int nleaf, const int *tag, double int2pos,
                       long long n, double *x3, hipStream_t st)
{
    k_tree_active<<<grid_of(nleaf), kTreeBlock, 0, st>>>(out.leaf_rng, nleaf, w.idx, tag, out.leaf_active);
    k_tree_x3<<<grid_of(n), kTreeBlock, 0, st>>>(w.x[0], w.x[1], w.x[2], int2pos, n, x3);
    return hipGetLastError();
}

}  // namespace pp2p

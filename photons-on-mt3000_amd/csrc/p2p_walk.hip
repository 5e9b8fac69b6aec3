// Interaction lists on the device: the symmetric dual walk of the kd-tree
// (walk_task_p2p, src/fmm.c:656-792, with the acceptance test of
// src/fmm.c:411-479) run level-synchronously instead of recursively.
//
// A frame is a (target, source) pair of tree cells; >= 0 names a node,
// < 0 a leaf as -(id + 1).  The recursive walk calls, from a frame:
//   node self pair   -> (s0,s0) (s0,s1) (s1,s0) (s1,s1)
//   leaf-node pair   -> opened (accept == 0): the node's two sons, else nothing
//   node-node pair   -> opened: the two sons of the wider box (sum of widths)
//   leaf-leaf pair   -> emits the task when the target leaf holds a local
//                       particle (active >= 0)
// Round r holds the frames at call depth r in depth-first order (children
// are placed at an exclusive scan of the per-frame child counts).  Three
// passes turn the rounds into the task list in the order the recursive walk
// emits it, without moving a frame after its round:
//   forward   frames of round r+1 from round r (one host sync per round for
//             the next round's size)
//   backward  tasks below each frame: its own (0 or 1) plus its children's,
//             from the exclusive scan x of the next round's counts
//   positions each frame's first task position, passed from parent to
//             child as parent + (x[child] - x[first child]); the emitting
//             frames write their task there
// The geometry is FP32 with each operation rounded on its own (no
// contraction), as the host's cpuType arithmetic.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "p2p_walk.h"

#pragma clang fp contract(off)

namespace pp2p {
namespace {

constexpr int kWalkBlock = 256;

// 1: well separated (multipole), 0: open, -1: beyond the cutoff.  Operation
// order and types as src/fmm.c:411-479: the half-sum of widths in FP64, the
// distance test on the gap vector, the opening test on the centre distance.
__device__ __forceinline__ int accept(const float4 wi, const float4 wj, float dx, float dy, float dz,
                                      float open_angle, float cutoff)
{
    const float w0 = (float)((double)(wi.x + wj.x) * 0.5);
    const float w1 = (float)((double)(wi.y + wj.y) * 0.5);
    const float w2 = (float)((double)(wi.z + wj.z) * 0.5);
    float m0 = dx < 0.0f ? -dx : dx;
    float m1 = dy < 0.0f ? -dy : dy;
    float m2 = dz < 0.0f ? -dz : dz;
    m0 -= w0;
    m1 -= w1;
    m2 -= w2;
    if (m0 <= 0.0f)
        m0 = 0.0f;
    if (m1 <= 0.0f)
        m1 = 0.0f;
    if (m2 <= 0.0f)
        m2 = 0.0f;
    const float dd2 = dx * dx + dy * dy + dz * dz;
    if ((double)(m0 + m1 + m2) < 1.0e-15)
        return 0;
    const float dm2 = m0 * m0 + m1 * m1 + m2 * m2;
    if (dm2 > cutoff * cutoff)
        return -1;
    float wmax = w0;
    if (w1 > wmax)
        wmax = w1;
    if (w2 > wmax)
        wmax = w2;
    wmax = fabsf(wmax);
    wmax *= 2.0f;
    return (wmax * wmax < open_angle * open_angle * dd2) ? 1 : 0;
}

// Child count (0, 2 or 4) of frame f, the children into ch[] when ch is not
// null; *emit = 1 for a leaf-leaf frame whose target is active, 2 for a
// well-separated pair (acceptance 1: an M2L task of walk_task_m2l).
__device__ __forceinline__ int expand_frame(const WalkTree &t, int2 f, int2 *ch, int *emit)
{
    const int im = f.x, jm = f.y;
    *emit = 0;
    if (im < 0 && jm < 0) {
        *emit = t.leaf_active[-im - 1] >= 0;
        return 0;
    }
    if (im == jm) {
        const int2 s = t.son[im];
        if (ch) {
            ch[0] = make_int2(s.x, s.x);
            ch[1] = make_int2(s.x, s.y);
            ch[2] = make_int2(s.y, s.x);
            ch[3] = make_int2(s.y, s.y);
        }
        return 4;
    }
    float4 ac, aw, bc, bw;
    if (im < 0) {
        ac = t.leaf_center[-im - 1];
        aw = t.leaf_width[-im - 1];
    } else {
        ac = t.node_center[im];
        aw = t.node_width[im];
    }
    if (jm < 0) {
        bc = t.leaf_center[-jm - 1];
        bw = t.leaf_width[-jm - 1];
    } else {
        bc = t.node_center[jm];
        bw = t.node_width[jm];
    }
    // leaf-node: accept(leaf, node); node-leaf: accept(leaf, node) on the
    // node-minus-leaf distance; the half-sum of widths is symmetric
    const int acc = accept(aw, bw, ac.x - bc.x, ac.y - bc.y, ac.z - bc.z, t.open_angle, t.cutoff);
    if (acc != 0) {
        *emit = acc == 1 ? 2 : 0;
        return 0;
    }
    if (ch) {
        if (im >= 0 && (jm < 0 || aw.x + aw.y + aw.z > bw.x + bw.y + bw.z)) {
            const int2 s = t.son[im];
            ch[0] = make_int2(s.x, jm);
            ch[1] = make_int2(s.y, jm);
        } else {
            const int2 s = t.son[jm];
            ch[0] = make_int2(im, s.x);
            ch[1] = make_int2(im, s.y);
        }
    }
    return 2;
}

struct KidsOf {
    __host__ __device__ unsigned operator()(unsigned char c) const { return c & 7u; }
};
struct EmitOf {
    __host__ __device__ unsigned long long operator()(unsigned char c) const { return (c >> 3) & 1u; }
};

inline unsigned grid_of(long long n) { return (unsigned)((n + kWalkBlock - 1) / kWalkBlock); }

// forward: code[i] = children | task << 3 | well-separated << 4, and the
// block's number of well-separated frames into bs[block]
__global__ __launch_bounds__(kWalkBlock) void k_walk_code(WalkTree t, const int2 *__restrict__ f, long long n,
                                                        unsigned char *__restrict__ code, int *__restrict__ bs)
{
    const long long i = (long long)blockIdx.x * kWalkBlock + threadIdx.x;
    int emit = 0;
    if (i < n) {
        const int k = expand_frame(t, f[i], nullptr, &emit);
        code[i] = (unsigned char)(k | emit << 3);  // emit 1: bit 3, emit 2: bit 4
    }
    const int nsep = __syncthreads_count(emit == 2);
    if (threadIdx.x == 0)
        bs[blockIdx.x] = nsep;
}

// off[n] = off[n-1] + children of the last frame (the next round's size)
__global__ void k_walk_off_end(unsigned *__restrict__ off, const unsigned char *__restrict__ code, long long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        off[n] = off[n - 1] + (code[n - 1] & 7u);
}

__global__ __launch_bounds__(kWalkBlock) void k_walk_children(WalkTree t, const int2 *__restrict__ f, long long n,
                                                            const unsigned *__restrict__ off,
                                                            int2 *__restrict__ next)
{
    const long long i = (long long)blockIdx.x * kWalkBlock + threadIdx.x;
    if (i >= n)
        return;
    int2 ch[4];
    int emit;
    const int k = expand_frame(t, f[i], ch, &emit);
    const unsigned o = off[i];
    for (int c = 0; c < k; ++c)
        next[o + c] = ch[c];
}

// backward: tasks at and below each frame
__global__ __launch_bounds__(kWalkBlock) void k_walk_below(const unsigned char *__restrict__ code,
                                                         const unsigned *__restrict__ off,
                                                         const int *__restrict__ xn, long long n,
                                                         int *__restrict__ cnt)
{
    const long long i = (long long)blockIdx.x * kWalkBlock + threadIdx.x;
    if (i >= n)
        return;
    const unsigned c = code[i];
    int v = (c >> 3) & 1;
    if (c & 7u) {
        const unsigned o = off[i];
        v += xn[o + (c & 7u)] - xn[o];
    }
    cnt[i] = v;
}

__global__ void k_walk_x_end(int *__restrict__ x, const int *__restrict__ cnt, long long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        x[n] = x[n - 1] + cnt[n - 1];
}

// positions: each frame's first task position; emitting frames write their task
__global__ __launch_bounds__(kWalkBlock) void k_walk_emit(const int2 *__restrict__ f,
                                                        const unsigned char *__restrict__ code,
                                                        const unsigned *__restrict__ off,
                                                        const int *__restrict__ xn, long long n,
                                                        const int *__restrict__ pos,
                                                        int *__restrict__ pos_next, int *__restrict__ task_t,
                                                        int *__restrict__ task_s)
{
    const long long i = (long long)blockIdx.x * kWalkBlock + threadIdx.x;
    if (i >= n)
        return;
    const unsigned c = code[i];
    const int p = pos[i];
    if ((c >> 3) & 1) {
        const int2 fr = f[i];
        task_t[p] = -fr.x - 1;
        task_s[p] = -fr.y - 1;
        return;
    }
    const unsigned k = c & 7u;
    if (k) {
        const unsigned o = off[i];
        const int x0 = xn[o];
        for (unsigned j = 0; j < k; ++j)
            pos_next[o + j] = p + (xn[o + j] - x0);
    }
}

// the well-separated frames of block b written as cells (leaf id, or
// nleaf + node id) at out + bso[b] + their rank in the block
__global__ __launch_bounds__(kWalkBlock) void k_walk_m2l(const int2 *__restrict__ f,
                                                       const unsigned char *__restrict__ code, long long n,
                                                       const int *__restrict__ bso, int nleaf,
                                                       int *__restrict__ ct, int *__restrict__ cs)
{
    __shared__ int wave_base[kWalkBlock / 64];
    const long long i = (long long)blockIdx.x * kWalkBlock + threadIdx.x;
    const bool sep = i < n && ((code[i] >> 4) & 1u);
    const unsigned long long m = __ballot(sep);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        wave_base[wave] = __popcll(m);
    __syncthreads();
    if (!sep)
        return;
    int o = bso[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w)
        o += wave_base[w];
    const int2 v = f[i];
    ct[o] = v.x < 0 ? -v.x - 1 : nleaf + v.x;
    cs[o] = v.y < 0 ? -v.y - 1 : nleaf + v.y;
}

}  // namespace

long long walk_blocks(long long n) { return grid_of(n); }

hipError_t walk_m2l(const int2 *f, const unsigned char *code, long long n, const int *bs, int *bso, int nleaf,
                    int *ct, int *cs, void *temp, size_t *temp_bytes, hipStream_t st)
{
    const size_t nb = grid_of(n);
    if (!temp)
        return rocprim::exclusive_scan(nullptr, *temp_bytes, bs, bso, 0, nb, rocprim::plus<int>(), st);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, bs, bso, 0, nb, rocprim::plus<int>(), st);
    if (e != hipSuccess)
        return e;
    k_walk_m2l<<<(unsigned)nb, kWalkBlock, 0, st>>>(f, code, n, bso, nleaf, ct, cs);
    return hipGetLastError();
}

hipError_t walk_code(const WalkTree &t, const int2 *f, long long n, unsigned char *code, unsigned *off, int *bs,
                     unsigned long long *emitted, unsigned long long *separated, void *temp, size_t *temp_bytes,
                     hipStream_t st)
{
    auto kids = rocprim::make_transform_iterator(code, KidsOf());
    auto emits = rocprim::make_transform_iterator(code, EmitOf());
    const size_t nb = grid_of(n);
    if (!temp) {
        size_t a = 0, b = 0, c = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, a, kids, off, 0u, (size_t)n, rocprim::plus<unsigned>(), st);
        if (e == hipSuccess)
            e = rocprim::reduce(nullptr, b, emits, emitted, 0ull, (size_t)n, rocprim::plus<unsigned long long>(), st);
        if (e == hipSuccess)
            e = rocprim::reduce(nullptr, c, bs, separated, 0ull, nb, rocprim::plus<unsigned long long>(), st);
        *temp_bytes = a > b ? (a > c ? a : c) : (b > c ? b : c);
        return e;
    }
    k_walk_code<<<(unsigned)nb, kWalkBlock, 0, st>>>(t, f, n, code, bs);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, kids, off, 0u, (size_t)n, rocprim::plus<unsigned>(), st);
    if (e == hipSuccess)
        e = rocprim::reduce(temp, *temp_bytes, emits, emitted, 0ull, (size_t)n, rocprim::plus<unsigned long long>(),
                            st);
    if (e == hipSuccess)
        e = rocprim::reduce(temp, *temp_bytes, bs, separated, 0ull, nb, rocprim::plus<unsigned long long>(), st);
    if (e != hipSuccess)
        return e;
    k_walk_off_end<<<1, 64, 0, st>>>(off, code, n);
    return hipGetLastError();
}

hipError_t walk_children(const WalkTree &t, const int2 *f, long long n, const unsigned *off, int2 *next,
                         hipStream_t st)
{
    k_walk_children<<<grid_of(n), kWalkBlock, 0, st>>>(t, f, n, off, next);
    return hipGetLastError();
}

hipError_t walk_below(const unsigned char *code, const unsigned *off, const int *xn, long long n, int *cnt, int *x,
                      void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (!temp)
        return rocprim::exclusive_scan(nullptr, *temp_bytes, cnt, x, 0, (size_t)n, rocprim::plus<int>(), st);
    k_walk_below<<<grid_of(n), kWalkBlock, 0, st>>>(code, off, xn, n, cnt);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, cnt, x, 0, (size_t)n, rocprim::plus<int>(), st);
    if (e != hipSuccess)
        return e;
    k_walk_x_end<<<1, 64, 0, st>>>(x, cnt, n);
    return hipGetLastError();
}

hipError_t walk_emit(const int2 *f, const unsigned char *code, const unsigned *off, const int *xn, long long n,
                     const int *pos, int *pos_next, int *task_t, int *task_s, hipStream_t st)
{
    k_walk_emit<<<grid_of(n), kWalkBlock, 0, st>>>(f, code, off, xn, n, pos, pos_next, task_t, task_s);
    return hipGetLastError();
}

}  // namespace pp2p

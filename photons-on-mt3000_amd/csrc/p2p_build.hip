// Device construction of everything the force kernel reads besides the
// particles: leaf frames, the target CSR of the task list and the work items.
//
// Target CSR (what the reference builds with thrust in src/gpu_entry.cu:
// 132-138): a stable radix sort of (task_t, task_s) by target, then the
// unique targets (rows) and their offsets.  Work items: each row is cut into
// i-chunks of <= 64 particles and, for the queue form, into source ranges of
// about `split` pair-interactions; queue items are sorted by cost, largest
// first (longest-processing-time order for the persistent kernel).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "p2p_build.h"
#include "p2p_cut.h"

namespace pp2p {

// ---------------------------------------------------------------- frames

// One wave per leaf: FP64 bounding-box centre as the leaf frame origin,
// positions relative to it, everything scaled by 1/rc.  frame_abs: origin 0
// and FP32 absolute coordinates (reference-style).
__global__ __launch_bounds__(256) void k_leaf_frame(const double *__restrict__ x3, const int2 *__restrict__ rng,
                                                    int nleaf, double inv_rc, int frame_abs,
                                                    double4 *__restrict__ ref, float4 *__restrict__ pos)
{
    const int leaf = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (leaf >= nleaf)
        return;
    const int2 r = rng[leaf];
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (!frame_abs && r.y > 0) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int k = lane; k < r.y; k += 64) {
            const double *p = x3 + 3 * ((long long)r.x + k);
            for (int d = 0; d < 3; ++d) {
                lo[d] = fmin(lo[d], p[d]);
                hi[d] = fmax(hi[d], p[d]);
            }
        }
        for (int o = 32; o > 0; o >>= 1)
            for (int d = 0; d < 3; ++d) {
                lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
                hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
            }
        cx = 0.5 * (lo[0] + hi[0]);
        cy = 0.5 * (lo[1] + hi[1]);
        cz = 0.5 * (lo[2] + hi[2]);
    }
    if (lane == 0)
        ref[leaf] = make_double4(cx * inv_rc, cy * inv_rc, cz * inv_rc, 0.0);
    // the absolute frame keeps the input units: the positions are the
    // reference's float coordinates, unscaled (the kernel scales r^2)
    if (frame_abs)
        inv_rc = 1.0;
    for (int k = lane; k < r.y; k += 64) {
        const long long i = (long long)r.x + k;
        const double *p = x3 + 3 * i;
        pos[i] = make_float4((float)((p[0] - cx) * inv_rc), (float)((p[1] - cy) * inv_rc),
                             (float)((p[2] - cz) * inv_rc), 0.f);
    }
}

__global__ void k_i32_to_f64(const int *__restrict__ posi, long long n3, double int2pos, double *__restrict__ x3)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n3)
        x3[k] = int2pos * (double)posi[k];
}

// abs frame in the reference's quantisation: float(INT2POS * posi)
// (src/gpu_entry.cu:94-96), then scaled
__global__ void k_f64_round_f32(double *__restrict__ x3, long long n3)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n3)
        x3[k] = (double)(float)x3[k];
}

// ---------------------------------------------------------------- fills

__global__ void k_fill_f64(double *p, long long n, double v)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = v;
}

__global__ void k_fill_i32(int *p, long long n, int v)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = v;
}

static int fill_grid(long long n)
{
    long long g = (n + 255) / 256;
    return (int)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

// ---------------------------------------------------------------- CSR

__global__ void k_row_flags(const int *__restrict__ key, int n, int *__restrict__ flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        flag[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// rowidx: inclusive scan of the flags
__global__ void k_row_emit(const int *__restrict__ key, const int *__restrict__ flag, const int *__restrict__ rowidx,
                           int n, int *__restrict__ row_tgt, int *__restrict__ row_off)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    if (flag[k]) {
        row_tgt[rowidx[k] - 1] = key[k];
        row_off[rowidx[k] - 1] = k;
    }
    if (k == n - 1)
        row_off[rowidx[k]] = n;
}

// ---------------------------------------------------------------- active lists

// One wave per leaf: number of particles with act == level.
__global__ __launch_bounds__(256) void k_active_count(const unsigned char *__restrict__ act, int level,
                                                      const int2 *__restrict__ rng, int nleaf, int *__restrict__ cnt)
{
    const int leaf = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (leaf >= nleaf)
        return;
    const int2 r = rng[leaf];
    int c = 0;
    for (int k = lane; k < r.y; k += 64)
        c += act[(long long)r.x + k] == (unsigned char)level;
    for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, 64);
    if (lane == 0)
        cnt[leaf] = c;
}

// off: exclusive scan of the counts; ilist gets the active indices in order.
__global__ __launch_bounds__(256) void k_active_emit(const unsigned char *__restrict__ act, int level,
                                                     const int2 *__restrict__ rng, int nleaf,
                                                     const int *__restrict__ off, int *__restrict__ ilist)
{
    const int leaf = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (leaf >= nleaf)
        return;
    const int2 r = rng[leaf];
    int base = off[leaf];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int k0 = 0; k0 < r.y; k0 += 64) {
        const bool a = k0 + lane < r.y && act[(long long)r.x + k0 + lane] == (unsigned char)level;
        const unsigned long long m = __ballot(a);
        if (a)
            ilist[base + __popcll(m & below)] = r.x + k0 + lane;
        base += __popcll(m);
    }
}

// ---------------------------------------------------------------- i-order

// The order in which the force kernel takes a leaf's i-particles: along the
// longest axis of the leaf's box, so that the lanes of a wave hold nearby
// particles, whose table segments for one source are then close or equal
// (fewer LDS bank conflicts on the coefficient gathers).  One wave per leaf;
// leaves of more than kOrderMax particles keep their order.  The forces are
// written to each particle's own index, so only the lane assignment changes.
constexpr int kOrderMax = 256;

__device__ __forceinline__ void order_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__global__ __launch_bounds__(256) void k_leaf_order(const float4 *__restrict__ pos, const int2 *__restrict__ rng,
                                                   int nleaf, int *__restrict__ olist, int *__restrict__ ocnt,
                                                   int *__restrict__ ooff)
{
    __shared__ float key[4][kOrderMax];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l = blockIdx.x * 4 + wv;
    if (l >= nleaf)
        return;                                      // the whole wave; no block barrier below
    const int2 r = rng[l];
    if (lane == 0) {
        ocnt[l] = r.y;
        ooff[l] = r.x;
    }
    if (r.y <= 1 || r.y > kOrderMax) {
        for (int k = lane; k < r.y; k += 64)
            olist[r.x + k] = r.x + k;
        return;
    }
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int k = lane; k < r.y; k += 64) {
        const float4 q = pos[r.x + k];
        lo[0] = fminf(lo[0], q.x); hi[0] = fmaxf(hi[0], q.x);
        lo[1] = fminf(lo[1], q.y); hi[1] = fmaxf(hi[1], q.y);
        lo[2] = fminf(lo[2], q.z); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fminf(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], o, 64));
        }
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    for (int k = lane; k < r.y; k += 64) {
        const float4 q = pos[r.x + k];
        key[wv][k] = axis == 0 ? q.x : (axis == 1 ? q.y : q.z);
    }
    order_lds_sync();
    // rank by (key, index): a permutation of [0, n)
    for (int k = lane; k < r.y; k += 64) {
        const float kk = key[wv][k];
        int rank = 0;
        for (int j = 0; j < r.y; ++j) {
            const float kj = key[wv][j];
            rank += (kj < kk) || (kj == kk && j < k);
        }
        olist[r.x + rank] = r.x + k;
    }
}

// ---------------------------------------------------------------- items

struct RowView {
    const int *row_tgt, *row_off, *src;
    const int2 *rng;
    const unsigned *nsx;             // exclusive prefix of the source sizes (mod 2^32)
    const int *act_cnt, *act_off;   // null: all particles of the target
    float stage_ops;                 // > 0: most i-particles per item
    int ipl_max;                     // most i-particles per lane, 1..3
};

__device__ __forceinline__ void row_i(const RowView &v, int tgt, int &ibeg, int &ni)
{
    if (v.act_cnt) {
        ibeg = v.act_off[tgt];
        ni = v.act_cnt[tgt];
    } else {
        const int2 r = v.rng[tgt];
        ibeg = r.x;
        ni = r.y;
    }
}

// pairs of every row (sum over its tasks of n_i * n_s), and the rows' n_s
// summed into *staged (one atomic per wave; blockDim.x is a multiple of 64)
__global__ void k_row_cost(RowView v, int nrows, unsigned long long *__restrict__ cost,
                           unsigned long long *__restrict__ staged)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long ns = 0;
    if (r < nrows) {
        int ibeg, ni;
        row_i(v, v.row_tgt[r], ibeg, ni);
        ns = v.nsx[v.row_off[r + 1]] - v.nsx[v.row_off[r]];
        cost[r] = ns * (unsigned long long)ni;
    }
    for (int o = 32; o > 0; o >>= 1)
        ns += __shfl_xor(ns, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && ns > 0)
        atomicAdd(staged, ns);
}

__global__ void k_zero_u64(unsigned long long *p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = 0ull;
}

// Shape of an item with c i-particles: L lanes per phase, G = 64 / L phases
// (lane groups that split the sources; lanes beyond G * L idle), IPL
// i-particles per lane, L * IPL >= c slots.  The wave time per source particle
// is proportional to IPL / G; the smallest value wins, and a tie goes to more
// i-particles per lane (more independent pairs per step: 1% on the 256^3
// clustered list).  48 -> L 16 (G 4) x IPL 3, 40 -> L 20 (G 3) x IPL 2,
// 64 -> L 32 (G 2) x IPL 2, 1 -> L 1 (G 64) x 1.
__device__ __forceinline__ void item_shape(int c, int ipl_max, int &lanes, int &ipl)
{
    lanes = kWave;
    ipl = (c + kWave - 1) / kWave;
    float best = (float)ipl;
    for (int l = kWave - 1; l >= 1; --l) {
        const int p = (c + l - 1) / l;
        if (p > ipl_max)
            break;                                   // fewer lanes only need more slots
        const float t = (float)p / (float)(kWave / l);
        if (t < best * 0.999f || (t < best * 1.001f && p > ipl)) {
            best = t;
            lanes = l;
            ipl = p;
        }
    }
}

// i-chunks of a target with ni particles: k = ceil(ni / 256) near-equal chunks
// (chunk_model > 0 asks for chunks of at most that many particles instead)
__device__ __forceinline__ int n_chunks(int ni, float max_chunk, int ipl_max)
{
    int cap = kWave * ipl_max;
    if (max_chunk > 0.f && (int)max_chunk < cap)
        cap = max((int)max_chunk, 1);
    return (ni + cap - 1) / cap;
}

__device__ __forceinline__ void chunk_of(int ni, int k, int q, int &beg, int &size)
{
    const int base = ni / k, rem = ni % k;
    beg = q * base + min(q, rem);
    size = base + (q < rem ? 1 : 0);
}

// Source ranges of a row: greedy cut when the running work of the largest
// i-chunk, (n_i + kStagePairs) per source particle, would exceed `split`;
// split <= 0: one range.  Counting the staging keeps a small target with a
// long source list from becoming one long item (it would run last in the
// cost order and hold the launch open).  Every (range, chunk) pair is one
// item; its cost is the modelled work (IPL * 64 / G + kStagePairs) * sum n_s.
template <bool EMIT>
__device__ int row_ranges(const RowView &v, int r, long long split, int tgt, int ibeg, int ni,
                          Item *items, unsigned *cost, int out0)
{
    if (ni == 0)
        return 0;
    const int k = n_chunks(ni, v.stage_ops, v.ipl_max);
    const long long wi = (ni + k - 1) / k;
    const int k0 = v.row_off[r], k1 = v.row_off[r + 1];
    int nr = 0, out = out0;
    int beg = k0;
    // the row's chunks have two sizes at most (ni / k and one more): their
    // shapes once per row, not once per item (the shape search divides in a
    // loop of up to 63 steps)
    int lanes_lo = 0, ipl_lo = 0, lanes_hi = 0, ipl_hi = 0;
    if (EMIT) {
        const int base = ni / k;
        if (base > 0)
            item_shape(base, v.ipl_max, lanes_lo, ipl_lo);
        if (ni % k)
            item_shape(base + 1, v.ipl_max, lanes_hi, ipl_hi);
    }
    auto emit = [&](int s_beg, int s_end, long long ns) {
        for (int q = 0; q < k; ++q) {
            int cb, cs;
            chunk_of(ni, k, q, cb, cs);
            const bool hi = cs > ni / k;
            const int lanes = hi ? lanes_hi : lanes_lo, ipl = hi ? ipl_hi : ipl_lo;
            items[out] = Item{tgt, ibeg + cb, cs, r, s_beg, s_end, lanes, ipl};
            if (cost)
                cost[out] = (unsigned)min(ns * ((long long)(kWave * ipl / (kWave / lanes)) + kStagePairs),
                                          0xffffffffll);
            ++out;
        }
    };
    // the cut after `beg`: a binary search on the prefix of the source sizes
    const unsigned *P = v.nsx;
    for (;;) {
        const int cut = split > 0 ? range_cut(PrefixArray{P}, beg, k1, wi + kStagePairs, split) : k1;
        ++nr;
        if (cut >= k1) {
            if (EMIT)
                emit(beg, k1, (long long)(P[k1] - P[beg]));
            break;
        }
        if (EMIT)
            emit(beg, cut, (long long)(P[cut] - P[beg]));
        beg = cut;
    }
    return nr * k;
}

__global__ void k_items_count(RowView v, int nrows, long long split, int *__restrict__ count)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows)
        return;
    int ibeg, ni;
    const int tgt = v.row_tgt[r];
    row_i(v, tgt, ibeg, ni);
    count[r] = row_ranges<false>(v, r, split, tgt, ibeg, ni, nullptr, nullptr, 0);
}

// off: exclusive scan of the counts (nrows + 1 entries)
__global__ void k_items_emit(RowView v, int nrows, long long split, const int *__restrict__ off,
                             Item *__restrict__ items, unsigned *__restrict__ cost)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows)
        return;
    int ibeg, ni;
    const int tgt = v.row_tgt[r];
    row_i(v, tgt, ibeg, ni);
    row_ranges<true>(v, r, split, tgt, ibeg, ni, items, cost, off[r]);
}

__global__ void k_gather_items(const Item *__restrict__ in, const int *__restrict__ order, int n, Item *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = in[order[k]];
}

__global__ void k_iota(int *p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        p[k] = k;
}

struct SrcSize {
    const int2 *rng;
    __host__ __device__ unsigned operator()(int s) const { return (unsigned)rng[s].y; }
};

__global__ void k_zero_u32(unsigned *p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = 0u;
}

// ---------------------------------------------------------------- host side

static int grid_of(long long n, int b = 256) { return (int)((n + b - 1) / b); }

hipError_t launch_leaf_frame(const double *x3, const int2 *rng, int nleaf, double inv_rc, bool frame_abs,
                             double4 *ref, float4 *pos, hipStream_t st)
{
    if (nleaf <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_leaf_frame, dim3(grid_of(nleaf, 4)), dim3(256), 0, st, x3, rng, nleaf, inv_rc,
                       frame_abs ? 1 : 0, ref, pos);
    return hipGetLastError();
}

hipError_t launch_i32_to_f64(const int *posi, long long n3, double int2pos, bool round_f32, double *x3, hipStream_t st)
{
    if (n3 <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_i32_to_f64, dim3(grid_of(n3)), dim3(256), 0, st, posi, n3, int2pos, x3);
    if (round_f32)
        hipLaunchKernelGGL(k_f64_round_f32, dim3(grid_of(n3)), dim3(256), 0, st, x3, n3);
    return hipGetLastError();
}

hipError_t launch_round_f32(double *x3, long long n3, hipStream_t st)
{
    if (n3 <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_f64_round_f32, dim3(grid_of(n3)), dim3(256), 0, st, x3, n3);
    return hipGetLastError();
}

__global__ void k_f64_to_f32(const double *__restrict__ a, long long n, float *__restrict__ b)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        b[k] = (float)a[k];
}

hipError_t launch_f64_to_f32(const double *a, long long n, float *b, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(fill_grid(n)), dim3(256), 0, st, a, n, b);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, long long n, double v, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(fill_grid(n)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, long long n, int v, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(n)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t active_lists(const unsigned char *act, int level, const int2 *rng, int nleaf,
                        int *cnt, int *off, int *ilist, void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (temp == nullptr) {
        return rocprim::exclusive_scan(nullptr, *temp_bytes, cnt, off, 0, (size_t)nleaf + 1,
                                       rocprim::plus<int>(), st);
    }
    hipLaunchKernelGGL(k_active_count, dim3(grid_of(nleaf, 4)), dim3(256), 0, st, act, level, rng, nleaf, cnt);
    hipError_t e = rocprim::exclusive_scan(temp, *temp_bytes, cnt, off, 0, (size_t)nleaf + 1,
                                           rocprim::plus<int>(), st);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_active_emit, dim3(grid_of(nleaf, 4)), dim3(256), 0, st, act, level, rng, nleaf, off, ilist);
    return hipGetLastError();
}

hipError_t csr_sort(const int *key_in, const int *val_in, int *key_out, int *val_out, int n, int key_bits,
                    void *temp, size_t *temp_bytes, hipStream_t st)
{
    return rocprim::radix_sort_pairs(temp, *temp_bytes, key_in, key_out, val_in, val_out, (size_t)n,
                                     0, key_bits, st);
}

hipError_t csr_rows(const int *key_sorted, int n, int *flag, int *rowidx, int *row_tgt, int *row_off,
                    void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (temp == nullptr)
        return rocprim::inclusive_scan(nullptr, *temp_bytes, flag, rowidx, (size_t)n, rocprim::plus<int>(), st);
    hipLaunchKernelGGL(k_row_flags, dim3(grid_of(n)), dim3(256), 0, st, key_sorted, n, flag);
    hipError_t e = rocprim::inclusive_scan(temp, *temp_bytes, flag, rowidx, (size_t)n, rocprim::plus<int>(), st);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_row_emit, dim3(grid_of(n)), dim3(256), 0, st, key_sorted, flag, rowidx, n, row_tgt, row_off);
    return hipGetLastError();
}

hipError_t leaf_order(const float4 *pos, const int2 *rng, int nleaf, int *olist, int *ocnt, int *ooff, hipStream_t st)
{
    if (nleaf <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_leaf_order, dim3((unsigned)((nleaf + 3) / 4)), dim3(256), 0, st, pos, rng, nleaf, olist, ocnt,
                       ooff);
    return hipGetLastError();
}

hipError_t src_prefix(const int *src, const int2 *rng, long long n, unsigned *nsx, void *temp, size_t *temp_bytes,
                      hipStream_t st)
{
    auto sizes = rocprim::make_transform_iterator(src, SrcSize{rng});
    if (temp == nullptr)
        return rocprim::inclusive_scan(nullptr, *temp_bytes, sizes, nsx + 1, (size_t)n, rocprim::plus<unsigned>(),
                                       st);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, nsx);
    if (n <= 0)
        return hipGetLastError();
    return rocprim::inclusive_scan(temp, *temp_bytes, sizes, nsx + 1, (size_t)n, rocprim::plus<unsigned>(), st);
}

hipError_t rows_cost(const RowArgs &ra, unsigned long long *cost, unsigned long long *total,
                     void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (temp == nullptr)
        return rocprim::reduce(nullptr, *temp_bytes, cost, total, 0ull, (size_t)ra.nrows,
                               rocprim::plus<unsigned long long>(), st);
    RowView v{ra.row_tgt, ra.row_off, ra.src, ra.rng, ra.nsx, ra.act_cnt, ra.act_off, ra.stage_ops, ra.ipl_max};
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, total + 1);
    hipLaunchKernelGGL(k_row_cost, dim3(grid_of(ra.nrows)), dim3(256), 0, st, v, ra.nrows, cost, total + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return rocprim::reduce(temp, *temp_bytes, cost, total, 0ull, (size_t)ra.nrows,
                           rocprim::plus<unsigned long long>(), st);
}

hipError_t items_count(const RowArgs &ra, long long split, int *count, int *off,
                       void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (temp == nullptr)
        return rocprim::exclusive_scan(nullptr, *temp_bytes, count, off, 0, (size_t)ra.nrows + 1,
                                       rocprim::plus<int>(), st);
    RowView v{ra.row_tgt, ra.row_off, ra.src, ra.rng, ra.nsx, ra.act_cnt, ra.act_off, ra.stage_ops, ra.ipl_max};
    hipLaunchKernelGGL(k_items_count, dim3(grid_of(ra.nrows)), dim3(256), 0, st, v, ra.nrows, split, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return rocprim::exclusive_scan(temp, *temp_bytes, count, off, 0, (size_t)ra.nrows + 1,
                                   rocprim::plus<int>(), st);
}

hipError_t items_emit(const RowArgs &ra, long long split, const int *off, Item *items, unsigned *cost, hipStream_t st)
{
    RowView v{ra.row_tgt, ra.row_off, ra.src, ra.rng, ra.nsx, ra.act_cnt, ra.act_off, ra.stage_ops, ra.ipl_max};
    hipLaunchKernelGGL(k_items_emit, dim3(grid_of(ra.nrows)), dim3(256), 0, st, v, ra.nrows, split, off, items, cost);
    return hipGetLastError();
}

hipError_t items_sort(const unsigned *cost, unsigned *cost_sorted, int *idx, int *idx_sorted, const Item *items,
                      Item *items_sorted, int n, void *temp, size_t *temp_bytes, hipStream_t st)
{
    if (temp == nullptr)
        return rocprim::radix_sort_pairs_desc(nullptr, *temp_bytes, cost, cost_sorted, idx, idx_sorted,
                                              (size_t)n, 0, 32, st);
    hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(256), 0, st, idx, n);
    hipError_t e = rocprim::radix_sort_pairs_desc(temp, *temp_bytes, cost, cost_sorted, idx, idx_sorted,
                                                  (size_t)n, 0, 32, st);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_gather_items, dim3(grid_of(n)), dim3(256), 0, st, items, idx_sorted, n, items_sorted);
    return hipGetLastError();
}

}  // namespace pp2p

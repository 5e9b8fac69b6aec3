// Interaction-list producer: kd-tree over integer particle positions and the
// symmetric dual-tree walk that emits (target leaf, source leaf) P2P pairs.
//
// This is the input side of the P2P path; it restates the reference's
// single-threaded tree build and walk so that the lists are the ones the
// reference hands to task_compute_p2p_gpu:
//   mean-split partition       src/fmm.c:168-224 (bksort_inplace, INTXYZ)
//   recursive build, numbering src/fmm.c:228-270 (build_kdtree)
//   node/leaf boxes            src/fmm.c:273-327 (center_kdtree)
//   acceptance (MAC + cutoff)  src/fmm.c:411-479
//   dual walk                  src/fmm.c:656-792 (walk_task_p2p)
//   active targets             src/fmm.c:1226-1236 (leaf holds a tag==1 particle)
// Geometry (split, widths, centres, acceptance) is evaluated in FP32 exactly
// as the reference's cpuType arithmetic, the mean in FP64.  Particles are held
// as SoA (three int32 coordinate planes + the original index) instead of the
// 40-byte Body records; the partition moves them identically.
//
// Differences from the reference, by design: leaf and node ids start at 0
// (the reference offsets them by first_leaf = NPART and first_node); the build
// and the walk run their top levels in parallel (OpenMP tasks) and concatenate
// the per-task outputs in depth-first order, which reproduces the sequential
// numbering and order exactly.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

#include <omp.h>

#include "photons_host.h"

namespace {

struct Node {
    int son[2];      // >= 0: node id; < 0: leaf id as -(leaf + 1)
    float split;
    float width[3];
    float center[3];
};

struct Leaf {
    int ipart, npart;
    int active;      // 0: holds a local particle (target), -1: ghosts only
    float width[3];
    float center[3];
};

// Allocator that leaves new elements uninitialised, so that large buffers are
// first written (and their pages first touched) by the parallel loops that
// fill them rather than by a serial zero fill.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *) noexcept {}
    template <class U, class... A>
    void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
};
using IntBuf = std::vector<int, UninitAlloc<int>>;

struct Tree {
    long long n = 0;
    double int2pos = 0.0;
    int maxleaf = 0;
    IntBuf x[3];             // coordinate planes (permuted in place)
    IntBuf idx;              // original particle index
    std::vector<Node> node;
    std::vector<Leaf> leaf;
};

inline double coord(const Tree &t, int d, long long i) { return (double)(t.int2pos * t.x[d][i]); }

inline void swap_part(Tree &t, long long a, long long b)
{
    std::swap(t.x[0][a], t.x[0][b]);
    std::swap(t.x[1][a], t.x[1][b]);
    std::swap(t.x[2][a], t.x[2][b]);
    std::swap(t.idx[a], t.idx[b]);
}

// Partition of [p, p + len) about the (sampled) mean of coordinate D; the
// scan order, the sampling stride and the left count follow src/fmm.c:168-224.
void mean_split(Tree &t, int D, long long p, long long len, long long np[2], double *split)
{
    np[0] = np[1] = 0;
    if (len < 2) {
        np[1] = len;
        return;
    }
    if (len == 2) {
        np[0] = np[1] = 1;
        *split = 0.5 * (t.int2pos * (t.x[D][p] + t.x[D][p + 1]));
        if (t.x[D][p] > t.x[D][p + 1])
            swap_part(t, p, p + 1);
        return;
    }
    long long mul = 1;
    if (len > 100000)
        mul = len / 100000;
    double mean = 0.0;
    for (long long n = 0; n < len; n += mul)
        mean += coord(t, D, p + n);
    mean /= (double)len;
    mean *= (double)mul;

    long long but = len - 1;
    for (long long n = 0; n < but; ++n) {
        if (coord(t, D, p + n) > mean) {
            while (coord(t, D, p + but) > mean && but > n)
                --but;
            swap_part(t, p + n, p + but);
        }
    }
    np[0] = but;
    np[1] = len - but;
    *split = mean;
}

// Depth-first build; ids in creation order (leaves and nodes separately).
// A subtree is built into its own arena with local ids (its root is node 0,
// leaves counted from 0); big subtrees are built by parallel OpenMP tasks and
// the arenas are concatenated in depth-first order, which yields exactly the
// ids of the sequential build.
struct Arena {
    std::vector<Node> node;
    std::vector<Leaf> leaf;
};

void append_arena(Arena &dst, Arena &src, int &son_ref)
{
    const int node_shift = (int)dst.node.size();
    const int leaf_shift = (int)dst.leaf.size();
    for (Node &b : src.node)
        for (int n = 0; n < 2; ++n)
            b.son[n] = b.son[n] >= 0 ? b.son[n] + node_shift : b.son[n] - leaf_shift;
    son_ref = node_shift;
    dst.node.insert(dst.node.end(), src.node.begin(), src.node.end());
    dst.leaf.insert(dst.leaf.end(), src.leaf.begin(), src.leaf.end());
    Arena().node.swap(src.node);
    std::vector<Leaf>().swap(src.leaf);
}

void build_arena(Tree &t, int direct, long long ip, long long len, Arena &out, int depth)
{
    out.node.push_back(Node{{-1, -1}, 0.f, {0, 0, 0}, {0, 0, 0}});
    long long np[2];
    double split = 0.0;
    mean_split(t, direct, ip, len, np, &split);
    out.node[0].split = (float)split;
    const int nd = (direct + 1) % 3;
    const long long ipc[2] = {ip, ip + np[0]};
    const bool is_node[2] = {np[0] > t.maxleaf, np[1] > t.maxleaf};
    Arena child[2];
    const bool par = depth < 12 && len > 200000;
    for (int n = 0; n < 2; ++n) {
        if (!is_node[n])
            continue;
        if (par) {
#pragma omp task default(shared) firstprivate(n)
            build_arena(t, nd, ipc[n], np[n], child[n], depth + 1);
        } else {
            build_arena(t, nd, ipc[n], np[n], child[n], depth + 1);
        }
    }
    if (par) {
#pragma omp taskwait
    }
    for (int n = 0; n < 2; ++n) {
        if (is_node[n]) {
            int ref;
            append_arena(out, child[n], ref);
            out.node[0].son[n] = ref;
        } else {
            Leaf l{};
            l.ipart = (int)ipc[n];
            l.npart = (int)np[n];
            out.node[0].son[n] = -(int)out.leaf.size() - 1;
            out.leaf.push_back(l);
        }
    }
}

void build(Tree &t, int direct, long long ip, long long len, int nthreads)
{
    Arena root;
#pragma omp parallel num_threads(nthreads)
#pragma omp single
    build_arena(t, direct, ip, len, root, 0);
    t.node.swap(root.node);
    t.leaf.swap(root.leaf);
}

// Boxes from the split planes (src/fmm.c:273-327), FP32.
void boxes(Tree &t, int direct, int inode, float left[3], float right[3])
{
    Node &b = t.node[inode];
    for (int d = 0; d < 3; ++d) {
        b.width[d] = right[d] - left[d];
        b.center[d] = (float)(0.5 * (right[d] + left[d]));
    }
    const int nd = (direct + 1) % 3;
    for (int n = 0; n < 2; ++n) {
        const int s = b.son[n];
        if (s < 0) {
            Leaf &l = t.leaf[-s - 1];
            for (int d = 0; d < 3; ++d) {
                l.width[d] = b.width[d];
                l.center[d] = b.center[d];
            }
            if (n == 0) {
                l.width[direct] = b.split - left[direct];
                l.center[direct] = (float)(0.5 * (left[direct] + b.split));
            } else {
                l.width[direct] = right[direct] - b.split;
                l.center[direct] = (float)(0.5 * (right[direct] + b.split));
            }
        } else if (n == 0) {
            const float keep = right[direct];
            right[direct] = b.split;
            boxes(t, nd, s, left, right);
            right[direct] = keep;
        } else {
            const float keep = left[direct];
            left[direct] = b.split;
            boxes(t, nd, s, left, right);
            left[direct] = keep;
        }
    }
}

// 1: well separated (multipole), 0: open, -1: beyond the cutoff (src/fmm.c:411-479)
int accept(const float wi[3], const float wj[3], const float dist[3], float open_angle, float cutoff)
{
    float w[3], mn[3];
    for (int d = 0; d < 3; ++d) {
        w[d] = (float)((wi[d] + wj[d]) * 0.5);
        mn[d] = dist[d] < 0.0f ? -dist[d] : dist[d];
        mn[d] -= w[d];
        if (mn[d] <= 0.0f)
            mn[d] = 0.0f;
    }
    const float dd2 = dist[0] * dist[0] + dist[1] * dist[1] + dist[2] * dist[2];
    if (mn[0] + mn[1] + mn[2] < 1.0e-15)
        return 0;
    const float dm2 = mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2];
    if (dm2 > cutoff * cutoff)
        return -1;
    float wmax = w[0];
    if (w[1] > wmax)
        wmax = w[1];
    if (w[2] > wmax)
        wmax = w[2];
    wmax = (float)std::fabs(wmax);
    wmax *= 2;
    return (wmax * wmax < open_angle * open_angle * dd2) ? 1 : 0;
}

struct Walker {
    const Tree &t;
    float open_angle, cutoff;
    std::vector<int> *tt, *ts;
    // well-separated pairs (acceptance 1), the tasks of walk_task_m2l
    // (src/fmm.c:812-...): cells as leaf id, or nleaf + node id; null: not kept
    std::vector<int> *mt = nullptr, *ms = nullptr;

    void emit_m2l(int im, int jm)
    {
        if (!mt)
            return;
        const int nl = (int)t.leaf.size();
        mt->push_back(im < 0 ? -im - 1 : nl + im);
        ms->push_back(jm < 0 ? -jm - 1 : nl + jm);
    }

    // a: >= 0 node id, < 0 leaf -(id+1)
    void emit(int lt, int ls)
    {
        if (t.leaf[lt].active >= 0) {
            tt->push_back(lt);
            ts->push_back(ls);
        }
    }

    void walk(int im, int jm)
    {
        if (im == jm) {
            if (im < 0) {
                emit(-im - 1, -jm - 1);
                return;
            }
            const Node &b = t.node[im];
            walk(b.son[0], b.son[0]);
            walk(b.son[0], b.son[1]);
            walk(b.son[1], b.son[0]);
            walk(b.son[1], b.son[1]);
            return;
        }
        if (im < 0 && jm < 0) {
            emit(-im - 1, -jm - 1);
            return;
        }
        float dist[3];
        if (im < 0) {
            const Leaf &a = t.leaf[-im - 1];
            const Node &b = t.node[jm];
            for (int d = 0; d < 3; ++d)
                dist[d] = a.center[d] - b.center[d];
            const int f = accept(a.width, b.width, dist, open_angle, cutoff);
            if (f == 0) {
                walk(im, b.son[0]);
                walk(im, b.son[1]);
            } else if (f == 1) {
                emit_m2l(im, jm);
            }
            return;
        }
        if (jm < 0) {
            const Node &a = t.node[im];
            const Leaf &b = t.leaf[-jm - 1];
            for (int d = 0; d < 3; ++d)
                dist[d] = a.center[d] - b.center[d];
            const int f = accept(b.width, a.width, dist, open_angle, cutoff);
            if (f == 0) {
                walk(a.son[0], jm);
                walk(a.son[1], jm);
            } else if (f == 1) {
                emit_m2l(im, jm);
            }
            return;
        }
        const Node &a = t.node[im];
        const Node &b = t.node[jm];
        for (int d = 0; d < 3; ++d)
            dist[d] = a.center[d] - b.center[d];
        const int f = accept(a.width, b.width, dist, open_angle, cutoff);
        if (f == 0) {
            if (a.width[0] + a.width[1] + a.width[2] > b.width[0] + b.width[1] + b.width[2]) {
                walk(a.son[0], jm);
                walk(a.son[1], jm);
            } else {
                walk(im, b.son[0]);
                walk(im, b.son[1]);
            }
        } else if (f == 1) {
            emit_m2l(im, jm);
        }
    }
};

// Expands the walk breadth-first into independent (im, jm) frames in
// depth-first order, until there are enough to run in parallel.
void expand(const Tree &t, float open_angle, float cutoff, std::vector<std::pair<int, int>> &frames,
            size_t target)
{
    // one level of expansion = replace each frame by its children frames in
    // the order the recursive walk would visit them; leaf-leaf frames stay
    for (int round = 0; round < 24 && frames.size() < target; ++round) {
        std::vector<std::pair<int, int>> next;
        bool grew = false;
        for (auto f : frames) {
            const int im = f.first, jm = f.second;
            if (im < 0 && jm < 0) {
                next.push_back(f);
                continue;
            }
            if (im == jm) {
                const Node &b = t.node[im];
                next.push_back({b.son[0], b.son[0]});
                next.push_back({b.son[0], b.son[1]});
                next.push_back({b.son[1], b.son[0]});
                next.push_back({b.son[1], b.son[1]});
                grew = true;
                continue;
            }
            float dist[3];
            if (im < 0) {
                const Leaf &a = t.leaf[-im - 1];
                const Node &b = t.node[jm];
                for (int d = 0; d < 3; ++d)
                    dist[d] = a.center[d] - b.center[d];
                const int fl = accept(a.width, b.width, dist, open_angle, cutoff);
                if (fl == 0) {
                    next.push_back({im, b.son[0]});
                    next.push_back({im, b.son[1]});
                    grew = true;
                } else if (fl == 1) {
                    next.push_back(f);          // a well-separated pair: the walk records it
                }
                continue;
            }
            if (jm < 0) {
                const Node &a = t.node[im];
                const Leaf &b = t.leaf[-jm - 1];
                for (int d = 0; d < 3; ++d)
                    dist[d] = a.center[d] - b.center[d];
                const int fl = accept(b.width, a.width, dist, open_angle, cutoff);
                if (fl == 0) {
                    next.push_back({a.son[0], jm});
                    next.push_back({a.son[1], jm});
                    grew = true;
                } else if (fl == 1) {
                    next.push_back(f);
                }
                continue;
            }
            const Node &a = t.node[im];
            const Node &b = t.node[jm];
            for (int d = 0; d < 3; ++d)
                dist[d] = a.center[d] - b.center[d];
            const int fl = accept(a.width, b.width, dist, open_angle, cutoff);
            if (fl == 0) {
                if (a.width[0] + a.width[1] + a.width[2] > b.width[0] + b.width[1] + b.width[2]) {
                    next.push_back({a.son[0], jm});
                    next.push_back({a.son[1], jm});
                } else {
                    next.push_back({im, b.son[0]});
                    next.push_back({im, b.son[1]});
                }
                grew = true;
            } else if (fl == 1) {
                next.push_back(f);
            }
        }
        frames.swap(next);
        if (!grew)
            break;
    }
}

}  // namespace

struct phost_lists {
    Tree tree;
    // task pairs per walk frame, in frame order; off[k] = tasks before frame k
    std::vector<std::vector<int>> ft, fs;
    std::vector<long long> off;
    // well-separated pairs per walk frame (cells: leaf id, or nleaf + node id)
    std::vector<std::vector<int>> mt, ms;
    std::vector<long long> moff{0};
    int nthreads = 1;
};

extern "C" {

phost_lists *phost_build_tree(long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                              const float bdl[3], const float bdr[3], int nthreads)
{
    if (n <= 0 || !posi || !tag || maxleaf < 1 || !(int2pos > 0.0))
        return nullptr;
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    phost_lists *h = new phost_lists();
    h->nthreads = nthreads;
    h->off.assign(1, 0);
    Tree &t = h->tree;
    t.n = n;
    t.int2pos = int2pos;
    t.maxleaf = maxleaf;
    for (int d = 0; d < 3; ++d)
        t.x[d].resize(n);
    t.idx.resize(n);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long long i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d)
            t.x[d][i] = posi[3 * i + d];
        t.idx[i] = (int)i;
    }
    build(t, 0, 0, n, nthreads);
    float left[3] = {bdl[0], bdl[1], bdl[2]}, right[3] = {bdr[0], bdr[1], bdr[2]};
    boxes(t, 0, 0, left, right);
    // targets: leaves holding a local particle (src/fmm.c:1226-1236)
    for (Leaf &l : t.leaf) {
        l.active = -1;
        for (int k = 0; k < l.npart; ++k)
            if (tag[t.idx[l.ipart + k]] == 1) {
                l.active = 0;
                break;
            }
    }
    return h;
}

int phost_walk(phost_lists *h, float open_angle, float cutoff_radius, int nthreads)
{
    if (!h)
        return -1;
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    const Tree &t = h->tree;
    // expand the top into frames and walk each frame in parallel; the frames
    // in order are the sequential depth-first order, and the per-frame lists
    // are copied out at their prefix offsets by phost_get_tasks
    std::vector<std::pair<int, int>> frames{{0, 0}};
    expand(t, open_angle, cutoff_radius, frames, (size_t)nthreads * 64);
    h->ft.assign(frames.size(), {});
    h->fs.assign(frames.size(), {});
    h->mt.assign(frames.size(), {});
    h->ms.assign(frames.size(), {});
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (long long k = 0; k < (long long)frames.size(); ++k) {
        Walker w{t, open_angle, cutoff_radius, &h->ft[k], &h->fs[k], &h->mt[k], &h->ms[k]};
        w.walk(frames[k].first, frames[k].second);
    }
    h->off.assign(frames.size() + 1, 0);
    h->moff.assign(frames.size() + 1, 0);
    for (size_t k = 0; k < frames.size(); ++k) {
        h->off[k + 1] = h->off[k] + (long long)h->ft[k].size();
        h->moff[k + 1] = h->moff[k] + (long long)h->mt[k].size();
    }
    h->nthreads = nthreads;
    return 0;
}

int phost_walk_stream(phost_lists *h, float open_angle, float cutoff_radius, int nthreads, int len_task,
                      int leaf_base, int *buf_t, int *buf_s, phost_flush_fn flush, int overlap)
{
    if (!h || len_task < 1 || !buf_t || !buf_s || !flush)
        return -1;
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    const Tree &t = h->tree;
    std::vector<std::pair<int, int>> frames{{0, 0}};
    expand(t, open_angle, cutoff_radius, frames, (size_t)nthreads * 64);
    const long long nf = (long long)frames.size();
    h->ft.assign(nf, {});
    h->fs.assign(nf, {});
    h->mt.assign(nf, {});
    h->ms.assign(nf, {});
    std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[nf]);
    for (long long k = 0; k < nf; ++k)
        done[k].store(0, std::memory_order_relaxed);
    // the walkers: frames in parallel, each marked done when its lists are whole
    auto walk_all = [&]() {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (long long k = 0; k < nf; ++k) {
            Walker w{t, open_angle, cutoff_radius, &h->ft[k], &h->fs[k], &h->mt[k], &h->ms[k]};
            w.walk(frames[k].first, frames[k].second);
            done[k].store(1, std::memory_order_release);
        }
    };
    std::thread walkers;
    if (overlap)
        walkers = std::thread(walk_all);
    else
        walk_all();
    // this thread: the frames' tasks in frame order into flushes of len_task
    int fill = 0;
    for (long long k = 0; k < nf; ++k) {
        while (!done[k].load(std::memory_order_acquire))
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        const std::vector<int> &ft = h->ft[k], &fs = h->fs[k];
        size_t i = 0;
        while (i < ft.size()) {
            const size_t m = std::min(ft.size() - i, (size_t)(len_task - fill));
            for (size_t j = 0; j < m; ++j) {
                buf_t[fill + j] = ft[i + j] + leaf_base;
                buf_s[fill + j] = fs[i + j] + leaf_base;
            }
            fill += (int)m;
            i += m;
            if (fill == len_task) {
                flush(fill, buf_t, buf_s, 0);
                fill = 0;
            }
        }
    }
    flush(fill, buf_t, buf_s, 1);
    if (walkers.joinable())
        walkers.join();
    h->off.assign(nf + 1, 0);
    h->moff.assign(nf + 1, 0);
    for (long long k = 0; k < nf; ++k) {
        h->off[k + 1] = h->off[k] + (long long)h->ft[k].size();
        h->moff[k + 1] = h->moff[k] + (long long)h->mt[k].size();
    }
    h->nthreads = nthreads;
    return 0;
}

phost_lists *phost_build(long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                         float open_angle, float cutoff_radius, const float bdl[3], const float bdr[3],
                         int nthreads)
{
    phost_lists *h = phost_build_tree(n, posi, tag, int2pos, maxleaf, bdl, bdr, nthreads);
    if (h)
        phost_walk(h, open_angle, cutoff_radius, nthreads);
    return h;
}

long long phost_count(const phost_lists *h, int what)
{
    if (!h)
        return -1;
    switch (what) {
    case PHOST_NPART: return h->tree.n;
    case PHOST_NLEAF: return (long long)h->tree.leaf.size();
    case PHOST_NNODE: return (long long)h->tree.node.size();
    case PHOST_NTASK: return h->off.back();
    case PHOST_NM2L: return h->moff.back();
    default: return -1;
    }
}

void phost_get_perm(const phost_lists *h, int *perm)
{
    std::memcpy(perm, h->tree.idx.data(), sizeof(int) * h->tree.n);
}

void phost_get_leaves(const phost_lists *h, int *ip, int *np, int *active, float *center, float *width)
{
    const auto &L = h->tree.leaf;
    for (size_t l = 0; l < L.size(); ++l) {
        if (ip) ip[l] = L[l].ipart;
        if (np) np[l] = L[l].npart;
        if (active) active[l] = L[l].active;
        for (int d = 0; d < 3; ++d) {
            if (center) center[3 * l + d] = L[l].center[d];
            if (width) width[3 * l + d] = L[l].width[d];
        }
    }
}

void phost_get_nodes(const phost_lists *h, int *son, float *center, float *width)
{
    const auto &N = h->tree.node;
    for (size_t k = 0; k < N.size(); ++k) {
        for (int n = 0; n < 2; ++n)
            if (son) son[2 * k + n] = N[k].son[n];
        for (int d = 0; d < 3; ++d) {
            if (center) center[3 * k + d] = N[k].center[d];
            if (width) width[3 * k + d] = N[k].width[d];
        }
    }
}

void phost_get_tasks(const phost_lists *h, int *task_t, int *task_s)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(h->nthreads)
    for (long long k = 0; k < (long long)h->ft.size(); ++k) {
        std::memcpy(task_t + h->off[k], h->ft[k].data(), sizeof(int) * h->ft[k].size());
        std::memcpy(task_s + h->off[k], h->fs[k].data(), sizeof(int) * h->fs[k].size());
    }
}

void phost_get_m2l(const phost_lists *h, int *task_t, int *task_s)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(h->nthreads)
    for (long long k = 0; k < (long long)h->mt.size(); ++k) {
        std::memcpy(task_t + h->moff[k], h->mt[k].data(), sizeof(int) * h->mt[k].size());
        std::memcpy(task_s + h->moff[k], h->ms[k].data(), sizeof(int) * h->ms[k].size());
    }
}

void phost_free(phost_lists *h) { delete h; }

}  // extern "C"

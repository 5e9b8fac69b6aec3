// Host context of libphotons_p2p.so: device buffers, uploads, the prepare
// step (target CSR + work items) and the launches.  One HIP stream per
// context; every call returns a pp2p_status and records a message on error.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/photons_p2p.h"
#include "forcelaw.h"
#include "p2p_build.h"
#include "p2p_launch.h"
#include "p2p_m2l.h"
#include "p2p_pm.h"
#include "p2p_scratch.h"
#include "p2p_tree.h"
#include "p2p_walk.h"

using namespace pp2p;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(PP2P_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// grow-only device buffer
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= cap)
            return hipSuccess;
        size_t want = std::max(n, cap + cap / 4);
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(want, 1));
        if (e == hipSuccess)
            cap = want;
        return e;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// grow-only pinned host buffer (staging of host-side repacking)
struct PinnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap)
            return hipSuccess;
        release();
        hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault);
        if (e == hipSuccess)
            cap = bytes;
        else
            p = nullptr;
        return e;
    }
    void release()
    {
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// fn(b, e) over [0, n) in up to 16 threads when n is large
template <typename Fn>
void parallel_ranges(long long n, Fn fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const long long nth = n < (1ll << 20) ? 1 : std::min<long long>(hw ? hw : 1, 16);
    if (nth <= 1) {
        fn(0ll, n);
        return;
    }
    std::vector<std::thread> th;
    for (long long k = 0; k < nth; ++k)
        th.emplace_back(fn, n * k / nth, n * (k + 1) / nth);
    for (auto &t : th)
        t.join();
}

}  // namespace

struct pp2p_ctx;
template <typename Fill>
static hipError_t staged_h2d(pp2p_ctx *c, void *dst, size_t units, size_t unit, size_t at, Fill fill);

struct pp2p_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;         // this device's static scratch (pp2p_compute_noalloc)
    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr, ev_p0 = nullptr, ev_p1 = nullptr;
    int num_cu = 0;

    pp2p_params params{};
    bool have_params = false;
    pp2p_tuning tuning{};
    int segs = 32;
    LawSetup law;
    DevBuf<float4> table;            // force-law table of the current law
    int table_f4 = 0;                // its length in float4 units

    // particles
    long long npart = 0;
    int nleaf = 0;
    bool have_particles = false;
    DevBuf<float4> pos;
    DevBuf<double4> ref;
    DevBuf<int2> rng;
    DevBuf<double> acc;
    DevBuf<float> acc32;        // FP32 copy of acc for pp2p_get_acc_f32_soa
    DevBuf<double> x3;          // FP64 staging of positions
    DevBuf<int> posi;           // int32 staging
    PinnedBuf h_stage;          // host packing of strided records
    std::vector<int2> h_rng;

    // active subset
    bool has_act = false;
    // i-particle order within leaves for the force kernel (no active mask set)
    DevBuf<int> olist, ocnt, ooff;
    bool order_ready = false;
    int act_level = 0;
    DevBuf<unsigned char> act;
    DevBuf<int> act_cnt, act_off, ilist;

    // task list (device copy, appended by pp2p_add_tasks)
    long long ntask = 0;
    DevBuf<int> task_t, task_s;
    long long pairs_estimate = 0;   // sum n_t n_s of the list, host side

    // CSR
    DevBuf<int> key_sorted, src, flag, rowidx, row_tgt, row_off;
    int nrows = 0;
    bool csr_ready = false;

    // items
    DevBuf<unsigned long long> row_cost, total;
    DevBuf<unsigned> nsx;               // prefix of the source sizes over the CSR
    DevBuf<int> item_count, item_off;
    DevBuf<Item> items_rows, items_q;     // row order / cost-sorted
    DevBuf<unsigned> cost, cost_sorted;
    DevBuf<int> idx, idx_sorted;
    int nitems = 0;
    long long split = 0;
    bool items_ready = false;
    std::vector<int> h_item_off;           // per row, into items_rows
    unsigned long long list_pairs = 0;

    DevBuf<unsigned> counter;
    DevBuf<unsigned long long> pair_count;
    DevBuf<unsigned char> temp;

    // device walk: the tree, and per round its frames, codes, child offsets
    // and task offsets (kept for the backward and position passes)
    struct WalkRound {
        DevBuf<int2> f;
        DevBuf<unsigned char> code;
        DevBuf<unsigned> off;
        DevBuf<int> x;
        DevBuf<int> bs, bso;            // well-separated frames per block, their scan
        long long n = 0;
    };
    DevBuf<int2> wson;
    DevBuf<float4> wnc, wnw, wlc, wlw;
    DevBuf<int> wact;
    std::vector<WalkRound> wr;
    DevBuf<int> wcnt, wpos_a, wpos_b;
    DevBuf<unsigned long long> wemit, wsep;
    // the walk's well-separated pairs as cells (leaf id, or nleaf + node id)
    DevBuf<int> wm2l_t, wm2l_s;
    long long walk_m2l_n = -1;      // -1: none since the last walk
    int walk_nnode = -1;        // nodes of the tree in the walk arrays (-1: none)

    // device kd-tree (pp2p_build_tree): per level its segments and sons, and
    // the per-particle arrays of the build; the finished tree goes to the
    // walk arrays above, its leaf ranges to rng
    struct TreeLevel {
        DevBuf<int2> seg, kid, leaf, below, idb;
        DevBuf<float> split;
        DevBuf<double> mean;
        DevBuf<int> np0, kk;
        DevBuf<float4> lo, hi;
        int nseg = 0, nleaf = 0;
    };
    std::vector<TreeLevel> tl;
    DevBuf<int> tx0, tx1, tx2, tidx, tsegof, tsmall, tsx, ttbl, ttag, tposi;
    DevBuf<int> tnodef, tleaff, tnoff, tloff;
    double tree_int2pos = 0.0;  // grid step of the last pp2p_build_tree
    bool tree_ready = false;    // walk arrays, tidx and rng hold the last pp2p_build_tree
    int tree_levels = 0;        // levels of that tree (tl[0 .. tree_levels))

    // M2L task loop and L2P: cells, their expansions, the task CSR
    DevBuf<float4> mcenter;
    DevBuf<float> mM, mL;
    bool fmm_ready = false;     // mM / mL hold the expansions of the device tree's cells
    DevBuf<int> mtt, mts, mkey, msrc, mflag, mrowidx, mrow_tgt, mrow_off;
    hipEvent_t ev_m0 = nullptr, ev_m1 = nullptr;

    // particle mesh: positions, tags, mass mesh, spectrum, potential, result
    DevBuf<int> pm_posi, pm_tag;
    DevBuf<double> pm_rho, pm_pot, pm_grad;
    DevBuf<double2> pm_spec;
    DevBuf<float> pm_acc;
    hipfftHandle pm_fwd = 0, pm_inv = 0;
    int pm_nside = 0;                  // the plans' mesh size (0: none)
    bool pm_ready = false;             // pm_rho / pm_pot hold the last call's meshes

    float last_kernel_ms = 0.f, last_prepare_ms = 0.f;
    bool kernel_timed = false, prepare_timed = false;
};

static int check_ctx(pp2p_ctx *c)
{
    if (!c)
        return fail(PP2P_ERR_ARG, "null context");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess)
        return fail(PP2P_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return PP2P_OK;
}

static int ensure_temp(pp2p_ctx *c, size_t bytes)
{
    HIPCHK(c->temp.reserve(std::max<size_t>(bytes, 256)));
    return PP2P_OK;
}

extern "C" {

const char *pp2p_version(void) { return "photons_p2p 0.1 (gfx950)"; }

int pp2p_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

const char *pp2p_last_error(void) { return g_last_error.c_str(); }

pp2p_ctx *pp2p_create(int device, int *status)
{
    int st = PP2P_OK;
    pp2p_ctx *c = nullptr;
    int n = pp2p_device_count();
    if (n <= 0) {
        st = fail(PP2P_ERR_NOGPU, "no HIP device visible");
    } else if (device < 0 || device >= n) {
        st = fail(PP2P_ERR_ARG, "device index out of range");
    } else {
        c = new pp2p_ctx();
        c->device = device;
        hipDeviceProp_t prop;
        if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
            hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&c->ev_k0) != hipSuccess || hipEventCreate(&c->ev_k1) != hipSuccess ||
            hipEventCreate(&c->ev_p0) != hipSuccess || hipEventCreate(&c->ev_p1) != hipSuccess ||
            c->counter.reserve(1) != hipSuccess) {
            st = fail(PP2P_ERR_DEVICE, "context creation failed");
            pp2p_destroy(c);
            c = nullptr;
        } else {
            c->num_cu = prop.multiProcessorCount;
            // once per device, with the device current; a failure only
            // disables pp2p_compute_noalloc
            if (scratch_address(device, &c->scratch) != hipSuccess)
                c->scratch = nullptr;
        }
    }
    if (status)
        *status = st;
    return c;
}

void pp2p_destroy(pp2p_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    c->table.release(); c->pos.release(); c->ref.release(); c->rng.release(); c->acc.release();
    c->x3.release(); c->posi.release(); c->act.release(); c->act_cnt.release(); c->act_off.release();
    c->ilist.release(); c->task_t.release(); c->task_s.release(); c->key_sorted.release(); c->src.release();
    c->flag.release(); c->rowidx.release(); c->row_tgt.release(); c->row_off.release(); c->row_cost.release();
    c->total.release(); c->item_count.release(); c->item_off.release(); c->items_rows.release();
    c->items_q.release(); c->cost.release(); c->cost_sorted.release();
    c->wson.release(); c->wnc.release(); c->wnw.release(); c->wlc.release(); c->wlw.release(); c->wact.release();
    for (auto &r : c->wr) {
        r.f.release(); r.code.release(); r.off.release(); r.x.release(); r.bs.release(); r.bso.release();
    }
    c->wcnt.release(); c->wpos_a.release(); c->wpos_b.release(); c->wemit.release();
    c->wsep.release(); c->wm2l_t.release(); c->wm2l_s.release();
    for (auto &l : c->tl) {
        l.seg.release(); l.kid.release(); l.leaf.release(); l.below.release(); l.idb.release(); l.split.release();
        l.mean.release(); l.np0.release(); l.kk.release(); l.lo.release(); l.hi.release();
    }
    c->tx0.release(); c->tx1.release(); c->tx2.release(); c->tidx.release(); c->tsegof.release();
    c->tsmall.release(); c->tsx.release(); c->ttbl.release(); c->ttag.release(); c->tposi.release();
    c->tnodef.release(); c->tleaff.release(); c->tnoff.release(); c->tloff.release();
    c->mcenter.release(); c->mM.release(); c->mL.release();
    for (DevBuf<int> *b : {&c->mtt, &c->mts, &c->mkey, &c->msrc, &c->mflag, &c->mrowidx, &c->mrow_tgt, &c->mrow_off})
        b->release();
    if (c->ev_m0) (void)hipEventDestroy(c->ev_m0);
    if (c->ev_m1) (void)hipEventDestroy(c->ev_m1);
    c->pm_posi.release(); c->pm_tag.release(); c->pm_rho.release(); c->pm_pot.release(); c->pm_grad.release(); c->pm_spec.release();
    c->pm_acc.release();
    c->olist.release(); c->ocnt.release(); c->ooff.release();
    if (c->pm_fwd) (void)hipfftDestroy(c->pm_fwd);
    if (c->pm_inv) (void)hipfftDestroy(c->pm_inv);
    c->nsx.release();
    c->idx.release(); c->idx_sorted.release(); c->counter.release(); c->temp.release();
    c->pair_count.release();
    c->acc32.release();
    c->h_stage.release();
    if (c->ev_k0) (void)hipEventDestroy(c->ev_k0);
    if (c->ev_k1) (void)hipEventDestroy(c->ev_k1);
    if (c->ev_p0) (void)hipEventDestroy(c->ev_p0);
    if (c->ev_p1) (void)hipEventDestroy(c->ev_p1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int pp2p_set_tuning(pp2p_ctx *c, const pp2p_tuning *t)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!t || t->waves_per_cu < 0 || t->split_pairs < 0 || t->ipl_max < 0 || t->ipl_max > kIplMax)
        return fail(PP2P_ERR_ARG, "invalid tuning");
    const int segs = t->segs_per_octave;
    if (segs != 0 && segs != 16 && segs != 32)
        return fail(PP2P_ERR_ARG, "tuning: segments per octave must be 0, 16 or 32");
    c->tuning = *t;
    const int new_segs = segs ? segs : 32;
    if (new_segs != c->segs) {
        c->segs = new_segs;
        if (c->have_params) {
            pp2p_params p = c->params;
            c->have_params = false;
            return pp2p_set_params(c, &p);
        }
    }
    c->items_ready = false;
    return PP2P_OK;
}

int pp2p_get_tuning(pp2p_ctx *c, pp2p_tuning *t)
{
    if (!c || !t)
        return fail(PP2P_ERR_ARG, "null argument");
    *t = c->tuning;
    return PP2P_OK;
}

int pp2p_set_params(pp2p_ctx *c, const pp2p_params *p)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!p)
        return fail(PP2P_ERR_ARG, "null params");
    if (p->law != PP2P_LAW_TABLE && p->law != PP2P_LAW_POLY && p->law != PP2P_LAW_LINEAR)
        return fail(PP2P_ERR_ARG, "unknown law");
    if (p->frame != PP2P_FRAME_LEAF && p->frame != PP2P_FRAME_ABS32)
        return fail(PP2P_ERR_ARG, "unknown frame");
    if (!(p->mass > 0.0) || !(p->box >= 0.0))
        return fail(PP2P_ERR_ARG, "mass must be > 0 and box >= 0");
    const char *err = nullptr;
    LawSetup ls;
    if (!law_setup(p->softening, p->eps, p->rs, p->ucut, c->segs, &ls, &err))
        return fail(PP2P_ERR_ARG, err);
    if (p->law == PP2P_LAW_TABLE && !ls.table_finite)
        return fail(PP2P_ERR_ARG, "force law: eps / rc too small for the table law (its cubic coefficients "
                                  "overflow FP32 below eps / rc ~ 6e-5, its octaves wrap below 2^-15.5); "
                                  "use the polynomial law");
    if (p->law == PP2P_LAW_LINEAR && !ls.lin_ok)
        return fail(PP2P_ERR_ARG, "force law: eps / rc too small for the linear law (its octaves of s "
                                  "outgrow the LDS below eps / rc ~ 2^-12.5); use the table or polynomial law");
    if (p->box > 0.0 && p->box < 2.0 * ls.rc)
        return fail(PP2P_ERR_ARG, "periodic box shorter than twice the cutoff radius");
    // frame change invalidates the converted positions
    const bool frame_changed = c->have_params && (c->params.frame != p->frame || c->law.rc != ls.rc);
    c->law = ls;
    c->params = *p;
    c->have_params = true;
    // the device copy holds the table of the chosen law, in float4 units
    std::vector<float> tab = p->law == PP2P_LAW_LINEAR ? ls.lin : ls.table;
    c->table_f4 = (int)((tab.size() + 3) / 4);
    tab.resize(4 * (size_t)c->table_f4, 0.0f);
    HIPCHK(c->table.reserve(std::max(c->table_f4, 1)));
    if (c->table_f4)
        HIPCHK(hipMemcpyAsync(c->table.p, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (frame_changed)
        c->have_particles = false;
    return PP2P_OK;
}

static int upload_leaves(pp2p_ctx *c, long long npart, int nleaf, const int *leaf_ip, const int *leaf_np)
{
    if (nleaf < 0 || (nleaf > 0 && (!leaf_ip || !leaf_np)))
        return fail(PP2P_ERR_ARG, "invalid leaf arrays");
    c->h_rng.resize(nleaf);
    for (int l = 0; l < nleaf; ++l) {
        if (leaf_np[l] < 0 || leaf_ip[l] < 0 || (long long)leaf_ip[l] + leaf_np[l] > npart)
            return fail(PP2P_ERR_ARG, "leaf " + std::to_string(l) + " outside the particle range");
        c->h_rng[l] = make_int2(leaf_ip[l], leaf_np[l]);
    }
    // Every leaf owns its particles: the leaf frames and the i-order lists are
    // written per particle, so two leaves sharing a particle would overwrite
    // each other's entries.  Non-empty ranges must be disjoint (the kd-tree
    // leaves of the reference are).
    {
        std::vector<int> ord;
        ord.reserve(nleaf);
        for (int l = 0; l < nleaf; ++l)
            if (leaf_np[l] > 0)
                ord.push_back(l);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return leaf_ip[a] < leaf_ip[b]; });
        for (size_t k = 1; k < ord.size(); ++k) {
            const int a = ord[k - 1], b = ord[k];
            if ((long long)leaf_ip[a] + leaf_np[a] > leaf_ip[b])
                return fail(PP2P_ERR_ARG, "leaves " + std::to_string(a) + " and " + std::to_string(b) +
                                              " share particles (leaf ranges must be disjoint)");
        }
    }
    HIPCHK(c->rng.reserve(std::max(nleaf, 1)));
    if (nleaf)
        HIPCHK(hipMemcpyAsync(c->rng.p, c->h_rng.data(), sizeof(int2) * nleaf, hipMemcpyHostToDevice, c->stream));
    return PP2P_OK;
}

static int finish_particles(pp2p_ctx *c, long long npart, int nleaf)
{
    HIPCHK(c->pos.reserve(std::max<long long>(npart, 1)));
    HIPCHK(c->ref.reserve(std::max(nleaf, 1)));
    HIPCHK(c->acc.reserve(std::max<long long>(3 * npart, 1)));
    const bool abs = c->params.frame == PP2P_FRAME_ABS32;
    if (abs)
        HIPCHK(launch_round_f32(c->x3.p, 3 * npart, c->stream));
    HIPCHK(launch_leaf_frame(c->x3.p, c->rng.p, nleaf, 1.0 / c->law.rc, abs, c->ref.p, c->pos.p, c->stream));
    HIPCHK(c->olist.reserve(std::max<long long>(npart, 1)));
    HIPCHK(c->ocnt.reserve(nleaf + 1));
    HIPCHK(c->ooff.reserve(nleaf + 1));
    HIPCHK(leaf_order(c->pos.p, c->rng.p, nleaf, c->olist.p, c->ocnt.p, c->ooff.p, c->stream));
    HIPCHK(launch_fill_i32(c->ocnt.p + nleaf, 1, 0, c->stream));
    c->order_ready = true;
    HIPCHK(launch_fill_f64(c->acc.p, 3 * npart, 0.0, c->stream));
    c->npart = npart;
    c->nleaf = nleaf;
    c->have_particles = true;
    c->tree_ready = false;
    c->has_act = false;
    c->ntask = 0;
    c->csr_ready = c->items_ready = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_set_particles_f64(pp2p_ctx *c, long long npart, const double *pos, int nleaf, const int *leaf_ip,
                           const int *leaf_np)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_params)
        return fail(PP2P_ERR_STATE, "set_params before set_particles");
    if (npart < 0 || npart > (1ll << 31) - 1 || (npart > 0 && !pos))
        return fail(PP2P_ERR_ARG, "invalid particle array");
    if ((st = upload_leaves(c, npart, nleaf, leaf_ip, leaf_np)))
        return st;
    HIPCHK(c->x3.reserve(std::max<long long>(3 * npart, 1)));
    if (npart)
        HIPCHK(hipMemcpyAsync(c->x3.p, pos, sizeof(double) * 3 * npart, hipMemcpyHostToDevice, c->stream));
    return finish_particles(c, npart, nleaf);
}

// Host-to-device copy of `units` records of `unit` bytes through the
// context's pinned buffer (from byte `at`, reserved by the caller) in four
// pieces: fill(stage, u0, u1) writes records [u0, u1) to stage + u0 * unit
// (called on up to 16 threads, each with its own sub-range), and the DMA of a
// piece overlaps the filling of the next.  Copies straight from pageable
// memory ran at about half the rate.  The buffer stays in use until the
// stream has passed the copies.
extern "C++" {
template <typename Fill>
static hipError_t staged_h2d(pp2p_ctx *c, void *dst, size_t units, size_t unit, size_t at, Fill fill)
{
    char *stage = static_cast<char *>(c->h_stage.p) + at;
    constexpr int kPieces = 4;
    for (int k = 0; k < kPieces; ++k) {
        const size_t u0 = units * k / kPieces, u1 = units * (k + 1) / kPieces;
        if (u1 <= u0)
            continue;
        parallel_ranges((long long)(u1 - u0), [&](long long i0, long long i1) {
            fill(stage, u0 + (size_t)i0, u0 + (size_t)i1);
        });
        hipError_t e = hipMemcpyAsync(static_cast<char *>(dst) + u0 * unit, stage + u0 * unit, (u1 - u0) * unit,
                                      hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}
}  // extern "C++"

int pp2p_set_particles_i32(pp2p_ctx *c, long long npart, const void *posi, size_t stride, double int2pos, int nleaf,
                           const int *leaf_ip, const int *leaf_np)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_params)
        return fail(PP2P_ERR_STATE, "set_params before set_particles");
    if (npart < 0 || npart > (1ll << 31) - 1 || (npart > 0 && !posi) || stride < 12 || !(int2pos > 0.0))
        return fail(PP2P_ERR_ARG, "invalid particle array");
    if ((st = upload_leaves(c, npart, nleaf, leaf_ip, leaf_np)))
        return st;
    HIPCHK(c->posi.reserve(std::max<long long>(3 * npart, 1)));
    HIPCHK(c->x3.reserve(std::max<long long>(3 * npart, 1)));
    if (npart) {
        // contiguous 12-byte rows through a pinned buffer kept by the context:
        // records wider than the three coordinates are packed on the way (as
        // the reference repacks part[i].posi, src/gpu_entry.cu:85-105); a
        // pitched copy of npart 12-byte rows from pageable memory is orders of
        // magnitude slower
        HIPCHK(c->h_stage.reserve(12 * (size_t)npart));
        const unsigned char *b = static_cast<const unsigned char *>(posi);
        HIPCHK(staged_h2d(c, c->posi.p, (size_t)npart, 12, 0, [&](char *stage, size_t u0, size_t u1) {
            if (stride == 12) {
                std::memcpy(stage + 12 * u0, b + 12 * u0, 12 * (u1 - u0));
                return;
            }
            for (size_t i = u0; i < u1; ++i)
                std::memcpy(stage + 12 * i, b + i * stride, 12);
        }));
        HIPCHK(launch_i32_to_f64(c->posi.p, 3 * npart, int2pos, false, c->x3.p, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));      // before the staging buffer is reused
    }
    return finish_particles(c, npart, nleaf);
}

int pp2p_set_active(pp2p_ctx *c, const unsigned char *a, int level)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before set_active");
    c->items_ready = false;
    if (!a) {
        c->has_act = false;
        return PP2P_OK;
    }
    if (level < 0 || level > 255)
        return fail(PP2P_ERR_ARG, "active level outside [0, 255]");
    HIPCHK(c->act.reserve(std::max<long long>(c->npart, 1)));
    HIPCHK(c->act_cnt.reserve(c->nleaf + 1));
    HIPCHK(c->act_off.reserve(c->nleaf + 1));
    HIPCHK(c->ilist.reserve(std::max<long long>(c->npart, 1)));
    if (c->npart)
        HIPCHK(hipMemcpyAsync(c->act.p, a, c->npart, hipMemcpyHostToDevice, c->stream));
    size_t tb = 0;
    HIPCHK(active_lists(c->act.p, level, c->rng.p, c->nleaf, c->act_cnt.p, c->act_off.p, c->ilist.p, nullptr, &tb,
                        c->stream));
    if ((st = ensure_temp(c, tb)))
        return st;
    tb = c->temp.cap;
    HIPCHK(launch_fill_i32(c->act_cnt.p + c->nleaf, 1, 0, c->stream));
    HIPCHK(active_lists(c->act.p, level, c->rng.p, c->nleaf, c->act_cnt.p, c->act_off.p, c->ilist.p, c->temp.p, &tb,
                        c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_act = true;
    c->act_level = level;
    return PP2P_OK;
}

static int append_tasks(pp2p_ctx *c, long long n, const int *tt, const int *ts)
{
    if (n < 0 || (n > 0 && (!tt || !ts)))
        return fail(PP2P_ERR_ARG, "invalid task arrays");
    if (c->ntask + n > (1ll << 31) - 1)
        return fail(PP2P_ERR_ARG, "more than 2^31-1 tasks in one list");
    {
        // the first task naming a leaf outside [0, nleaf), in parallel over
        // ranges of a long list
        const unsigned nl = (unsigned)c->nleaf;
        auto first_bad = [&](long long b, long long e) {
            for (long long k = b; k < e; ++k)
                if ((unsigned)tt[k] >= nl || (unsigned)ts[k] >= nl)
                    return k;
            return e;
        };
        long long bad = n;
        const unsigned hw = std::thread::hardware_concurrency();
        const int nth = n < (1ll << 20) ? 1 : (int)std::min<unsigned>(hw ? hw : 1, 16);
        if (nth <= 1) {
            bad = first_bad(0, n);
        } else {
            std::vector<long long> found(nth, n);
            std::vector<std::thread> th;
            for (int k = 0; k < nth; ++k)
                th.emplace_back([&, k] {
                    const long long b = n * k / nth, e = n * (k + 1) / nth;
                    const long long f = first_bad(b, e);
                    found[k] = f < e ? f : n;
                });
            for (auto &t : th)
                t.join();
            bad = *std::min_element(found.begin(), found.end());
        }
        if (bad < n)
            return fail(PP2P_ERR_ARG, "task " + std::to_string(bad) + " names a leaf outside [0, nleaf)");
    }
    const size_t need = (size_t)(c->ntask + n);
    if (need > c->task_t.cap) {
        // keep the existing tasks across the grow
        DevBuf<int> nt, ns;
        HIPCHK(nt.reserve(need));
        HIPCHK(ns.reserve(need));
        if (c->ntask) {
            HIPCHK(hipMemcpyAsync(nt.p, c->task_t.p, sizeof(int) * c->ntask, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(ns.p, c->task_s.p, sizeof(int) * c->ntask, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->task_t.release();
        c->task_s.release();
        c->task_t = nt;
        c->task_s = ns;
    }
    if (n) {
        HIPCHK(hipMemcpyAsync(c->task_t.p + c->ntask, tt, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->task_s.p + c->ntask, ts, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));   // the caller may reuse its buffers on return
    }
    c->ntask += n;
    c->csr_ready = c->items_ready = false;
    return PP2P_OK;
}

int pp2p_set_tasks(pp2p_ctx *c, long long n, const int *tt, const int *ts)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before set_tasks");
    c->ntask = 0;
    return append_tasks(c, n, tt, ts);
}

int pp2p_add_tasks(pp2p_ctx *c, long long n, const int *tt, const int *ts)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before add_tasks");
    return append_tasks(c, n, tt, ts);
}

static int upload_boxes(pp2p_ctx *c, DevBuf<float4> &dst, int n, const float *xyz)
{
    std::vector<float4> h((size_t)std::max(n, 1));
    for (int k = 0; k < n; ++k)
        h[k] = make_float4(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], 0.f);
    HIPCHK(dst.reserve(h.size()));
    HIPCHK(hipMemcpyAsync(dst.p, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

// The walk over the tree in the walk arrays (c->walk_nnode nodes, c->nleaf
// leaves): replaces the context's list.
static int run_walk(pp2p_ctx *c, float open_angle, float cutoff_radius, long long *ntask)
{
    const int nnode = c->walk_nnode;
    int st = PP2P_OK;
    WalkTree t{c->wson.p, c->wnc.p, c->wnw.p, c->wlc.p, c->wlw.p, c->wact.p, open_angle, cutoff_radius};

    c->ntask = 0;
    c->csr_ready = c->items_ready = false;
    auto &wr = c->wr;
    // forward: the frames of each round
    if (wr.empty())
        wr.resize(1);
    HIPCHK(wr[0].f.reserve(1));
    const int2 root = nnode > 0 ? make_int2(0, 0) : make_int2(-1, -1);
    HIPCHK(hipMemcpyAsync(wr[0].f.p, &root, sizeof(int2), hipMemcpyHostToDevice, c->stream));
    wr[0].n = 1;
    // per round: frames that emit a task (the list length is their sum) and
    // well-separated frames (the M2L pairs)
    HIPCHK(c->wemit.reserve(2 * (size_t)nnode + 4));
    HIPCHK(c->wsep.reserve(2 * (size_t)nnode + 4));
    c->walk_m2l_n = -1;
    int last = 0;
    for (int r = 0;; ++r) {
        if (r > 2 * nnode + 2)
            return fail(PP2P_ERR_ARG, "tree walk did not end");
        const long long n = wr[r].n;
        // child offsets are 32-bit: a round holds at most 2^29 frames
        if (n > (1ll << 29))
            return fail(PP2P_ERR_ARG, "tree walk round larger than 2^29 frames");
        size_t tb = 0;
        HIPCHK(walk_code(t, wr[r].f.p, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &tb, c->stream));
        if ((st = ensure_temp(c, tb)))
            return st;
        tb = c->temp.cap;
        HIPCHK(wr[r].code.reserve(n));
        HIPCHK(wr[r].off.reserve(n + 1));
        HIPCHK(wr[r].bs.reserve(walk_blocks(n)));
        HIPCHK(walk_code(t, wr[r].f.p, n, wr[r].code.p, wr[r].off.p, wr[r].bs.p, c->wemit.p + r, c->wsep.p + r,
                         c->temp.p, &tb, c->stream));
        unsigned nn = 0;
        HIPCHK(hipMemcpyAsync(&nn, wr[r].off.p + n, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (nn == 0) {
            last = r;
            break;
        }
        if ((int)wr.size() < r + 2)
            wr.resize(r + 2);
        HIPCHK(wr[r + 1].f.reserve(nn));
        HIPCHK(walk_children(t, wr[r].f.p, n, wr[r].off.p, wr[r + 1].f.p, c->stream));
        wr[r + 1].n = nn;
    }
    std::vector<unsigned long long> em(last + 1), sep(last + 1);
    HIPCHK(hipMemcpyAsync(em.data(), c->wemit.p, sizeof(unsigned long long) * (last + 1), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(sep.data(), c->wsep.p, sizeof(unsigned long long) * (last + 1), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long emitted = 0, nsep = 0;
    for (int r = 0; r <= last; ++r) {
        emitted += em[r];
        nsep += sep[r];
    }
    // the well-separated pairs, round by round in frame order, as cells
    if (nsep > (1ull << 31) - 1)
        return fail(PP2P_ERR_ARG, "more than 2^31-1 well-separated pairs");
    {
        size_t tsel = 0;
        for (int r = 0; r <= last; ++r) {
            size_t tb = 0;
            HIPCHK(walk_m2l(nullptr, nullptr, wr[r].n, nullptr, nullptr, 0, nullptr, nullptr, nullptr, &tb,
                            c->stream));
            tsel = std::max(tsel, tb);
            HIPCHK(wr[r].bso.reserve(walk_blocks(wr[r].n)));
        }
        if ((st = ensure_temp(c, tsel)))
            return st;
        HIPCHK(c->wm2l_t.reserve(std::max<unsigned long long>(nsep, 1)));
        HIPCHK(c->wm2l_s.reserve(std::max<unsigned long long>(nsep, 1)));
        unsigned long long o = 0;
        for (int r = 0; r <= last; ++r) {
            if (!sep[r])
                continue;
            size_t tb = c->temp.cap;
            HIPCHK(walk_m2l(wr[r].f.p, wr[r].code.p, wr[r].n, wr[r].bs.p, wr[r].bso.p, c->nleaf, c->wm2l_t.p + o,
                            c->wm2l_s.p + o, c->temp.p, &tb, c->stream));
            o += sep[r];
        }
        c->walk_m2l_n = (long long)nsep;
    }
    if (emitted > (1ull << 31) - 1)
        return fail(PP2P_ERR_ARG, "more than 2^31-1 tasks in one list");
    // backward: tasks at and below each frame, last round first (every
    // buffer the passes share is sized before the first launch)
    long long nmax = 1;
    size_t tmax = 0;
    for (int r = 0; r <= last; ++r) {
        size_t tb = 0;
        HIPCHK(walk_below(nullptr, nullptr, nullptr, wr[r].n, nullptr, nullptr, nullptr, &tb, c->stream));
        tmax = std::max(tmax, tb);
        nmax = std::max(nmax, wr[r].n);
        HIPCHK(wr[r].x.reserve(wr[r].n + 1));
    }
    if ((st = ensure_temp(c, tmax)))
        return st;
    HIPCHK(c->wcnt.reserve(nmax));
    HIPCHK(c->wpos_a.reserve(nmax));
    HIPCHK(c->wpos_b.reserve(nmax));
    for (int r = last; r >= 0; --r) {
        const int *xn = r < last ? wr[r + 1].x.p : nullptr;
        size_t tb = c->temp.cap;
        HIPCHK(walk_below(wr[r].code.p, wr[r].off.p, xn, wr[r].n, c->wcnt.p, wr[r].x.p, c->temp.p, &tb, c->stream));
    }
    const long long total = (long long)emitted;
    // positions: first task of each frame, round by round; emitting frames write
    HIPCHK(c->task_t.reserve(std::max(total, 1ll)));
    HIPCHK(c->task_s.reserve(std::max(total, 1ll)));
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(c->wpos_a.p, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    DevBuf<int> *pc = &c->wpos_a, *pn = &c->wpos_b;
    for (int r = 0; r <= last; ++r) {
        const int *xn = r < last ? wr[r + 1].x.p : nullptr;
        HIPCHK(walk_emit(wr[r].f.p, wr[r].code.p, wr[r].off.p, xn, wr[r].n, pc->p, r < last ? pn->p : nullptr,
                         c->task_t.p, c->task_s.p, c->stream));
        std::swap(pc, pn);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ntask = total;
    if (ntask)
        *ntask = c->ntask;
    return PP2P_OK;
}

int pp2p_walk_tasks(pp2p_ctx *c, int nnode, const int *node_son, const float *node_center, const float *node_width,
                    const float *leaf_center, const float *leaf_width, const int *leaf_active, float open_angle,
                    float cutoff_radius, long long *ntask)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before walk_tasks");
    const int nleaf = c->nleaf;
    if (nnode < 0 || (nnode > 0 && (!node_son || !node_center || !node_width)) || !leaf_center || !leaf_width ||
        !leaf_active)
        return fail(PP2P_ERR_ARG, "invalid tree arrays");
    if (nnode == 0 && nleaf != 1)
        return fail(PP2P_ERR_ARG, "a tree without nodes holds exactly one leaf");
    if (!(open_angle >= 0.f) || !(cutoff_radius >= 0.f) || std::isinf(cutoff_radius))
        return fail(PP2P_ERR_ARG, "invalid open_angle or cutoff_radius");
    // sons are created after their parent (ids increase down the tree), so
    // every round moves at least one cell of each opened frame one level down
    // and the walk ends within 2 * nnode + 2 rounds
    for (int k = 0; k < nnode; ++k)
        for (int h = 0; h < 2; ++h) {
            const int s = node_son[2 * k + h];
            if (s >= 0 ? (s <= k || s >= nnode) : (-(long long)s - 1 >= nleaf))
                return fail(PP2P_ERR_ARG, "node " + std::to_string(k) + " has an invalid son");
        }
    HIPCHK(c->wson.reserve(std::max(nnode, 1)));
    if (nnode)
        HIPCHK(hipMemcpyAsync(c->wson.p, node_son, sizeof(int2) * nnode, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c->wact.reserve(nleaf));
    HIPCHK(hipMemcpyAsync(c->wact.p, leaf_active, sizeof(int) * nleaf, hipMemcpyHostToDevice, c->stream));
    if ((st = upload_boxes(c, c->wnc, nnode, node_center)) || (st = upload_boxes(c, c->wnw, nnode, node_width)) ||
        (st = upload_boxes(c, c->wlc, nleaf, leaf_center)) || (st = upload_boxes(c, c->wlw, nleaf, leaf_width)))
        return st;
    c->walk_nnode = nnode;
    c->tree_ready = false;
    c->fmm_ready = false;
    return run_walk(c, open_angle, cutoff_radius, ntask);
}

int pp2p_build_tree(pp2p_ctx *c, long long n, const int *posi, const int *tag, double int2pos, int maxleaf,
                    const float *bdl, const float *bdr, int *nleaf_out, int *nnode_out)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_params)
        return fail(PP2P_ERR_STATE, "set_params before build_tree");
    if (n < 1 || n > (1ll << 31) - 2 || !posi || !tag || !bdl || !bdr)
        return fail(PP2P_ERR_ARG, "invalid particle arrays");
    if (maxleaf < 1 || !(int2pos > 0.0) || std::isinf(int2pos))
        return fail(PP2P_ERR_ARG, "invalid maxleaf or int2pos");
    const hipStream_t sm = c->stream;
    HIPCHK(c->tposi.reserve(3 * n));
    HIPCHK(c->ttag.reserve(n));
    for (DevBuf<int> *b : {&c->tx0, &c->tx1, &c->tx2, &c->tidx, &c->tsegof, &c->ttbl})
        HIPCHK(b->reserve(n));
    HIPCHK(c->tsmall.reserve(n + 1));
    HIPCHK(c->tsx.reserve(n + 1));
    {
        const size_t np3 = 3 * (size_t)n;
        HIPCHK(c->h_stage.reserve(sizeof(int) * (np3 + (size_t)n)));
        HIPCHK(staged_h2d(c, c->tposi.p, np3, sizeof(int), 0, [&](char *stage, size_t u0, size_t u1) {
            std::memcpy(stage + sizeof(int) * u0, posi + u0, sizeof(int) * (u1 - u0));
        }));
        HIPCHK(staged_h2d(c, c->ttag.p, (size_t)n, sizeof(int), sizeof(int) * np3, [&](char *stage, size_t u0, size_t u1) {
            std::memcpy(stage + sizeof(int) * u0, tag + u0, sizeof(int) * (u1 - u0));
        }));
    }
    int *const xs[3] = {c->tx0.p, c->tx1.p, c->tx2.p};
    const TreeScratch w{{c->tx0.p, c->tx1.p, c->tx2.p}, c->tidx.p, c->tsegof.p, c->tsmall.p, c->tsx.p, c->ttbl.p};
    HIPCHK(tree_init(c->tposi.p, n, xs, c->tidx.p, c->tsegof.p, sm));
    auto &tl = c->tl;
    if (tl.empty())
        tl.resize(1);
    HIPCHK(tl[0].seg.reserve(1));
    const int2 root = make_int2(0, (int)n);
    HIPCHK(hipMemcpyAsync(tl[0].seg.p, &root, sizeof(int2), hipMemcpyHostToDevice, sm));
    tl[0].nseg = 1;
    // levels: split every segment, then its sons become the next level or leaves
    int last = 0;
    for (int l = 0;; ++l) {
        if (l >= 1024)
            return fail(PP2P_ERR_ARG, "tree deeper than 1024 levels (many coincident particles?)");
        HIPCHK(hipStreamSynchronize(sm));      // nothing in flight while buffers grow
        auto &L = tl[l];
        const int ns = L.nseg;
        HIPCHK(L.mean.reserve(ns));
        HIPCHK(L.split.reserve(ns));
        HIPCHK(L.np0.reserve(ns));
        HIPCHK(L.kk.reserve(ns));
        HIPCHK(L.kid.reserve(ns));
        for (DevBuf<int> *b : {&c->tnodef, &c->tleaff, &c->tnoff, &c->tloff})
            HIPCHK(b->reserve(2 * (size_t)ns + 1));
        size_t t1 = 0, t2 = 0;
        HIPCHK(tree_split(w, L.seg.p, ns, l % 3, int2pos, n, nullptr, nullptr, nullptr, nullptr, nullptr, &t1, sm));
        HIPCHK(tree_children(L.seg.p, ns, nullptr, maxleaf, nullptr, nullptr, nullptr, nullptr, nullptr, &t2, sm));
        if ((st = ensure_temp(c, std::max(t1, t2))))
            return st;
        t1 = t2 = c->temp.cap;
        HIPCHK(tree_split(w, L.seg.p, ns, l % 3, int2pos, n, L.mean.p, L.split.p, L.np0.p, L.kk.p, c->temp.p, &t1,
                          sm));
        HIPCHK(tree_children(L.seg.p, ns, L.np0.p, maxleaf, c->tnodef.p, c->tleaff.p, c->tnoff.p, c->tloff.p,
                             c->temp.p, &t2, sm));
        int tot[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&tot[0], c->tnoff.p + 2 * ns, sizeof(int), hipMemcpyDeviceToHost, sm));
        HIPCHK(hipMemcpyAsync(&tot[1], c->tloff.p + 2 * ns, sizeof(int), hipMemcpyDeviceToHost, sm));
        HIPCHK(hipStreamSynchronize(sm));
        if ((int)tl.size() < l + 2)
            tl.resize(l + 2);
        auto &N = tl[l + 1];
        HIPCHK(N.seg.reserve(std::max(tot[0], 1)));
        HIPCHK(tl[l].leaf.reserve(std::max(tot[1], 1)));
        N.nseg = tot[0];
        tl[l].nleaf = tot[1];
        HIPCHK(tree_next(w, tl[l].seg.p, ns, tl[l].np0.p, c->tnodef.p, c->tnoff.p, c->tloff.p, N.seg.p, tl[l].leaf.p,
                         tl[l].kid.p, n, sm));
        if (tot[0] == 0) {
            last = l;
            break;
        }
    }
    HIPCHK(hipStreamSynchronize(sm));
    // subtree sizes, deepest level first
    for (int l = last; l >= 0; --l) {
        HIPCHK(tl[l].below.reserve(tl[l].nseg));
        HIPCHK(tl[l].idb.reserve(tl[l].nseg));
        HIPCHK(tl[l].lo.reserve(tl[l].nseg));
        HIPCHK(tl[l].hi.reserve(tl[l].nseg));
    }
    for (int l = last; l >= 0; --l)
        HIPCHK(tree_below(tl[l].kid.p, tl[l].nseg, l < last ? tl[l + 1].below.p : nullptr, tl[l].below.p, sm));
    int2 counts;
    HIPCHK(hipMemcpyAsync(&counts, tl[0].below.p, sizeof(int2), hipMemcpyDeviceToHost, sm));
    HIPCHK(hipStreamSynchronize(sm));
    const int nnode = counts.x, nleaf = counts.y;
    // ids and boxes, root first
    HIPCHK(c->wson.reserve(nnode));
    HIPCHK(c->wnc.reserve(nnode));
    HIPCHK(c->wnw.reserve(nnode));
    HIPCHK(c->wlc.reserve(nleaf));
    HIPCHK(c->wlw.reserve(nleaf));
    HIPCHK(c->wact.reserve(nleaf));
    HIPCHK(c->rng.reserve(nleaf));
    HIPCHK(c->x3.reserve(3 * n));
    const int2 idb0 = make_int2(0, 0);
    const float4 lo0 = make_float4(bdl[0], bdl[1], bdl[2], 0.f), hi0 = make_float4(bdr[0], bdr[1], bdr[2], 0.f);
    HIPCHK(hipMemcpyAsync(tl[0].idb.p, &idb0, sizeof(int2), hipMemcpyHostToDevice, sm));
    HIPCHK(hipMemcpyAsync(tl[0].lo.p, &lo0, sizeof(float4), hipMemcpyHostToDevice, sm));
    HIPCHK(hipMemcpyAsync(tl[0].hi.p, &hi0, sizeof(float4), hipMemcpyHostToDevice, sm));
    const TreeOut out{c->wson.p, c->wnc.p, c->wnw.p, c->wlc.p, c->wlw.p, c->rng.p, c->wact.p};
    auto view = [&](int l) {
        if (l > last)
            return TreeLevelView{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        auto &L = tl[l];
        return TreeLevelView{L.nseg, L.kid.p, L.leaf.p, L.split.p, L.below.p, L.idb.p, L.lo.p, L.hi.p};
    };
    for (int l = 0; l <= last; ++l)
        HIPCHK(tree_place(view(l), view(l + 1), l % 3, out, sm));
    HIPCHK(tree_finish(w, out, nleaf, c->ttag.p, int2pos, n, c->x3.p, sm));
    if ((st = finish_particles(c, n, nleaf)))
        return st;
    c->walk_nnode = nnode;
    c->tree_ready = true;
    c->tree_int2pos = int2pos;
    c->tree_levels = last + 1;
    c->fmm_ready = false;
    c->walk_m2l_n = -1;
    if (nleaf_out)
        *nleaf_out = nleaf;
    if (nnode_out)
        *nnode_out = nnode;
    return PP2P_OK;
}

int pp2p_walk_tree(pp2p_ctx *c, float open_angle, float cutoff_radius, long long *ntask)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->tree_ready)
        return fail(PP2P_ERR_STATE, "build_tree before walk_tree");
    if (!(open_angle >= 0.f) || !(cutoff_radius >= 0.f) || std::isinf(cutoff_radius))
        return fail(PP2P_ERR_ARG, "invalid open_angle or cutoff_radius");
    return run_walk(c, open_angle, cutoff_radius, ntask);
}

static int copy_xyz(pp2p_ctx *c, const float4 *src, int n, float *dst)
{
    std::vector<float4> h((size_t)n);
    if (n)
        HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k) {
        dst[3 * k] = h[k].x;
        dst[3 * k + 1] = h[k].y;
        dst[3 * k + 2] = h[k].z;
    }
    return PP2P_OK;
}

int pp2p_get_tree(pp2p_ctx *c, int *perm, int *leaf_ip, int *leaf_np, int *leaf_active, float *leaf_center,
                  float *leaf_width, int *node_son, float *node_center, float *node_width)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->tree_ready)
        return fail(PP2P_ERR_STATE, "no tree built on the device");
    const int nleaf = c->nleaf, nnode = c->walk_nnode;
    if (perm)
        HIPCHK(hipMemcpyAsync(perm, c->tidx.p, sizeof(int) * c->npart, hipMemcpyDeviceToHost, c->stream));
    if (leaf_active)
        HIPCHK(hipMemcpyAsync(leaf_active, c->wact.p, sizeof(int) * nleaf, hipMemcpyDeviceToHost, c->stream));
    if (node_son)
        HIPCHK(hipMemcpyAsync(node_son, c->wson.p, sizeof(int2) * nnode, hipMemcpyDeviceToHost, c->stream));
    if (leaf_ip || leaf_np) {
        std::vector<int2> r((size_t)nleaf);
        HIPCHK(hipMemcpyAsync(r.data(), c->rng.p, sizeof(int2) * nleaf, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int l = 0; l < nleaf; ++l) {
            if (leaf_ip)
                leaf_ip[l] = r[l].x;
            if (leaf_np)
                leaf_np[l] = r[l].y;
        }
    }
    if ((leaf_center && (st = copy_xyz(c, c->wlc.p, nleaf, leaf_center))) ||
        (leaf_width && (st = copy_xyz(c, c->wlw.p, nleaf, leaf_width))) ||
        (node_center && (st = copy_xyz(c, c->wnc.p, nnode, node_center))) ||
        (node_width && (st = copy_xyz(c, c->wnw.p, nnode, node_width))))
        return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

static int upload_centers(pp2p_ctx *c, int n, const float *xyz)
{
    return upload_boxes(c, c->mcenter, n, xyz);
}

static int m2l_core(pp2p_ctx *c, int ncell, long long ntask, const int *task_t, const int *task_s, float rs,
                    double *ms_out);
// host task arrays to the device copies mtt / mts
static int upload_m2l_tasks(pp2p_ctx *c, long long ntask, const int *task_t, const int *task_s)
{
    if (ntask <= 0)
        return PP2P_OK;
    HIPCHK(c->mtt.reserve(ntask));
    HIPCHK(c->mts.reserve(ntask));
    HIPCHK(hipMemcpyAsync(c->mtt.p, task_t, sizeof(int) * ntask, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->mts.p, task_s, sizeof(int) * ntask, hipMemcpyHostToDevice, c->stream));
    return PP2P_OK;
}

int pp2p_m2l(pp2p_ctx *c, int ncell, const float *center, const float *M, long long ntask, const int *task_t,
             const int *task_s, float rs, float *L, double *kernel_ms)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (ncell < 1 || !center || !M || !L || ntask < 0 || (ntask > 0 && (!task_t || !task_s)))
        return fail(PP2P_ERR_ARG, "invalid cell or task arrays");
    if (!(rs > 0.f) || std::isinf(rs))
        return fail(PP2P_ERR_ARG, "invalid split radius");
    if (ntask > (1ll << 31) - 1 || (long long)ncell * 20 > (1ll << 31) - 1)
        return fail(PP2P_ERR_ARG, "more than 2^31-1 tasks or coefficients");
    for (long long k = 0; k < ntask; ++k)
        if (task_t[k] < 0 || task_t[k] >= ncell || task_s[k] < 0 || task_s[k] >= ncell)
            return fail(PP2P_ERR_ARG, "m2l task " + std::to_string(k) + " names a cell outside [0, ncell)");
    const hipStream_t sm = c->stream;
    if ((st = upload_centers(c, ncell, center)))
        return st;
    const size_t nco = (size_t)20 * ncell;
    HIPCHK(c->mM.reserve(nco));
    HIPCHK(c->mL.reserve(nco));
    HIPCHK(hipMemcpyAsync(c->mM.p, M, sizeof(float) * nco, hipMemcpyHostToDevice, sm));
    HIPCHK(hipMemcpyAsync(c->mL.p, L, sizeof(float) * nco, hipMemcpyHostToDevice, sm));
    c->fmm_ready = false;
    double ms = 0.0;
    if ((st = upload_m2l_tasks(c, ntask, task_t, task_s)) || (st = m2l_core(c, ncell, ntask, c->mtt.p, c->mts.p, rs, &ms)))
        return st;
    HIPCHK(hipMemcpyAsync(L, c->mL.p, sizeof(float) * nco, hipMemcpyDeviceToHost, sm));
    HIPCHK(hipStreamSynchronize(sm));
    if (kernel_ms)
        *kernel_ms = ms;
    return PP2P_OK;
}

// The M2L loop over host task arrays on the cells in c->mcenter / c->mM,
// accumulating into c->mL; *ms: the kernel's device time.
static int m2l_core(pp2p_ctx *c, int ncell, long long ntask, const int *task_t, const int *task_s, float rs,
                    double *ms_out)
{
    int st;
    const hipStream_t sm = c->stream;
    const int n = (int)ntask;
    int nrows = 0;
    if (n > 0) {
        for (DevBuf<int> *b : {&c->mkey, &c->msrc, &c->mflag, &c->mrowidx, &c->mrow_tgt})
            HIPCHK(b->reserve(n));
        HIPCHK(c->mrow_off.reserve(n + 1));
        int bits = 1;
        while ((1ll << bits) < (long long)ncell)
            ++bits;
        size_t tb1 = 0, tb2 = 0;
        HIPCHK(csr_sort(task_t, task_s, c->mkey.p, c->msrc.p, n, bits, nullptr, &tb1, sm));
        HIPCHK(csr_rows(c->mkey.p, n, c->mflag.p, c->mrowidx.p, c->mrow_tgt.p, c->mrow_off.p, nullptr, &tb2, sm));
        if ((st = ensure_temp(c, std::max(tb1, tb2))))
            return st;
        tb1 = tb2 = c->temp.cap;
        HIPCHK(csr_sort(task_t, task_s, c->mkey.p, c->msrc.p, n, bits, c->temp.p, &tb1, sm));
        HIPCHK(csr_rows(c->mkey.p, n, c->mflag.p, c->mrowidx.p, c->mrow_tgt.p, c->mrow_off.p, c->temp.p, &tb2,
                        sm));
        HIPCHK(hipMemcpyAsync(&nrows, c->mrowidx.p + n - 1, sizeof(int), hipMemcpyDeviceToHost, sm));
        HIPCHK(hipStreamSynchronize(sm));
    }
    if (!c->ev_m0) {
        HIPCHK(hipEventCreate(&c->ev_m0));
        HIPCHK(hipEventCreate(&c->ev_m1));
    }
    const M2LArgs a{c->mcenter.p, c->mM.p, c->mrow_tgt.p, c->mrow_off.p, c->msrc.p, nrows, rs, c->mL.p};
    HIPCHK(hipEventRecord(c->ev_m0, sm));
    HIPCHK(launch_m2l(a, sm));
    HIPCHK(hipEventRecord(c->ev_m1, sm));
    HIPCHK(hipEventSynchronize(c->ev_m1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev_m0, c->ev_m1));
    *ms_out = ms;
    return PP2P_OK;
}

int pp2p_l2p(pp2p_ctx *c, const float *L, const float *center)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before l2p");
    if (c->nleaf > 0 && (!L || !center))
        return fail(PP2P_ERR_ARG, "invalid expansion arrays");
    if (c->nleaf == 0)
        return PP2P_OK;
    if ((st = upload_centers(c, c->nleaf, center)))
        return st;
    const size_t nco = (size_t)20 * c->nleaf;
    HIPCHK(c->mL.reserve(nco));
    HIPCHK(hipMemcpyAsync(c->mL.p, L, sizeof(float) * nco, hipMemcpyHostToDevice, c->stream));
    c->fmm_ready = false;
    const L2PArgs a{c->rng.p, c->mcenter.p, c->mL.p, c->x3.p, c->nleaf, c->npart, c->acc.p};
    HIPCHK(launch_l2p(a, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_fmm_short(pp2p_ctx *c, long long nm2l, const int *m2l_t, const int *m2l_s, double *device_ms)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->tree_ready)
        return fail(PP2P_ERR_STATE, "build_tree before fmm_short");
    if (!c->have_params)
        return fail(PP2P_ERR_STATE, "set_params before fmm_short");
    const int nleaf = c->nleaf, nnode = c->walk_nnode;
    const long long ncell = (long long)nleaf + nnode;
    if (ncell * 20 > (1ll << 31) - 1)
        return fail(PP2P_ERR_ARG, "too many cells");
    const bool walked = nm2l < 0;
    if (walked && (c->walk_m2l_n < 0 || m2l_t || m2l_s))
        return fail(PP2P_ERR_STATE, "no well-separated pairs from pp2p_walk_tree on this tree");
    if (!walked && nm2l > 0 && (!m2l_t || !m2l_s))
        return fail(PP2P_ERR_ARG, "invalid m2l task arrays");
    for (long long k = 0; k < nm2l; ++k)
        if (m2l_t[k] < 0 || m2l_t[k] >= ncell || m2l_s[k] < 0 || m2l_s[k] >= ncell)
            return fail(PP2P_ERR_ARG, "m2l task " + std::to_string(k) + " names a cell outside [0, ncell)");
    const hipStream_t sm = c->stream;
    HIPCHK(c->mcenter.reserve(ncell));
    HIPCHK(c->mM.reserve(20 * ncell));
    HIPCHK(c->mL.reserve(20 * ncell));
    if (!c->ev_m0) {
        HIPCHK(hipEventCreate(&c->ev_m0));
        HIPCHK(hipEventCreate(&c->ev_m1));
    }
    const FmmTree t{nleaf, nnode, c->wson.p, c->wnc.p, c->wlc.p, c->rng.p, c->x3.p};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, sm));
    // upward: P2M of the leaves, M2M level by level from the deepest
    HIPCHK(launch_cell_centers(t, c->mcenter.p, sm));
    HIPCHK(launch_p2m(t, (float)c->params.mass, c->mM.p, sm));
    for (int l = c->tree_levels - 1; l >= 0; --l)
        HIPCHK(launch_m2m_level(t, c->tl[l].idb.p, c->tl[l].nseg, c->mM.p, sm));
    HIPCHK(launch_zero_f32(c->mL.p, 20 * ncell, sm));
    HIPCHK(hipEventRecord(e1, sm));
    HIPCHK(hipEventSynchronize(e1));
    float up_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&up_ms, e0, e1));
    // M2L over the well-separated pairs
    double m2l_ms = 0.0;
    if (!walked)
        st = upload_m2l_tasks(c, nm2l, m2l_t, m2l_s);
    if (!st)
        st = walked ? m2l_core(c, (int)ncell, c->walk_m2l_n, c->wm2l_t.p, c->wm2l_s.p, (float)c->params.rs, &m2l_ms)
                    : m2l_core(c, (int)ncell, nm2l, c->mtt.p, c->mts.p, (float)c->params.rs, &m2l_ms);
    if (st) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return st;
    }
    // downward: L2L level by level from the root, then L2P into the accumulator
    HIPCHK(hipEventRecord(e0, sm));
    for (int l = 0; l < c->tree_levels; ++l)
        HIPCHK(launch_l2l_level(t, c->tl[l].idb.p, c->tl[l].nseg, c->mL.p, sm));
    const L2PArgs a{c->rng.p, c->wlc.p, c->mL.p, c->x3.p, nleaf, c->npart, c->acc.p};
    HIPCHK(launch_l2p(a, sm));
    HIPCHK(hipEventRecord(e1, sm));
    HIPCHK(hipEventSynchronize(e1));
    float down_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&down_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->fmm_ready = true;
    if (device_ms)
        *device_ms = up_ms + m2l_ms + down_ms;
    return PP2P_OK;
}

int pp2p_get_expansions(pp2p_ctx *c, float *M, float *L)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->fmm_ready)
        return fail(PP2P_ERR_STATE, "no fmm_short since the last tree or m2l call");
    const size_t n = (size_t)20 * ((size_t)c->nleaf + c->walk_nnode);
    if (M)
        HIPCHK(hipMemcpyAsync(M, c->mM.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    if (L)
        HIPCHK(hipMemcpyAsync(L, c->mL.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_get_m2l_tasks(pp2p_ctx *c, long long *n, int *m2l_t, int *m2l_s)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (c->walk_m2l_n < 0)
        return fail(PP2P_ERR_STATE, "no walk since the last tree");
    if (n)
        *n = c->walk_m2l_n;
    if (c->walk_m2l_n && m2l_t)
        HIPCHK(hipMemcpyAsync(m2l_t, c->wm2l_t.p, sizeof(int) * c->walk_m2l_n, hipMemcpyDeviceToHost, c->stream));
    if (c->walk_m2l_n && m2l_s)
        HIPCHK(hipMemcpyAsync(m2l_s, c->wm2l_s.p, sizeof(int) * c->walk_m2l_n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_get_tasks(pp2p_ctx *c, long long *ntask, int *task_t, int *task_s)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (ntask)
        *ntask = c->ntask;
    if (c->ntask && task_t)
        HIPCHK(hipMemcpyAsync(task_t, c->task_t.p, sizeof(int) * c->ntask, hipMemcpyDeviceToHost, c->stream));
    if (c->ntask && task_s)
        HIPCHK(hipMemcpyAsync(task_s, c->task_s.p, sizeof(int) * c->ntask, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

// Device storage of one list's CSR and work items: the context's grow-only
// buffers (pp2p_compute, pp2p_prepare), or consecutive slices of one device
// region reserved beforehand, so that building a list allocates nothing.
extern "C++" {
class Space {
public:
    explicit Space(pp2p_ctx *c) : c_(c) {}
    Space(pp2p_ctx *c, char *base, size_t cap) : c_(c), base_(base), cap_(cap) {}
    bool scratch() const { return base_ != nullptr; }

    // n (at least 1) elements of T: the context buffer b, grown if needed, or
    // the next slice of the scratch
    template <typename T>
    int get(DevBuf<T> &b, size_t n, T **out)
    {
        n = std::max<size_t>(n, 1);
        if (!scratch()) {
            HIPCHK(b.reserve(n));
            *out = b.p;
            return PP2P_OK;
        }
        return carve(n * sizeof(T), reinterpret_cast<void **>(out));
    }
    // storage for the next rocPRIM call only (later slices may reuse it:
    // everything runs in order on the context stream)
    int temp(size_t bytes, void **out, size_t *len)
    {
        bytes = std::max<size_t>(bytes, 256);
        if (!scratch()) {
            HIPCHK(c_->temp.reserve(bytes));
            *out = c_->temp.p;
            *len = c_->temp.cap;
            return PP2P_OK;
        }
        const size_t at = align(used_);
        if (at + bytes > cap_)
            return overflow(at + bytes);
        *out = base_ + at;
        *len = bytes;
        return PP2P_OK;
    }
    int carve(size_t bytes, void **out)
    {
        const size_t at = align(used_);
        if (at + bytes > cap_)
            return overflow(at + bytes);
        *out = base_ + at;
        used_ = at + bytes;
        return PP2P_OK;
    }

private:
    static size_t align(size_t x) { return (x + 255) & ~(size_t)255; }
    int overflow(size_t need) const
    {
        return fail(PP2P_ERR_ARG, "the list needs at least " + std::to_string(need) + " bytes of scratch, more than the " +
                                      std::to_string(cap_) + " of the preallocated region; use pp2p_compute");
    }
    pp2p_ctx *c_;
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};
}  // extern "C++"

// The CSR and work items of one list, wherever they live.
struct ListView {
    int *key_sorted = nullptr, *src = nullptr, *row_tgt = nullptr, *row_off = nullptr;
    unsigned *nsx = nullptr;
    Item *items_rows = nullptr, *items_q = nullptr;
    long long ntask = 0;
    int nrows = 0, nitems = 0;
    long long split = 0;
    unsigned long long pairs = 0;
};

static ListView ctx_view(pp2p_ctx *c)
{
    ListView L;
    L.key_sorted = c->key_sorted.p;
    L.src = c->src.p;
    L.row_tgt = c->row_tgt.p;
    L.row_off = c->row_off.p;
    L.nsx = c->nsx.p;
    L.items_rows = c->items_rows.p;
    L.items_q = c->items_q.p;
    L.ntask = c->ntask;
    L.nrows = c->nrows;
    L.nitems = c->nitems;
    L.split = c->split;
    L.pairs = c->list_pairs;
    return L;
}

// Target CSR of the list (task_t, task_s): stable sort by target, then the
// rows (src/gpu_entry.cu:132-138).  One host synchronisation (the row count).
static int csr_build(pp2p_ctx *c, Space &sp, const int *tt, const int *ts, int n, ListView &L)
{
    int *flag = nullptr, *rowidx = nullptr;
    int st;
    if ((st = sp.get(c->key_sorted, n, &L.key_sorted)) || (st = sp.get(c->src, n, &L.src)) ||
        (st = sp.get(c->flag, n, &flag)) || (st = sp.get(c->rowidx, n, &rowidx)) ||
        (st = sp.get(c->row_tgt, n, &L.row_tgt)) || (st = sp.get(c->row_off, (size_t)n + 1, &L.row_off)))
        return st;
    L.ntask = n;
    L.nrows = 0;
    if (n == 0)
        return PP2P_OK;
    int bits = 1;
    while ((1ll << bits) < (long long)c->nleaf)
        ++bits;
    size_t tb1 = 0, tb2 = 0;
    HIPCHK(csr_sort(tt, ts, L.key_sorted, L.src, n, bits, nullptr, &tb1, c->stream));
    HIPCHK(csr_rows(L.key_sorted, n, flag, rowidx, L.row_tgt, L.row_off, nullptr, &tb2, c->stream));
    void *tmp = nullptr;
    if ((st = sp.temp(std::max(tb1, tb2), &tmp, &tb1)))
        return st;
    tb2 = tb1;
    HIPCHK(csr_sort(tt, ts, L.key_sorted, L.src, n, bits, tmp, &tb1, c->stream));
    HIPCHK(csr_rows(L.key_sorted, n, flag, rowidx, L.row_tgt, L.row_off, tmp, &tb2, c->stream));
    int rows = 0;
    HIPCHK(hipMemcpyAsync(&rows, rowidx + n - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    L.nrows = rows;
    return PP2P_OK;
}

static int build_csr(pp2p_ctx *c)
{
    Space sp(c);
    ListView L;
    int st = csr_build(c, sp, c->task_t.p, c->task_s.p, (int)c->ntask, L);
    if (st)
        return st;
    c->nrows = L.nrows;
    c->csr_ready = true;
    return PP2P_OK;
}

static RowArgs row_args(pp2p_ctx *c, const ListView &L)
{
    RowArgs ra;
    ra.row_tgt = L.row_tgt;
    ra.row_off = L.row_off;
    ra.src = L.src;
    ra.rng = c->rng.p;
    ra.nsx = L.nsx;
    ra.act_cnt = c->has_act ? c->act_cnt.p : (c->order_ready ? c->ocnt.p : nullptr);
    ra.act_off = c->has_act ? c->act_off.p : (c->order_ready ? c->ooff.p : nullptr);
    ra.nrows = L.nrows;
    // chunk_model > 0: at most that many i-particles per item (default 256)
    ra.stage_ops = c->tuning.chunk_model > 0 ? (float)c->tuning.chunk_model : -1.0f;
    ra.ipl_max = c->tuning.ipl_max >= 1 && c->tuning.ipl_max <= kIplMax ? c->tuning.ipl_max : kIplMax;
    return ra;
}

static ForceVariant variant(pp2p_ctx *c, bool queue)
{
    ForceVariant v;
    v.law = c->params.law == PP2P_LAW_POLY ? LAW_POLY : c->params.law == PP2P_LAW_LINEAR ? LAW_LINEAR : LAW_LOGTAB;
    v.soft = c->params.softening == PP2P_SOFT_PLUMMER ? SOFT_PLUMMER : SOFT_REF;
    v.fold64 = c->params.frame == PP2P_FRAME_LEAF;
    v.indirect = c->has_act || c->order_ready;
    v.queue = queue;
    return v;
}

// persistent workgroups of the queue form of kernel variant v
static int queue_groups(pp2p_ctx *c, const ForceVariant &v)
{
    int per_cu = 0;
    const int wpg = force_waves_per_group(v.law);
    if (force_occupancy(v, c->table_f4, &per_cu) != hipSuccess || per_cu <= 0)
        per_cu = std::max(1, 16 / wpg);
    if (c->tuning.waves_per_cu > 0)
        per_cu = std::max(1, std::min(per_cu, (c->tuning.waves_per_cu + wpg - 1) / wpg));
    return per_cu * c->num_cu;
}

static int queue_groups(pp2p_ctx *c) { return queue_groups(c, variant(c, true)); }

static int queue_waves(pp2p_ctx *c)
{
    const ForceVariant v = variant(c, true);
    return queue_groups(c, v) * force_waves_per_group(v.law);
}

// Work items of a list: every CSR row cut into i-chunks and source ranges of
// about `split` pair-interactions.  Two copies: row order (GRID and
// PER_TARGET forms; h_item_off, when given, receives the items per row) and
// sorted by cost, largest first (QUEUE form).  Two host synchronisations
// (the pair total, the item count).
static int items_build(pp2p_ctx *c, Space &sp, ListView &L, std::vector<int> *h_item_off)
{
    const int nr = L.nrows;
    unsigned long long *row_cost = nullptr, *total = nullptr;
    int *item_count = nullptr, *item_off = nullptr;
    int st;
    if ((st = sp.get(c->row_cost, nr, &row_cost)) || (st = sp.get(c->total, 2, &total)) ||
        (st = sp.get(c->item_count, (size_t)nr + 1, &item_count)) ||
        (st = sp.get(c->item_off, (size_t)nr + 1, &item_off)))
        return st;
    L.nitems = 0;
    L.pairs = 0;
    if (h_item_off)
        h_item_off->assign(1, 0);
    if (nr == 0)
        return PP2P_OK;
    if ((st = sp.get(c->nsx, (size_t)L.ntask + 1, &L.nsx)))
        return st;
    RowArgs ra = row_args(c, L);
    size_t tb0 = 0, tb1 = 0, tb2 = 0;
    HIPCHK(src_prefix(L.src, c->rng.p, L.ntask, L.nsx, nullptr, &tb0, c->stream));
    HIPCHK(rows_cost(ra, row_cost, total, nullptr, &tb1, c->stream));
    HIPCHK(items_count(ra, 0, item_count, item_off, nullptr, &tb2, c->stream));
    void *tmp = nullptr;
    size_t tlen = 0;
    if ((st = sp.temp(std::max(tb0, std::max(tb1, tb2)), &tmp, &tlen)))
        return st;
    tb0 = tb1 = tb2 = tlen;
    HIPCHK(src_prefix(L.src, c->rng.p, L.ntask, L.nsx, tmp, &tb0, c->stream));
    HIPCHK(rows_cost(ra, row_cost, total, tmp, &tb1, c->stream));
    unsigned long long tot[2] = {0, 0};              // pairs, staged source particles
    HIPCHK(hipMemcpyAsync(tot, total, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    L.pairs = tot[0];

    long long split;
    if (c->tuning.split_pairs > 0) {
        split = c->tuning.split_pairs;
    } else {
        // about 8 items per persistent wave of the modelled work (pairs and
        // staging, as the cut counts it), within [2^14, 2^22]
        const long long waves = (long long)queue_waves(c);
        const unsigned long long work = tot[0] + (unsigned long long)kStagePairs * tot[1];
        split = (long long)(work / (unsigned long long)std::max(1ll, 8 * waves));
        split = std::max(split, 1ll << 14);
        split = std::min(split, 1ll << 22);
    }
    HIPCHK(launch_fill_i32(item_count + nr, 1, 0, c->stream));
    HIPCHK(items_count(ra, split, item_count, item_off, tmp, &tb2, c->stream));
    int nitems = 0;
    HIPCHK(hipMemcpyAsync(&nitems, item_off + nr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned *cost = nullptr, *cost_sorted = nullptr;
    int *idx = nullptr, *idx_sorted = nullptr;
    if ((st = sp.get(c->items_rows, nitems, &L.items_rows)) || (st = sp.get(c->items_q, nitems, &L.items_q)) ||
        (st = sp.get(c->cost, nitems, &cost)) || (st = sp.get(c->cost_sorted, nitems, &cost_sorted)) ||
        (st = sp.get(c->idx, nitems, &idx)) || (st = sp.get(c->idx_sorted, nitems, &idx_sorted)))
        return st;
    HIPCHK(items_emit(ra, split, item_off, L.items_rows, cost, c->stream));
    size_t tb3 = 0;
    HIPCHK(items_sort(cost, cost_sorted, idx, idx_sorted, L.items_rows, L.items_q, nitems, nullptr, &tb3,
                      c->stream));
    if ((st = sp.temp(tb3, &tmp, &tb3)))
        return st;
    HIPCHK(items_sort(cost, cost_sorted, idx, idx_sorted, L.items_rows, L.items_q, nitems, tmp, &tb3, c->stream));
    if (h_item_off) {
        h_item_off->resize(nr + 1);
        HIPCHK(hipMemcpyAsync(h_item_off->data(), item_off, sizeof(int) * (nr + 1), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    L.nitems = nitems;
    L.split = split;
    return PP2P_OK;
}

static int build_items(pp2p_ctx *c)
{
    Space sp(c);
    ListView L = ctx_view(c);
    int st = items_build(c, sp, L, &c->h_item_off);
    if (st)
        return st;
    c->list_pairs = L.pairs;
    c->nitems = L.nitems;
    c->split = L.split;
    return PP2P_OK;
}

static int prepare_form(pp2p_ctx *c, bool)
{
    int st;
    if (!c->csr_ready && (st = build_csr(c)))
        return st;
    if (!c->items_ready) {
        if ((st = build_items(c)))
            return st;
        c->items_ready = true;
    }
    return PP2P_OK;
}

int pp2p_prepare(pp2p_ctx *c)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before prepare");
    HIPCHK(hipEventRecord(c->ev_p0, c->stream));
    c->csr_ready = c->items_ready = false;
    if ((st = prepare_form(c, true)))
        return st;
    HIPCHK(hipEventRecord(c->ev_p1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev_p1));
    HIPCHK(hipEventElapsedTime(&c->last_prepare_ms, c->ev_p0, c->ev_p1));
    return PP2P_OK;
}

static ForceArgs force_args(pp2p_ctx *c, bool queue)
{
    ForceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.items = queue ? c->items_q.p : c->items_rows.p;
    a.nitems = c->nitems;
    a.item_base = 0;
    a.counter = c->counter.p;
    a.pos = c->pos.p;
    a.ref = c->ref.p;
    a.rng = c->rng.p;
    a.src = c->src.p;
    a.ilist = c->has_act ? c->ilist.p : (c->order_ready ? c->olist.p : nullptr);
    a.acc = c->acc.p;
    a.npart = c->npart;
    // the leaf frames hold positions in units of rc; the absolute FP32 frame
    // keeps the input units (the reference's float coordinates), so its sums
    // carry one more 1 / rc and its cull distance is in input units
    const bool abs = c->params.frame == PP2P_FRAME_ABS32;
    const double rc = c->law.rc;
    a.box = c->params.box / rc;
    a.out_scale = c->params.mass / (rc * rc) / (abs ? rc : 1.0);
    a.table = c->table.p;
    a.table_len = c->table_f4;
    a.cull_lim = c->tuning.no_cull ? 3.0e38f : (float)(1.0009765625 * (abs ? rc * rc : 1.0));
    return a;
}

// enqueue the launches of one evaluation (no synchronisation; graph-safe)
// QUEUE launches of fewer than 16 items per persistent wave deal the grid's
// first items out instead of starting every wave on the queue counter
static void set_deal(ForceArgs &a, int groups, int wpg)
{
    a.deal = (long long)a.nitems < 16ll * groups * wpg ? groups * wpg : 0;
}

static int enqueue_compute(pp2p_ctx *c, int exec)
{
    const bool queue = exec == PP2P_EXEC_QUEUE;
    ForceArgs a = force_args(c, queue);
    const ForceVariant v = variant(c, queue);
    const int wpg = force_waves_per_group(v.law);
    if (a.nitems == 0)
        return PP2P_OK;
    if (queue) {
        HIPCHK(launch_reset_counter(c->counter.p, c->stream));
        const int groups = std::min(queue_groups(c, v), (a.nitems + wpg - 1) / wpg);
        set_deal(a, groups, wpg);
        HIPCHK(launch_force(a, c->law.lc, v, groups, c->stream));
    } else if (exec == PP2P_EXEC_GRID) {
        HIPCHK(launch_force(a, c->law.lc, v, (a.nitems + wpg - 1) / wpg, c->stream));
    } else {
        for (int r = 0; r < c->nrows; ++r) {
            const int b = c->h_item_off[r], e = c->h_item_off[r + 1];
            if (e <= b)
                continue;
            a.item_base = b;
            a.nitems = e;
            HIPCHK(launch_force(a, c->law.lc, v, (e - b + wpg - 1) / wpg, c->stream));
        }
    }
    return PP2P_OK;
}

int pp2p_compute(pp2p_ctx *c, int exec)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before compute");
    if (exec != PP2P_EXEC_QUEUE && exec != PP2P_EXEC_GRID && exec != PP2P_EXEC_PER_TARGET)
        return fail(PP2P_ERR_ARG, "unknown execution form");
    if ((st = prepare_form(c, exec == PP2P_EXEC_QUEUE)))
        return st;
    HIPCHK(hipEventRecord(c->ev_k0, c->stream));
    if ((st = enqueue_compute(c, exec)))
        return st;
    HIPCHK(hipEventRecord(c->ev_k1, c->stream));
    c->kernel_timed = true;
    return PP2P_OK;
}

int pp2p_compute_noalloc(pp2p_ctx *c, long long ntask, const int *task_t, const int *task_s, double *acc)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before compute_noalloc");
    if (ntask < 0 || ntask > (1ll << 31) - 1 || (ntask > 0 && (!task_t || !task_s)) || !acc)
        return fail(PP2P_ERR_ARG, "compute_noalloc: 0 <= ntask < 2^31, device task arrays and a device accumulator");
    if (!c->scratch)
        return fail(PP2P_ERR_DEVICE, "compute_noalloc: the scratch array of this device has no address");
    Space sp(c, static_cast<char *>(c->scratch), kScratchBytes);
    int *flag = nullptr;
    unsigned *counter = nullptr;
    if ((st = sp.carve(sizeof(int), reinterpret_cast<void **>(&flag))) ||
        (st = sp.carve(sizeof(unsigned), reinterpret_cast<void **>(&counter))))
        return st;
    // the scratch keeps the last call's contents: clear the flag first
    HIPCHK(launch_fill_i32(flag, 1, 0, c->stream));
    HIPCHK(launch_check_tasks(task_t, task_s, ntask, c->nleaf, flag, c->stream));
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bad)
        return fail(PP2P_ERR_ARG, "compute_noalloc: a task names a leaf outside [0, nleaf)");
    ListView L;
    if ((st = csr_build(c, sp, task_t, task_s, (int)ntask, L)) || (st = items_build(c, sp, L, nullptr)))
        return st;
    ForceArgs a = force_args(c, true);
    a.items = L.items_q;
    a.nitems = L.nitems;
    a.src = L.src;
    a.counter = counter;
    a.acc = acc;
    const ForceVariant v = variant(c, true);
    const int wpg = force_waves_per_group(v.law);
    HIPCHK(hipEventRecord(c->ev_k0, c->stream));
    if (a.nitems > 0) {
        HIPCHK(launch_reset_counter(counter, c->stream));
        const int groups = std::min(queue_groups(c, v), (a.nitems + wpg - 1) / wpg);
        set_deal(a, groups, wpg);
        HIPCHK(launch_force(a, c->law.lc, v, groups, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev_k1, c->stream));
    c->kernel_timed = true;
    return PP2P_OK;
}

int pp2p_zero_acc(pp2p_ctx *c)
{
    int st = check_ctx(c);
    if (st)
        return st;
    HIPCHK(launch_fill_f64(c->acc.p, 3 * c->npart, 0.0, c->stream));
    return PP2P_OK;
}

int pp2p_run_graph(pp2p_ctx *c, int exec, int nsteps, int nreplay)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before run_graph");
    if (nsteps < 1 || nreplay < 1)
        return fail(PP2P_ERR_ARG, "nsteps and nreplay must be >= 1");
    if (exec != PP2P_EXEC_QUEUE && exec != PP2P_EXEC_GRID && exec != PP2P_EXEC_PER_TARGET)
        return fail(PP2P_ERR_ARG, "unknown execution form");
    if ((st = prepare_form(c, exec == PP2P_EXEC_QUEUE)))
        return st;
    // occupancy queries are not capturable: resolve them before capture
    (void)queue_groups(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int cst = PP2P_OK;
    for (int s = 0; s < nsteps && cst == PP2P_OK; ++s) {
        if (launch_fill_f64(c->acc.p, 3 * c->npart, 0.0, c->stream) != hipSuccess)
            cst = fail(PP2P_ERR_DEVICE, "fill launch failed under capture");
        else
            cst = enqueue_compute(c, exec);
    }
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (cst)
        return cst;
    if (e != hipSuccess)
        return fail(PP2P_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(PP2P_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    HIPCHK(hipEventRecord(c->ev_k0, c->stream));
    for (int r = 0; r < nreplay; ++r) {
        e = hipGraphLaunch(gexec, c->stream);
        if (e != hipSuccess)
            break;
    }
    hipError_t e2 = hipEventRecord(c->ev_k1, c->stream);
    hipError_t e3 = hipStreamSynchronize(c->stream);
    (void)hipGraphExecDestroy(gexec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return fail(PP2P_ERR_DEVICE, "graph replay failed");
    c->kernel_timed = true;
    return PP2P_OK;
}

int pp2p_sync(pp2p_ctx *c)
{
    int st = check_ctx(c);
    if (st)
        return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_get_acc_f64(pp2p_ctx *c, double *out)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!out && c->npart)
        return fail(PP2P_ERR_ARG, "null output");
    std::vector<double> planes((size_t)3 * c->npart);
    if (c->npart)
        HIPCHK(hipMemcpyAsync(planes.data(), c->acc.p, sizeof(double) * 3 * c->npart, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (long long i = 0; i < c->npart; ++i)
        for (int d = 0; d < 3; ++d)
            out[3 * i + d] = planes[(size_t)d * c->npart + i];
    return PP2P_OK;
}

int pp2p_get_acc_f32_soa(pp2p_ctx *c, float *out)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!out && c->npart)
        return fail(PP2P_ERR_ARG, "null output");
    // rounded to FP32 on the device: half the bytes over the bus, no host pass
    if (c->npart) {
        HIPCHK(c->acc32.reserve((size_t)3 * c->npart));
        HIPCHK(launch_f64_to_f32(c->acc.p, 3 * c->npart, c->acc32.p, c->stream));
        HIPCHK(hipMemcpyAsync(out, c->acc32.p, sizeof(float) * 3 * c->npart, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

int pp2p_list_stats(pp2p_ctx *c, long long *ntask, long long *nrows, long long *nitems, long long *npairs)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before list_stats");
    if ((st = prepare_form(c, true)))
        return st;
    if (ntask) *ntask = c->ntask;
    if (nrows) *nrows = c->nrows;
    if (nitems) *nitems = c->nitems;
    if (npairs) *npairs = (long long)c->list_pairs;
    return PP2P_OK;
}

int pp2p_get_csr(pp2p_ctx *c, long long *nrows, int *key_sorted, int *src, int *row_tgt, int *row_off)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before get_csr");
    if (!c->csr_ready && (st = build_csr(c)))
        return st;
    const size_t n = (size_t)c->ntask, r = (size_t)c->nrows;
    if (n && key_sorted)
        HIPCHK(hipMemcpyAsync(key_sorted, c->key_sorted.p, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    if (n && src)
        HIPCHK(hipMemcpyAsync(src, c->src.p, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    if (r && row_tgt)
        HIPCHK(hipMemcpyAsync(row_tgt, c->row_tgt.p, sizeof(int) * r, hipMemcpyDeviceToHost, c->stream));
    if (row_off) {
        if (r)
            HIPCHK(hipMemcpyAsync(row_off, c->row_off.p, sizeof(int) * (r + 1), hipMemcpyDeviceToHost, c->stream));
        else
            row_off[0] = 0;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nrows)
        *nrows = (long long)r;
    return PP2P_OK;
}

int pp2p_count_inside(pp2p_ctx *c, long long *n_inside)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before count_inside");
    if (!n_inside)
        return fail(PP2P_ERR_ARG, "null output");
    if ((st = prepare_form(c, true)))
        return st;
    HIPCHK(c->pair_count.reserve(1));
    unsigned long long zero = 0, total = 0;
    HIPCHK(hipMemcpyAsync(c->pair_count.p, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    ForceArgs a = force_args(c, true);
    a.count = c->pair_count.p;
    a.table_len = 0;                       // the counting pass reads no table
    if (a.nitems > 0) {
        HIPCHK(launch_reset_counter(c->counter.p, c->stream));
        ForceVariant v = variant(c, true);
        v.law = LAW_LOGTAB;                // the counting kernels have the table law's workgroup shape
        const int groups = std::min(queue_groups(c, v), (a.nitems + kWavesPerGroup - 1) / kWavesPerGroup);
        set_deal(a, groups, kWavesPerGroup);
        HIPCHK(launch_count(a, c->law.lc, v, groups, c->stream));
    }
    HIPCHK(hipMemcpyAsync(&total, c->pair_count.p, sizeof(total), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *n_inside = (long long)total;
    return PP2P_OK;
}

int pp2p_time_steps(pp2p_ctx *c, int exec, int nsteps, double *total_ms, double *force_ms)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->have_particles)
        return fail(PP2P_ERR_STATE, "set_particles before time_steps");
    if (nsteps < 1 || !total_ms)
        return fail(PP2P_ERR_ARG, "nsteps >= 1 and a result pointer are required");
    if (exec != PP2P_EXEC_QUEUE && exec != PP2P_EXEC_GRID && exec != PP2P_EXEC_PER_TARGET)
        return fail(PP2P_ERR_ARG, "unknown execution form");
    if ((st = prepare_form(c, exec == PP2P_EXEC_QUEUE)))
        return st;
    (void)queue_groups(c);
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps, nullptr);
    for (auto &e : ev)
        HIPCHK(hipEventCreate(&e));
    auto cleanup = [&]() {
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
    };
    st = PP2P_OK;
    if (hipEventRecord(c->ev_k0, c->stream) != hipSuccess)
        st = fail(PP2P_ERR_DEVICE, "event record failed");
    for (int s = 0; s < nsteps && st == PP2P_OK; ++s) {
        if (launch_fill_f64(c->acc.p, 3 * c->npart, 0.0, c->stream) != hipSuccess ||
            hipEventRecord(ev[2 * s], c->stream) != hipSuccess) {
            st = fail(PP2P_ERR_DEVICE, "step launch failed");
            break;
        }
        st = enqueue_compute(c, exec);
        if (st == PP2P_OK && hipEventRecord(ev[2 * s + 1], c->stream) != hipSuccess)
            st = fail(PP2P_ERR_DEVICE, "event record failed");
    }
    if (st == PP2P_OK && hipEventRecord(c->ev_k1, c->stream) != hipSuccess)
        st = fail(PP2P_ERR_DEVICE, "event record failed");
    if (st == PP2P_OK && hipEventSynchronize(c->ev_k1) != hipSuccess)
        st = fail(PP2P_ERR_DEVICE, "event synchronize failed");
    if (st != PP2P_OK) {
        (void)hipStreamSynchronize(c->stream);
        cleanup();
        return st;
    }
    float ms = 0.f;
    double fsum = 0.0;
    hipError_t e = hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1);
    for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
        float f = 0.f;
        e = hipEventElapsedTime(&f, ev[2 * s], ev[2 * s + 1]);
        fsum += f;
    }
    cleanup();
    if (e != hipSuccess)
        return fail(PP2P_ERR_DEVICE, std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
    *total_ms = ms;
    if (force_ms)
        *force_ms = fsum;
    c->kernel_timed = true;
    return PP2P_OK;
}

double pp2p_last_kernel_ms(pp2p_ctx *c)
{
    if (!c || !c->kernel_timed)
        return -1.0;
    if (hipEventSynchronize(c->ev_k1) != hipSuccess)
        return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1) != hipSuccess)
        return -1.0;
    return ms;
}

int pp2p_law_eval(const pp2p_params *p, int segs, long long n, const float *s, float *F, int *table_len)
{
    if (!p || n < 0 || (n > 0 && (!s || !F)))
        return fail(PP2P_ERR_ARG, "invalid law_eval arguments");
    LawSetup ls;
    const char *err = nullptr;
    if (!law_setup(p->softening, p->eps, p->rs, p->ucut, segs ? segs : 32, &ls, &err))
        return fail(PP2P_ERR_ARG, err);
    if (p->law == PP2P_LAW_LINEAR) {
        if (!ls.lin_ok)
            return fail(PP2P_ERR_ARG, "force law: eps / rc too small for the linear law");
        for (long long k = 0; k < n; ++k) {
            // the kernel's s lies in [smin, 1]
            const float sc = std::fmin(std::fmax(s[k], ls.lc.smin), 1.0f);
            F[k] = law_linear_eval(ls, sc);
        }
        if (table_len)
            *table_len = (int)(ls.lin.size() / 2);
        return PP2P_OK;
    }
    if (!ls.table_finite)
        return fail(PP2P_ERR_ARG, "force law: eps / rc too small for the table law (its cubic coefficients "
                                  "overflow FP32 below eps / rc ~ 6e-5, its octaves wrap below 2^-15.5)");
    for (long long k = 0; k < n; ++k) {
        const float sc = std::fmin(std::fmax(s[k], 0.0f), 1.0f);     // the kernel's clamp to s <= 1
        F[k] = law_logtab_eval(ls, sc);
    }
    if (table_len)
        *table_len = ls.table_len;
    return PP2P_OK;
}

double pp2p_last_prepare_ms(pp2p_ctx *c) { return c ? c->last_prepare_ms : -1.0; }

}  // extern "C"

// ------------------------------------------------------------ particle mesh

#define FFTCHK(expr)                                                                      \
    do {                                                                                  \
        hipfftResult r_ = (expr);                                                         \
        if (r_ != HIPFFT_SUCCESS)                                                         \
            return fail(PP2P_ERR_DEVICE, std::string(#expr) + ": hipFFT error " + std::to_string((int)r_)); \
    } while (0)

static int pm_plans(pp2p_ctx *c, int nside)
{
    if (c->pm_nside == nside)
        return PP2P_OK;
    if (c->pm_fwd) (void)hipfftDestroy(c->pm_fwd);
    if (c->pm_inv) (void)hipfftDestroy(c->pm_inv);
    c->pm_fwd = c->pm_inv = 0;
    c->pm_nside = 0;
    FFTCHK(hipfftPlan3d(&c->pm_fwd, nside, nside, nside, HIPFFT_D2Z));
    FFTCHK(hipfftPlan3d(&c->pm_inv, nside, nside, nside, HIPFFT_Z2D));
    FFTCHK(hipfftSetStream(c->pm_fwd, c->stream));
    FFTCHK(hipfftSetStream(c->pm_inv, c->stream));
    c->pm_nside = nside;
    return PP2P_OK;
}

static int pm_check(float box, int nside, float split_radius, float mass, double int2pos)
{
    if (nside < 8 || nside > 1024 || (nside & 1))
        return fail(PP2P_ERR_ARG, "mesh size must be even and in [8, 1024]");
    if (!(box > 0.f) || std::isinf(box) || !(int2pos > 0.0) || std::isinf(int2pos) || !(split_radius >= 0.f) ||
        std::isinf(split_radius) || std::isnan(mass) || std::isinf(mass))
        return fail(PP2P_ERR_ARG, "invalid box, grid step, split radius or mass");
    return PP2P_OK;
}

// Deposit, transforms, gradient and interpolation of a on the library stream
// (acc32: store FP32 SoA; else add into acc64); *device_ms: their device time.
static int pm_run(pp2p_ctx *c, const PmArgs &a, float split_radius, float *acc32, double *acc64, double *device_ms)
{
    int st;
    const hipStream_t sm = c->stream;
    const int nside = a.nside;
    const long long ncell = (long long)nside * nside * nside;
    const long long nspec = (long long)nside * nside * (nside / 2 + 1);
    c->pm_ready = false;
    if ((st = pm_plans(c, nside)))
        return st;
    HIPCHK(c->pm_rho.reserve(ncell));
    HIPCHK(c->pm_pot.reserve(ncell));
    HIPCHK(c->pm_grad.reserve(3 * ncell));
    HIPCHK(c->pm_spec.reserve(nspec));
    if (!c->ev_m0) {
        HIPCHK(hipEventCreate(&c->ev_m0));
        HIPCHK(hipEventCreate(&c->ev_m1));
    }
    HIPCHK(hipEventRecord(c->ev_m0, sm));
    HIPCHK(pm_zero(c->pm_rho.p, ncell, sm));
    HIPCHK(pm_deposit(a, c->pm_rho.p, sm));
    FFTCHK(hipfftExecD2Z(c->pm_fwd, c->pm_rho.p, reinterpret_cast<hipfftDoubleComplex *>(c->pm_spec.p)));
    HIPCHK(pm_green(nside, split_radius, a.box, c->pm_spec.p, sm));
    FFTCHK(hipfftExecZ2D(c->pm_inv, reinterpret_cast<hipfftDoubleComplex *>(c->pm_spec.p), c->pm_pot.p));
    HIPCHK(pm_interp(a, c->pm_pot.p, c->pm_grad.p, acc32, acc64, sm));
    HIPCHK(hipEventRecord(c->ev_m1, sm));
    HIPCHK(hipEventSynchronize(c->ev_m1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev_m0, c->ev_m1));
    if (device_ms)
        *device_ms = ms;
    c->pm_ready = true;
    return PP2P_OK;
}

int pp2p_pm_forces(pp2p_ctx *c, long long npart, const int *posi, const int *tag, double int2pos, float box,
                   int nside, float split_radius, float mass, float *acc, double *device_ms)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (npart < 0 || (npart > 0 && (!posi || !acc)))
        return fail(PP2P_ERR_ARG, "invalid particle arrays");
    if ((st = pm_check(box, nside, split_radius, mass, int2pos)))
        return st;
    if (npart > (1ll << 31) - 1)
        return fail(PP2P_ERR_ARG, "more than 2^31-1 particles");
    const hipStream_t sm = c->stream;
    HIPCHK(c->pm_posi.reserve(std::max(3 * npart, 1ll)));
    HIPCHK(c->pm_tag.reserve(std::max(npart, 1ll)));
    HIPCHK(c->pm_acc.reserve(std::max(3 * npart, 1ll)));
    if (npart > 0) {
        HIPCHK(hipMemcpyAsync(c->pm_posi.p, posi, sizeof(int) * 3 * npart, hipMemcpyHostToDevice, sm));
        if (tag)
            HIPCHK(hipMemcpyAsync(c->pm_tag.p, tag, sizeof(int) * npart, hipMemcpyHostToDevice, sm));
    }
    const int *p = c->pm_posi.p;
    PmArgs a{p, p + 1, p + 2, 3, tag ? c->pm_tag.p : nullptr, nullptr, npart, int2pos, box, mass, nside};
    if ((st = pm_run(c, a, split_radius, c->pm_acc.p, nullptr, device_ms)))
        return st;
    if (npart > 0) {
        HIPCHK(hipMemcpyAsync(acc, c->pm_acc.p, sizeof(float) * 3 * npart, hipMemcpyDeviceToHost, sm));
        HIPCHK(hipStreamSynchronize(sm));
    }
    return PP2P_OK;
}

int pp2p_pm_add_tree(pp2p_ctx *c, float box, int nside, float split_radius, float mass, double *device_ms)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->tree_ready)
        return fail(PP2P_ERR_STATE, "build_tree before pm_add_tree");
    if ((st = pm_check(box, nside, split_radius, mass, c->tree_int2pos)))
        return st;
    // the tree's particles in tree order: coordinate planes, tags through the
    // permutation; the accelerations go into the accumulator's planes
    PmArgs a{c->tx0.p, c->tx1.p, c->tx2.p, 1, c->ttag.p, c->tidx.p, c->npart, c->tree_int2pos, box, mass, nside};
    return pm_run(c, a, split_radius, nullptr, c->acc.p, device_ms);
}

int pp2p_pm_get_mesh(pp2p_ctx *c, int which, double *out)
{
    int st = check_ctx(c);
    if (st)
        return st;
    if (!c->pm_ready)
        return fail(PP2P_ERR_STATE, "pm_forces before pm_get_mesh");
    if ((which != 0 && which != 1) || !out)
        return fail(PP2P_ERR_ARG, "which must be 0 (mass) or 1 (potential), with an output array");
    const long long ncell = (long long)c->pm_nside * c->pm_nside * c->pm_nside;
    HIPCHK(hipMemcpyAsync(out, which == 0 ? c->pm_rho.p : c->pm_pot.p, sizeof(double) * ncell,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PP2P_OK;
}

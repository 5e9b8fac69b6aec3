// libp2p_hip.so: the reference's P2P entry points over the native library.
//
// The functions read the host program's particle/leaf arrays and force-law
// globals exactly where the reference's entry points read them
// (src/gpu_entry.cu:32-209, :211-386, src/gpu_active.impl.h:1-321,
// src/gpu_func.impl.h:184-275, :588-730) and drive one pp2p context:
// the first call of a step uploads the integer positions and leaf ranges,
// every flush appends its tasks, the last call launches the persistent force
// kernel over all of them.  Results stay on the device until
// devMemcpyDH_updateAcc adds them into acc[].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/p2p_hip.h"
#include "../../include/photons_p2p.h"

// State of the host program (defined there, src/global.c in the reference).
extern "C" {
extern Body *part;
extern Pack *leaf;
extern int first_leaf, last_leaf, NPART, NPACK, MAXNPACK;
extern int PROC_RANK, PROC_SIZE, NEW_RANK;
extern P2PPack *packlist;
extern float SoftenScale, MASSPART, splitRadius;
extern double INT2POS;
extern float *acc[3];
extern int *task_t, *task_s;
}

namespace {

const double kUcut = 2.8125;   // GPUCUTOFF, src/gpu_kernel.impl.h:86

struct Step {
    pp2p_ctx *ctx = nullptr;
    int device = -1;
    int call_cnt = 0;              // calls of the current step (static in the reference)
    long long tot_tasks = 0;
    int active_level = -1;         // level of the current active step, -1 none
    double last_kernel_s = 0.0;
    std::vector<float> out;        // hostmalloc_out: 3 planes of the copied results
    long long out_n = 0;
    std::vector<int> lip, lnp, tt, ts;
    // p2ppack_compute_*: per pack results, by pack index
    std::vector<std::vector<double>> pack_acc;
    bool pack_mode = false;
    int status = 0;
    std::string error;
};

Step g;

void set_error(int st, const std::string &msg)
{
    g.status = st;
    g.error = msg;
    std::fprintf(stderr, "[p2p_hip %d] %s\n", PROC_RANK, msg.c_str());
}

bool check(int st, const char *what)
{
    if (st == PP2P_OK)
        return true;
    set_error(st, std::string(what) + ": " + pp2p_last_error());
    return false;
}

// device choice of initGPU (src/gpu_func.impl.h:428-454) for one device per rank
bool ensure_ctx()
{
    if (g.ctx)
        return true;
    const int n = pp2p_device_count();
    if (n <= 0) {
        set_error(PP2P_ERR_NOGPU, "no HIP device visible");
        return false;
    }
    NEW_RANK = PROC_RANK;
    g.device = (PROC_RANK >= 0 ? PROC_RANK : 0) % n;
    int st = 0;
    g.ctx = pp2p_create(g.device, &st);
    if (!g.ctx) {
        set_error(st, std::string("pp2p_create: ") + pp2p_last_error());
        return false;
    }
    return true;
}

// first call of a step: parameters, particles and leaves of the whole rank
bool begin_step(bool with_act, int active_level)
{
    if (!ensure_ctx())
        return false;
    pp2p_params p;
    std::memset(&p, 0, sizeof(p));
    p.softening = PP2P_SOFT_REF;        // max(r, SoftenScale), src/gpu_kernel.impl.h:66-81
    p.law = PP2P_LAW_TABLE;
    p.frame = PP2P_FRAME_LEAF;
    p.eps = SoftenScale;
    p.rs = splitRadius;
    p.ucut = kUcut;
    p.mass = MASSPART;
    p.box = 0.0;                        // periodic images arrive as ghost particles
    if (!check(pp2p_set_params(g.ctx, &p), "set_params"))
        return false;
    const int nleaf = last_leaf - first_leaf;
    if (nleaf < 0 || NPART < 0 || (nleaf > 0 && !leaf) || (NPART > 0 && !part)) {
        set_error(PP2P_ERR_ARG, "Pack info error: invalid leaf or particle range");
        return false;
    }
    g.lip.resize(nleaf);
    g.lnp.resize(nleaf);
    for (int l = 0; l < nleaf; ++l) {
        g.lip[l] = leaf[first_leaf + l].ipart;
        g.lnp[l] = leaf[first_leaf + l].npart;
    }
    const void *posi = NPART > 0 ? (const void *)part[0].posi : nullptr;
    if (!check(pp2p_set_particles_i32(g.ctx, NPART, posi, sizeof(Body), INT2POS, nleaf, g.lip.data(),
                                      g.lnp.data()),
               "set_particles"))
        return false;
    if (with_act) {
        std::vector<unsigned char> a(NPART);
        for (int i = 0; i < NPART; ++i)
            a[i] = (unsigned char)part[i].act;       // uint_8b, src/gpu_active.impl.h:56-58
        if (!check(pp2p_set_active(g.ctx, a.data(), active_level), "set_active"))
            return false;
    }
    if (!check(pp2p_set_tasks(g.ctx, 0, nullptr, nullptr), "set_tasks"))
        return false;
    g.tot_tasks = 0;
    g.pack_mode = false;
    return true;
}

// leaf ids relative to first_leaf, without touching the caller's arrays
bool add_tasks(int ntask, const int *tt, const int *ts, int ileaf)
{
    if (ntask <= 0)
        return true;
    if (!tt || !ts) {
        set_error(PP2P_ERR_ARG, "null task arrays");
        return false;
    }
    g.tt.resize(ntask);
    g.ts.resize(ntask);
    for (int k = 0; k < ntask; ++k) {
        g.tt[k] = tt[k] - ileaf;
        g.ts[k] = ts[k] - ileaf;
    }
    if (!check(pp2p_add_tasks(g.ctx, ntask, g.tt.data(), g.ts.data()), "add_tasks"))
        return false;
    g.tot_tasks += ntask;
    return true;
}

// launches asynchronously: the host continues (M2L, PM) while it runs
bool finish_step()
{
    return check(pp2p_compute(g.ctx, PP2P_EXEC_QUEUE), "compute");
}

void *flush_call(int ntask, int *lt, int *ls, int last_time, bool act, int active_level)
{
    g.status = 0;
    const bool first = g.call_cnt == 0;
    ++g.call_cnt;
    P2PPack *pk = packlist;
    if (first) {
        ++NPACK;                                   // src/gpu_entry.cu:66
        if (!begin_step(act, active_level))
            return nullptr;
        g.active_level = act ? active_level : -1;
    }
    if (pk) {
        pk->ntask = ntask;
        pk->task_t = lt;
        pk->task_s = ls;
        pk->ileaf = first_leaf;
        pk->nleaf = last_leaf - first_leaf;
        pk->ipart = 0;
        pk->npart = NPART;
        pk->rankid = 0;
    }
    if (!add_tasks(ntask, lt, ls, first_leaf))
        return nullptr;
    if (last_time) {
        finish_step();
        // results of the whole step live in this pack: keep it visible to
        // devMemcpyDH_updateAcc even when the last flush carried no task
        if (pk)
            pk->ntask = (int)(g.tot_tasks > 0x7fffffff ? 0x7fffffff : g.tot_tasks);
        g.call_cnt = 0;
    }
    return nullptr;
}

void *pack_call(P2PPack *packs, int npack, bool act, int active_level)
{
    g.status = 0;
    if (npack <= 0)
        return nullptr;
    if (!packs) {
        set_error(PP2P_ERR_ARG, "null pack list");
        return nullptr;
    }
    if (!begin_step(act, active_level))
        return nullptr;
    g.pack_acc.assign(npack, {});
    std::vector<double> all;
    for (int n = 0; n < npack; ++n) {
        const P2PPack &pk = packs[n];
        if (pk.ntask == 0)
            continue;
        if (pk.ipart < 0 || pk.npart < 0 || (long long)pk.ipart + pk.npart > NPART) {
            set_error(PP2P_ERR_ARG, "pack particle range outside [0, NPART)");
            return nullptr;
        }
        if (!check(pp2p_zero_acc(g.ctx), "zero_acc") || !check(pp2p_set_tasks(g.ctx, 0, nullptr, nullptr), "set_tasks"))
            return nullptr;
        g.tot_tasks = 0;
        if (!add_tasks(pk.ntask, pk.task_t, pk.task_s, first_leaf) || !finish_step())
            return nullptr;
        all.resize((size_t)3 * NPART);
        if (!check(pp2p_get_acc_f64(g.ctx, all.data()), "get_acc"))
            return nullptr;
        auto &dst = g.pack_acc[n];
        dst.assign((size_t)3 * pk.npart, 0.0);
        for (int i = 0; i < pk.npart; ++i)
            for (int d = 0; d < 3; ++d)
                dst[(size_t)d * pk.npart + i] = all[3 * ((size_t)pk.ipart + i) + d];
    }
    g.pack_mode = true;
    g.active_level = act ? active_level : -1;
    return nullptr;
}

bool level_selected(int j, int active_level)
{
    return active_level == -1 || (active_level && part[j].act == active_level);
}

}  // namespace

extern "C" {

void *task_compute_p2p_gpu(int ntask, int *ltask_t, int *ltask_s, int last_time)
{
    return flush_call(ntask, ltask_t, ltask_s, last_time, false, 0);
}

void *task_compute_p2p_act_gpu(int ntask, int *ltask_t, int *ltask_s, int last_time, int active_level)
{
    flush_call(ntask, ltask_t, ltask_s, last_time, true, active_level);
    if (last_time && g.status == 0) {
        // src/gpu_active.impl.h:135-150: collect at once, level-filtered
        hostmalloc_out(NPART);
        devMemcpyDH_updateAcc(packlist, NPACK, active_level);
        hostfree_out();
    }
    return nullptr;
}

void *p2ppack_compute_gpu(P2PPack *packs, int npack) { return pack_call(packs, npack, false, 0); }

void *p2ppack_compute_act_gpu(P2PPack *packs, int npack, int active_level)
{
    pack_call(packs, npack, true, active_level);
    if (g.status == 0 && npack > 0) {
        hostmalloc_out(NPART);
        devMemcpyDH_updateAcc(packs, npack, active_level);
        hostfree_out();
    }
    return nullptr;
}

void hostmalloc_task(int len)
{
    task_t = (int *)std::malloc(sizeof(int) * (size_t)(len > 0 ? len : 1));
    task_s = (int *)std::malloc(sizeof(int) * (size_t)(len > 0 ? len : 1));
    if (!task_t || !task_s)
        set_error(PP2P_ERR_ALLOC, "hostmalloc_task: out of host memory");
}

void hostrealloc_task(int len)
{
    int *a = (int *)std::realloc(task_t, sizeof(int) * (size_t)(len > 0 ? len : 1));
    int *b = (int *)std::realloc(task_s, sizeof(int) * (size_t)(len > 0 ? len : 1));
    if (a)
        task_t = a;
    if (b)
        task_s = b;
    if (!a || !b)
        set_error(PP2P_ERR_ALLOC, "hostrealloc_task: out of host memory");
}

void hostfree_task(void)
{
    std::free(task_t);
    std::free(task_s);
    task_t = task_s = nullptr;
}

// The reference stages FP32 positions on the host (h_part_pos); positions
// here go to the device as integers directly, so there is nothing to stage.
void hostmalloc_in(int, int) {}
void hostfree_in(void) {}

void hostmalloc_out(int npart)
{
    g.out_n = npart > 0 ? npart : 0;
    g.out.assign((size_t)3 * g.out_n, 0.0f);
}

void hostfree_out(void)
{
    g.out.clear();
    g.out.shrink_to_fit();
    g.out_n = 0;
}

void dinfofree(void) {}

void devfree(void)
{
    if (g.ctx)
        pp2p_destroy(g.ctx);
    g.ctx = nullptr;
    g.call_cnt = 0;
    g.pack_acc.clear();
}

void updateAcc_fromHostBuff(P2PPack pack, int active_level)
{
    const long long n = pack.npart;
    if ((long long)g.out.size() < 3 * n) {
        set_error(PP2P_ERR_STATE, "updateAcc_fromHostBuff: hostmalloc_out buffer too small");
        return;
    }
    for (long long i = 0; i < n; ++i) {
        const int j = pack.ipart + (int)i;
        if (level_selected(j, active_level)) {
            acc[0][j] += g.out[i];
            acc[1][j] += g.out[i + n];
            acc[2][j] += g.out[i + 2 * n];
        }
    }
}

void devMemcpyDH_updateAcc(P2PPack *packs, int npack, int active_level)
{
    g.status = 0;
    if (!g.ctx || !packs)
        return;
    if (g.pack_mode) {
        for (int n = 0; n < npack && n < (int)g.pack_acc.size(); ++n) {
            if (packs[n].ntask == 0 || g.pack_acc[n].empty())
                continue;
            const long long np = packs[n].npart;
            if ((long long)g.out.size() < 3 * np)
                hostmalloc_out((int)np);
            for (long long k = 0; k < 3 * np; ++k)
                g.out[k] = (float)g.pack_acc[n][k];
            updateAcc_fromHostBuff(packs[n], active_level);
        }
        return;
    }
    std::vector<float> soa((size_t)3 * NPART);
    if (!check(pp2p_get_acc_f32_soa(g.ctx, soa.data()), "get_acc"))
        return;
    for (int n = 0; n < npack; ++n) {
        if (packs[n].ntask == 0)
            continue;
        const long long np = packs[n].npart, ip = packs[n].ipart;
        if (ip < 0 || ip + np > NPART) {
            set_error(PP2P_ERR_ARG, "pack particle range outside [0, NPART)");
            return;
        }
        if ((long long)g.out.size() < 3 * np)
            hostmalloc_out((int)np);
        for (int d = 0; d < 3; ++d)
            std::memcpy(&g.out[(size_t)d * np], &soa[(size_t)d * NPART + ip], sizeof(float) * np);
        updateAcc_fromHostBuff(packs[n], active_level);
    }
}

void sync_with_dev(void)
{
    if (g.ctx)
        check(pp2p_sync(g.ctx), "sync");
}

double get_last_kntime(int)
{
    if (!g.ctx)
        return 0.0;
    check(pp2p_sync(g.ctx), "sync");
    const double ms = pp2p_last_kernel_ms(g.ctx);
    g.last_kernel_s = ms > 0 ? ms * 1e-3 : 0.0;
    return g.last_kernel_s;
}

int p2p_hip_status(void) { return g.status; }
const char *p2p_hip_error(void) { return g.error.c_str(); }

}  // extern "C"

// libp2p_hip.so: the reference's P2P entry points over the native library.
//
// The functions read the host program's particle/leaf arrays and force-law
// globals exactly where the reference's entry points read them
// (src/gpu_entry.cu:32-209, :211-386, src/gpu_active.impl.h:1-321,
// src/gpu_func.impl.h:184-275, :588-730) and drive one pp2p context.
// As in the reference, every flush of the host walk (src/fmm.c:500-515) is
// computed while the walk goes on: the call copies the flush's tasks and
// returns, and a worker thread of the library uploads the positions and
// leaves (first call of a step), then each flush's tasks, builds its CSR and
// work queue and launches the persistent force kernel, the launches adding
// into one FP64 accumulator.  Flushes that queue up while the worker is busy
// go to the device as one list.  Results stay on the device until
// devMemcpyDH_updateAcc adds them into acc[]; every entry point that reads
// results or device state first waits for the queued work.
#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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

// Host array in pinned memory (hipHostMalloc), so that copies to and from
// the device run at the link's rate; pageable memory when pinning fails.
// Uninitialised: every entry used is written first.
template <typename T>
class HostArray {
public:
    HostArray() = default;
    HostArray(const HostArray &) = delete;
    HostArray &operator=(const HostArray &) = delete;
    HostArray(HostArray &&o) noexcept { swap(o); }
    HostArray &operator=(HostArray &&o) noexcept
    {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }
    ~HostArray() { release(); }
    T *get() const { return p_; }
    size_t size() const { return cap_; }
    void fit(size_t n)
    {
        if (n <= cap_)
            return;
        release();
        void *q = nullptr;
        if (hipHostMalloc(&q, sizeof(T) * n, hipHostMallocDefault) == hipSuccess && q) {
            p_ = static_cast<T *>(q);
            pinned_ = true;
        } else {
            (void)hipGetLastError();
            p_ = new T[n];
            pinned_ = false;
        }
        cap_ = n;
    }
    void release()
    {
        if (p_) {
            if (pinned_)
                (void)hipHostFree(p_);
            else
                delete[] p_;
        }
        p_ = nullptr;
        cap_ = 0;
    }

private:
    void swap(HostArray &o)
    {
        std::swap(p_, o.p_);
        std::swap(cap_, o.cap_);
        std::swap(pinned_, o.pinned_);
    }
    T *p_ = nullptr;
    size_t cap_ = 0;
    bool pinned_ = false;
};

struct Step {
    pp2p_ctx *ctx = nullptr;
    int device = -1;
    int call_cnt = 0;              // calls of the current step (static in the reference)
    long long tot_tasks = 0;       // tasks handed over in the current step (caller's count)
    int active_level = -1;         // level of the current active step, -1 none
    double kernel_ms = 0.0;        // device time of the step's force launches, summed (worker)
    bool failed = false;           // the step's set-up failed: its flushes are dropped (worker)
    std::vector<float> out;        // hostmalloc_out: 3 planes of the copied results
    long long out_n = 0;
    HostArray<float> soa;          // the rank's results, 3 planes of NPART (kept across steps)
    std::vector<int> lip, lnp;
    // p2ppack_compute_*: per pack results, by pack index
    std::vector<std::vector<double>> pack_acc;
    bool pack_mode = false;
    int status = 0;
    std::string error;
};

Step g;
std::mutex g_err;                  // the caller's thread and the worker both record errors

void set_error(int st, const std::string &msg)
{
    std::lock_guard<std::mutex> lk(g_err);
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
    p.frame = PP2P_FRAME_LEAF;
    p.eps = SoftenScale;
    p.rs = splitRadius;
    p.ucut = kUcut;
    p.mass = MASSPART;
    p.box = 0.0;                        // periodic images arrive as ghost particles
    // The table law refuses softening below eps / rc ~ 6e-5 (the cubic
    // coefficients of its smallest octaves leave the FP32 range); the
    // polynomial law has no such floor and takes over there.
    p.law = PP2P_LAW_TABLE;
    if (pp2p_law_eval(&p, 0, 0, nullptr, nullptr, nullptr) != PP2P_OK)
        p.law = PP2P_LAW_POLY;
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
    return check(pp2p_set_tasks(g.ctx, 0, nullptr, nullptr), "set_tasks");
}

// Task storage of one flush, kept for reuse: a flush of LEN_TASK tasks is
// tens of MB, and fresh memory would be pinned (or page-faulted) every time.
struct TaskBuf {
    HostArray<int> t, s;
    size_t n = 0;
    void fit(size_t want)
    {
        t.fit(want);
        s.fit(want);
        n = want;
    }
};

class BufPool {
public:
    TaskBuf take()
    {
        std::lock_guard<std::mutex> lk(m_);
        if (free_.empty())
            return TaskBuf{};
        TaskBuf b = std::move(free_.back());
        free_.pop_back();
        return b;
    }
    void give(TaskBuf &&b)
    {
        std::lock_guard<std::mutex> lk(m_);
        if (free_.size() < 8)
            free_.push_back(std::move(b));
    }

private:
    std::mutex m_;
    std::vector<TaskBuf> free_;
};

BufPool g_pool;

// One unit of device work, handed from the caller's thread to the worker.
struct Job {
    enum Kind { BEGIN, FLUSH } kind = FLUSH;
    bool act = false;              // BEGIN: active step
    int level = 0;                 // BEGIN: active level
    TaskBuf tasks;                 // FLUSH: leaf pairs, ids relative to first_leaf
};

// Runs on the worker.  BEGIN: the step's set-up.  A batch of FLUSH jobs: one
// list, one launch, added into the accumulator; the launch's device time is
// added to the step's.
void run_jobs(std::vector<Job> &batch)
{
    if (batch[0].kind == Job::BEGIN) {
        g.kernel_ms = 0.0;
        g.failed = !begin_step(batch[0].act, batch[0].level);
        return;
    }
    if (g.failed || !g.ctx)
        return;
    bool ok = check(pp2p_set_tasks(g.ctx, (long long)batch[0].tasks.n, batch[0].tasks.t.get(),
                                   batch[0].tasks.s.get()),
                    "set_tasks");
    for (size_t k = 1; ok && k < batch.size(); ++k)
        ok = check(pp2p_add_tasks(g.ctx, (long long)batch[k].tasks.n, batch[k].tasks.t.get(), batch[k].tasks.s.get()),
                   "add_tasks");
    // the tasks are on the device now: their host storage goes back to the pool
    for (Job &j : batch)
        g_pool.give(std::move(j.tasks));
    if (ok)
        ok = check(pp2p_compute(g.ctx, PP2P_EXEC_QUEUE), "compute");
    if (ok) {
        const double ms = pp2p_last_kernel_ms(g.ctx);        // waits for this launch
        if (ms > 0)
            g.kernel_ms += ms;
    }
    if (!ok)
        g.failed = true;
}

// The library's worker thread and its job queue.  The caller's thread posts
// jobs and returns; drain() waits until every posted job has run, which
// orders the worker's writes to g before the caller's reads.
class Worker {
public:
    ~Worker() { stop(); }
    void post(Job &&j)
    {
        std::lock_guard<std::mutex> lk(m_);
        if (!th_.joinable()) {
            quit_ = false;
            th_ = std::thread([this] { loop(); });
        }
        q_.push_back(std::move(j));
        cv_.notify_one();
    }
    void drain()
    {
        std::unique_lock<std::mutex> lk(m_);
        idle_.wait(lk, [this] { return q_.empty() && !busy_; });
    }
    void stop()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            if (!th_.joinable())
                return;
            quit_ = true;
            cv_.notify_one();
        }
        th_.join();
    }

private:
    void loop()
    {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return quit_ || !q_.empty(); });
            if (q_.empty())
                return;                                  // quit with nothing left to run
            std::vector<Job> batch;
            batch.push_back(std::move(q_.front()));
            q_.pop_front();
            // flushes waiting behind a flush go to the device with it
            while (batch[0].kind == Job::FLUSH && !q_.empty() && q_.front().kind == Job::FLUSH) {
                batch.push_back(std::move(q_.front()));
                q_.pop_front();
            }
            busy_ = true;
            lk.unlock();
            run_jobs(batch);
            lk.lock();
            busy_ = false;
            if (q_.empty())
                idle_.notify_all();
        }
    }
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_, idle_;
    std::deque<Job> q_;
    bool busy_ = false, quit_ = false;
};

// constructed after g, so destroyed (joined) before it
Worker worker;

void *flush_call(int ntask, int *lt, int *ls, int last_time, bool act, int active_level)
{
    const bool first = g.call_cnt == 0;
    ++g.call_cnt;
    P2PPack *pk = packlist;
    if (first) {
        worker.drain();                                  // nothing of an earlier step is left
        g.status = 0;
        ++NPACK;                                         // src/gpu_entry.cu:66
        g.tot_tasks = 0;
        g.pack_mode = false;
        g.active_level = act ? active_level : -1;
        Job b;
        b.kind = Job::BEGIN;
        b.act = act;
        b.level = active_level;
        worker.post(std::move(b));
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
    if (ntask > 0) {
        if (!lt || !ls) {
            worker.drain();
            set_error(PP2P_ERR_ARG, "null task arrays");
        } else {
            // leaf ids relative to first_leaf, copied: the caller refills its
            // buffers as soon as the call returns, and they are not rebased
            Job f;
            f.kind = Job::FLUSH;
            f.tasks = g_pool.take();
            f.tasks.fit((size_t)ntask);
            int *ft = f.tasks.t.get(), *fs = f.tasks.s.get();
            const int base = first_leaf;
            auto rebase = [=](long long b, long long e) {
                for (long long k = b; k < e; ++k) {
                    ft[k] = lt[k] - base;
                    fs[k] = ls[k] - base;
                }
            };
            // the walk waits for this copy: spread a long flush over threads
            const unsigned hw = std::thread::hardware_concurrency();
            const long long nth = ntask < (1 << 20) ? 1 : std::min<long long>(hw ? hw : 1, 8);
            if (nth <= 1) {
                rebase(0, ntask);
            } else {
                std::vector<std::thread> th;
                for (long long k = 0; k < nth; ++k)
                    th.emplace_back(rebase, ntask * k / nth, ntask * (k + 1) / nth);
                for (auto &t : th)
                    t.join();
            }
            worker.post(std::move(f));
            g.tot_tasks += ntask;
        }
    }
    if (last_time) {
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
    worker.drain();
    g.status = 0;
    if (npack <= 0)
        return nullptr;
    if (!packs) {
        set_error(PP2P_ERR_ARG, "null pack list");
        return nullptr;
    }
    if (!begin_step(act, active_level))
        return nullptr;
    g.tot_tasks = 0;
    g.kernel_ms = 0.0;
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
        if (!pk.task_t || !pk.task_s) {
            set_error(PP2P_ERR_ARG, "null task arrays");
            return nullptr;
        }
        std::vector<int> tt(pk.ntask), ts(pk.ntask);
        for (int k = 0; k < pk.ntask; ++k) {
            tt[k] = pk.task_t[k] - first_leaf;
            ts[k] = pk.task_s[k] - first_leaf;
        }
        if (!check(pp2p_zero_acc(g.ctx), "zero_acc") ||
            !check(pp2p_set_tasks(g.ctx, pk.ntask, tt.data(), ts.data()), "set_tasks") ||
            !check(pp2p_compute(g.ctx, PP2P_EXEC_QUEUE), "compute"))
            return nullptr;
        const double ms = pp2p_last_kernel_ms(g.ctx);
        if (ms > 0)
            g.kernel_ms += ms;
        g.tot_tasks += pk.ntask;
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

// acc[d][ip + i] += x[d][i] for the selected particles of [ip, ip + n), in
// parallel over disjoint ranges when n is large
void add_results(long long ip, long long n, const float *x, const float *y, const float *z, int active_level)
{
    auto run = [&](long long b, long long e) {
        for (long long i = b; i < e; ++i) {
            const int j = (int)(ip + i);
            if (level_selected(j, active_level)) {
                acc[0][j] += x[i];
                acc[1][j] += y[i];
                acc[2][j] += z[i];
            }
        }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const long long nth = n < (1 << 20) ? 1 : std::min<long long>(hw ? hw : 1, 16);
    if (nth <= 1) {
        run(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (long long k = 0; k < nth; ++k)
        th.emplace_back(run, n * k / nth, n * (k + 1) / nth);
    for (auto &t : th)
        t.join();
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
    if (last_time) {
        worker.drain();
        if (g.status != 0 || g.failed)
            return nullptr;
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
    worker.drain();
    if (g.ctx)
        pp2p_destroy(g.ctx);
    g.ctx = nullptr;
    g.call_cnt = 0;
    g.pack_acc.clear();
    g.soa.release();
}

void updateAcc_fromHostBuff(P2PPack pack, int active_level)
{
    const long long n = pack.npart;
    if ((long long)g.out.size() < 3 * n) {
        set_error(PP2P_ERR_STATE, "updateAcc_fromHostBuff: hostmalloc_out buffer too small");
        return;
    }
    add_results(pack.ipart, n, g.out.data(), g.out.data() + n, g.out.data() + 2 * n, active_level);
}

void devMemcpyDH_updateAcc(P2PPack *packs, int npack, int active_level)
{
    worker.drain();
    if (!g.ctx || !packs || (!g.pack_mode && g.failed))
        return;                                     // a failed step keeps its status and adds nothing
    g.status = 0;
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
    // the device rounds to FP32; the pinned host buffer is kept across steps
    g.soa.fit((size_t)3 * (NPART > 0 ? NPART : 1));
    if (!check(pp2p_get_acc_f32_soa(g.ctx, g.soa.get()), "get_acc"))
        return;
    const float *soa = g.soa.get();
    for (int n = 0; n < npack; ++n) {
        if (packs[n].ntask == 0)
            continue;
        const long long np = packs[n].npart, ip = packs[n].ipart;
        if (ip < 0 || ip + np > NPART) {
            set_error(PP2P_ERR_ARG, "pack particle range outside [0, NPART)");
            return;
        }
        add_results(ip, np, soa + ip, soa + (size_t)NPART + ip, soa + 2 * (size_t)NPART + ip, active_level);
    }
}

void sync_with_dev(void)
{
    worker.drain();
    if (g.ctx)
        check(pp2p_sync(g.ctx), "sync");
}

// seconds of device time of the step's force launches (one per flush batch)
double get_last_kntime(int)
{
    worker.drain();
    if (!g.ctx)
        return 0.0;
    check(pp2p_sync(g.ctx), "sync");
    return g.kernel_ms * 1e-3;
}

int p2p_hip_status(void)
{
    worker.drain();
    return g.status;
}

const char *p2p_hip_error(void)
{
    worker.drain();
    return g.error.c_str();
}

}  // extern "C"

#!/usr/bin/env python3
"""Benchmark of the short-range P2P path on one MI355X.

A step is one evaluation of every leaf pair of the workload's interaction
list (zero the accumulator + the persistent force launch over the prepared,
cost-sorted queue), with particles, leaves and the list resident in HBM.
The default workload is the 256^3 clustered configuration (half a
Zel'dovich-displaced grid, half in 2,000 Plummer spheres, L = 400 Mpc/h,
leaves <= 64, mesh 256^3, rs = 1.25 cells, cutoff u <= 2.8125, Plummer eps =
spacing / 40), lists from the kd-tree + dual walk with periodic ghost copies.

Metric: G pair-interactions/s = N_list / t_step, N_list = sum over leaf pairs
of n_t * n_s (the reference's own counter, src/gpu_entry.cu:50-61).
Roofline: algorithmic FLOPs = 48 N_in + 11 (N_list - N_in) (the per-pair
operation count of the reference pair function inside / beyond the cutoff)
over the force launch's device time, against the FP32 vector peak.
Accuracy: per-particle relative error against the FP64 oracle on a seeded
sample of target leaves.  cpu_baseline: the FP32 restatement of the
reference kernel (oracle) on a bounded sample of the same list, all host
threads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "photons-on-mt3000_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak
FLOPS_IN, FLOPS_OUT = 48, 11      # per-pair operation counts inside / beyond the cutoff
METRIC = "G pair-interactions/s + % FP32 vector peak at p99 force err <=1e-5, 1 MI355X"
# HBM bytes per force launch (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, factors from a
# 1 GiB copy kernel profiled in the same runs); see the named profile summary
MEASURED_TRAFFIC = {
    "clustered256": {"bytes": 6.706e10, "profile": "profiles/p2p_force_clustered256_traffic.md"},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str, n1: int | None, threads: int, softening: int = 1):
    from photons_p2p import workloads as W
    if name == "clustered":
        n = n1 or 256
        # 256^3 in 400 Mpc/h; 512^3 in 1 Gpc/h (the large single-GPU share)
        box = 1000.0 if n == 512 else 400.0 * n / 256
        return W.config_clustered(n, box=box, nthreads=threads, progress=lambda m: log(f"[bench] {m}"))
    if name == "plain":
        return W.config_plain(n1 or 64, nthreads=threads, softening=softening)
    if name == "precision":
        return W.config_precision(n1 or 128, nthreads=threads)
    if name == "uneven":
        return W.uneven_pairs()
    raise SystemExit(f"unknown workload {name}")


def sample_rows(w, nrows: int, seed: int):
    """Task mask selecting every leaf pair of a seeded sample of target leaves."""
    tg = np.unique(w.task_t)
    g = np.random.default_rng(seed)
    pick = g.choice(tg, size=min(nrows, tg.size), replace=False)
    return np.isin(w.task_t, pick)


def accuracy(w, acc, mask, threads, pos=None, act=None, level=1, softening=None):
    import pyoracle
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening if softening is None else softening,
                          box=w.box)
    csr = pyoracle.build_csr(w.task_t[mask], w.task_s[mask], w.nleaf)
    ref = pyoracle.p2p_f64(prm, w.pos if pos is None else pos, w.leaf_ip, w.leaf_np, csr, act=act, level=level,
                           nthreads=threads)
    tgt = csr[0]
    idx = np.concatenate([np.arange(w.leaf_ip[t], w.leaf_ip[t] + w.leaf_np[t]) for t in tgt])
    if act is not None:
        idx = idx[act[idx] == level]
    return pyoracle.rel_err(acc[idx], ref[idx]), idx.size


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, threads: int, budget_s: float):
    """FP32 restatement of the reference kernel on a bounded row sample:
    the sample grows until one pass takes about a third of the budget, then
    runs twice more; the best of the three passes counts."""
    import pyoracle
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening, box=w.box)
    pos32 = np.ascontiguousarray(w.pos.T.astype(np.float32))
    ntgt = np.unique(w.task_t).size
    nrows = 16
    while True:
        mask = sample_rows(w, nrows, seed=99)
        csr = pyoracle.build_csr(w.task_t[mask], w.task_s[mask], w.nleaf)
        npairs = int(np.sum(w.leaf_np[w.task_t[mask]].astype(np.int64) * w.leaf_np[w.task_s[mask]]))
        t0 = time.perf_counter()
        pyoracle.p2p_ref32(prm, pos32, w.leaf_ip, w.leaf_np, csr, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= budget_s / 6 or nrows >= ntgt:
            break
        nrows = min(ntgt, max(2 * nrows, int(nrows * budget_s / 3 / max(dt, 1e-3))))
    best = dt
    for _ in range(2):
        t0 = time.perf_counter()
        pyoracle.p2p_ref32(prm, pos32, w.leaf_ip, w.leaf_np, csr, nthreads=threads)
        best = min(best, time.perf_counter() - t0)
    return npairs / best, nrows, npairs, best


def copy_calibration(mib: int) -> dict:
    """Five device-to-device copies of `mib` MiB (hipMemcpy's copy kernel): a
    kernel whose HBM traffic is known exactly (mib MiB read + mib MiB written
    per launch), profiled beside the force kernel so that the FETCH_SIZE /
    WRITE_SIZE counters can be scaled to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    for rc in (hip.hipMalloc(ctypes.byref(a), n), hip.hipMalloc(ctypes.byref(b), n),
               hip.hipMemset(a, 1, n), hip.hipEventCreate(ctypes.byref(e0)), hip.hipEventCreate(ctypes.byref(e1))):
        if rc != 0:
            raise RuntimeError(f"copy calibration: HIP error {rc}")
    t = []
    for _ in range(5):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, n, 3, None)                # 3 = hipMemcpyDeviceToDevice
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        t.append(ms.value)
    hip.hipFree(a)
    hip.hipFree(b)
    ms = float(np.median(t))
    log(f"[bench] copy calibration: {mib} MiB in {ms:.3f} ms ({2 * (mib << 20) / ms / 1e6:.0f} GB/s)")
    return {"mib": mib, "ms": ms, "bytes_read": mib << 20, "bytes_written": mib << 20}


def dropin_protocol(w, threads: int, len_task: int) -> dict:
    """One force step through the drop-in library in the reference's protocol,
    from a host tree: the host walk hands LEN_TASK tasks at a time to
    task_compute_p2p_gpu (src/fmm.c:500-515) while it runs, then
    get_last_kntime, hostmalloc_out, devMemcpyDH_updateAcc, hostfree_out
    (src/gravity.c:441-447).  The host program's globals come from
    tests/support/libhostglobals.so (what src/global.c defines in the
    reference), the particles in Body records, the leaves in Pack records.
    Timed twice: with the walk and the device work overlapped (the drop-in's
    worker computes each flush as it arrives), and with the walk finished
    before the first flush (for comparison).  The positions stay in host
    memory and go to the device in the step, as in the reference."""
    import ctypes
    import photons_p2p as pp
    from photons_p2p import workloads as W
    m = w.meta
    host = ctypes.CDLL(os.path.join(ROOT, "tests", "support", "libhostglobals.so"), mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(pp.DROPIN_LIB)
    i32, vp = ctypes.c_int, ctypes.c_void_p
    lib.hostmalloc_task.argtypes = [i32]
    lib.hostmalloc_out.argtypes = [i32]
    lib.devMemcpyDH_updateAcc.argtypes = [vp, i32, i32]
    lib.get_last_kntime.argtypes = [i32]
    lib.get_last_kntime.restype = ctypes.c_double
    lib.p2p_hip_status.restype = i32
    lib.p2p_hip_error.restype = ctypes.c_char_p
    n, nl = int(w.npart), int(w.nleaf)
    body = np.zeros(n, dtype=[("tag", np.int32), ("act", np.int32), ("posi", np.int32, 3), ("vel", np.float32, 3),
                              ("id", np.uint64)])
    assert body.itemsize == 40
    body["tag"] = m["tag"]
    body["posi"] = m["posi"]
    leaves = np.zeros(nl, dtype=[("active", np.int32), ("npart", np.int32), ("ipart", np.int32),
                                 ("rest", np.uint8, 196)])
    assert leaves.itemsize == 208
    leaves["ipart"] = w.leaf_ip
    leaves["npart"] = w.leaf_np
    packs = np.zeros(9, dtype=[("raw", np.uint8, 40)])
    accs = [np.zeros(n, np.float32) for _ in range(3)]
    g = lambda name, t: t.in_dll(host, name)  # noqa: E731
    g("NPART", i32).value = n
    g("first_leaf", i32).value = n
    g("last_leaf", i32).value = n + nl
    g("NPACK", i32).value = 0
    g("MAXNPACK", i32).value = 9
    g("SoftenScale", ctypes.c_float).value = w.eps
    g("splitRadius", ctypes.c_float).value = w.rs
    g("MASSPART", ctypes.c_float).value = w.mass
    g("INT2POS", ctypes.c_double).value = m["int2pos"]
    vp.in_dll(host, "part").value = body.ctypes.data
    vp.in_dll(host, "leaf").value = leaves.ctypes.data - n * 208          # leaf[first_leaf + l]
    vp.in_dll(host, "packlist").value = packs.ctypes.data
    accp = (vp * 3).in_dll(host, "acc")
    for d in range(3):
        accp[d] = accs[d].ctypes.data
    hl = W._host_lib()
    P = lambda a: a.ctypes.data_as(vp)  # noqa: E731
    # the tree is built from the particles in their input order, as the
    # workload's was: the partition's outcome depends on that order, and the
    # leaves must be the Pack records' (the Body records are in tree order)
    perm = m["perm"]
    posi = np.empty_like(np.asarray(m["posi"], np.int32))
    posi[perm] = m["posi"]
    tag = np.empty_like(np.asarray(m["tag"], np.int32))
    tag[perm] = m["tag"]
    bdl = np.asarray(m["bdl"], np.float32)
    bdr = np.asarray(m["bdr"], np.float32)
    t0 = time.perf_counter()
    h = hl.phost_build_tree(n, P(posi), P(tag), m["int2pos"], m["maxleaf"], P(bdl), P(bdr), threads)
    tree_s = time.perf_counter() - t0
    ip = np.empty(nl, np.int32)
    npl = np.empty(nl, np.int32)
    hl.phost_get_leaves(h, P(ip), P(npl), None, None, None)
    hperm = np.empty(n, np.int32)
    hl.phost_get_perm(h, P(hperm))
    same_leaves = bool(np.array_equal(ip, w.leaf_ip) and np.array_equal(npl, w.leaf_np)
                       and np.array_equal(hperm, perm))
    if not same_leaves:
        hl.phost_free(h)
        raise RuntimeError("drop-in step: the host tree's leaves differ from the Pack records")
    fn = ctypes.cast(lib.task_compute_p2p_gpu, vp)
    res = {"len_task": len_task, "host_tree_s": tree_s, "leaves_identical": same_leaves, "host_threads": threads}
    try:
        for overlap in (1, 0, 1):
            for a in accs:
                a[:] = 0.0
            g("NPACK", i32).value = 0
            lib.hostmalloc_task(len_task)
            tt = vp.in_dll(host, "task_t").value
            ts = vp.in_dll(host, "task_s").value
            t1 = time.perf_counter()
            rc = hl.phost_walk_stream(h, np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]), threads,
                                      len_task, n, tt, ts, fn, overlap)
            t2 = time.perf_counter()
            kt = lib.get_last_kntime(0)
            t3 = time.perf_counter()
            lib.hostmalloc_out(n)
            lib.devMemcpyDH_updateAcc(packs.ctypes.data, g("NPACK", i32).value, -1)
            lib.hostfree_out()
            t4 = time.perf_counter()
            lib.hostfree_task()
            st = lib.p2p_hip_status()
            if rc != 0 or st != 0:
                raise RuntimeError(f"drop-in step failed: {st} {lib.p2p_hip_error().decode()}")
            key = "overlapped" if overlap else "walk_then_flushes"
            res[key] = {"walk_and_flush_calls_ms": 1e3 * (t2 - t1), "wait_for_device_ms": 1e3 * (t3 - t2),
                        "collect_ms": 1e3 * (t4 - t3), "step_ms": 1e3 * (t4 - t1 + tree_s),
                        "force_kernels_ms": 1e3 * kt}
    finally:
        hl.phost_free(h)
        lib.devfree()
    res["acc"] = np.stack(accs, axis=1).astype(np.float64)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clustered", choices=["clustered", "plain", "precision", "uneven"])
    ap.add_argument("--n1", type=int, default=None, help="particles per side (default: the config's)")
    ap.add_argument("--softening", default="plummer", choices=["plummer", "ref"],
                    help="plain workload: Plummer (eps = spacing / 40) or the reference's max(r, eps) rule "
                         "(eps = SoftenScale = 0.03 cell)")
    ap.add_argument("--exec", default="queue", choices=["queue", "grid", "per_target"])
    ap.add_argument("--law", default=None, choices=["table", "poly", "linear"],
                    help="force law (default: the tuned table's entry for the workload, else table)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-acc", action="store_true", help="skip the accuracy sample")
    ap.add_argument("--acc-rows", type=int, default=300)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--waves-per-cu", type=int, default=0, help="persistent waves per CU (0: occupancy max)")
    ap.add_argument("--split", type=int, default=0, help="queue item size in pair-interactions (0: auto)")
    ap.add_argument("--segs", type=int, default=0, help="table segments per octave (16/32, 0: 32)")
    ap.add_argument("--no-cull", action="store_true")
    ap.add_argument("--ipl-max", type=int, default=0, help="most i-particles per lane of an item (0: library default)")
    ap.add_argument("--graph", type=int, default=0,
                    help="capture this many steps in one hipGraph and time its replays (0: eager steps)")
    ap.add_argument("--autotune", action="store_true",
                    help="measure the tuning knobs on this workload first (else use photons_p2p/tuned.json)")
    ap.add_argument("--tuned-out", default=None, help="write the autotuned entry into this JSON table")
    ap.add_argument("--active", type=float, default=0.0,
                    help="active fraction (0: every particle); targets with an active particle only")
    ap.add_argument("--precision-forms", action="store_true",
                    help="precision workload: also evaluate FP32 absolute coordinates and report both errors")
    ap.add_argument("--compare-grid", action="store_true",
                    help="also time one launch with one wave per item in list order (no persistence)")
    ap.add_argument("--active-list", default=None,
                    help="comma-separated active fractions evaluated one after another on the same workload")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a buffer of this many MiB 5 times (HBM counter calibration)")
    ap.add_argument("--compare-plain", action="store_true",
                    help="also time the one-launch-per-target form the same way and report the ratio")
    ap.add_argument("--chunk-model", type=int, default=0, help="i-chunking: 0 default, -1 chunks of 64, >0 cost")
    ap.add_argument("--no-pm", action="store_true", help="skip the particle-mesh measurement")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in protocol step (host walk streaming flushes into libp2p_hip.so)")
    ap.add_argument("--no-active-sweep", action="store_true",
                    help="skip the default active-subset sweep of the clustered workloads")
    ap.add_argument("--no-device-walk", action="store_true",
                    help="skip building the tree and the list again on the device (tree workloads)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the P2P path is evaluated on one MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import photons_p2p as pp
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.time()
    w = make_workload(args.workload, args.n1, threads, softening=0 if args.softening == "ref" else 1)
    log(f"[bench] workload {w.name}: N={w.npart} leaves={w.nleaf} tasks={w.task_t.size} "
        f"built in {time.time() - t0:.1f}s")
    act = None
    if args.active > 0:
        from photons_p2p import workloads as W
        n_grid = (w.meta["n1"] ** 3) // 2 if w.name.startswith("clustered") else None
        act, tmask = W.active_subset(w, args.active, n_grid=n_grid)
        w.task_t, w.task_s = w.task_t[tmask], w.task_s[tmask]
        log(f"[bench] active fraction {args.active}: {int((act == 1).sum())} active particles, "
            f"{w.task_t.size} tasks to targets holding one")
    pos_eval = w.pos
    if args.workload == "precision":
        pos_eval = w.pos + w.meta["offset"]      # absolute coordinates ~1e3 box lengths from the origin
    exec_form = {"queue": pp.EXEC_QUEUE, "grid": pp.EXEC_GRID, "per_target": pp.EXEC_PER_TARGET}[args.exec]
    law_ids = {"table": pp.LAW_TABLE, "poly": pp.LAW_POLY, "linear": pp.LAW_LINEAR}
    law = law_ids[args.law or "table"]

    ctx = pp.Context(0)
    ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, law=law, box=w.box))
    ctx.set_tuning(waves_per_cu=args.waves_per_cu, split_pairs=args.split, segs_per_octave=args.segs,
                   no_cull=int(args.no_cull), chunk_model=args.chunk_model, ipl_max=args.ipl_max)
    ctx.set_particles(pos_eval, w.leaf_ip, w.leaf_np)
    if act is not None:
        ctx.set_active(act, 1)
    ctx.set_tasks(w.task_t, w.task_s)
    prep_ms = ctx.prepare()
    st = ctx.list_stats()
    from photons_p2p import autotune as AT
    sig = AT.signature(st["ntask"], w.npart, w.nleaf, st["npairs"])
    lawp = pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, law=law, box=w.box)
    tuned = None
    explicit = args.waves_per_cu or args.split or args.chunk_model or args.segs or args.no_cull or args.ipl_max
    if args.autotune:
        res = AT.tune(ctx, lawp, log=log)
        law = res["best"].get("law", law)
        tuned = {"source": "autotune", "signature": sig, **res}
        log(f"[bench] autotuned {sig}: {res['best']} ({res['ms']:.3f} ms/step)")
        if args.tuned_out:
            table = AT.load_table(args.tuned_out)
            table[sig] = {"best": res["best"], "ms": res["ms"], "workload": w.name}
            AT.save_table(table, args.tuned_out)
    elif not explicit:
        cfg = AT.lookup(sig)
        if cfg is not None:
            if args.law is not None:
                cfg = {**cfg, "law": law}          # an explicit --law wins over the table's
            AT.apply(ctx, lawp, cfg)
            law = cfg.get("law", law)
            tuned = {"source": "photons_p2p/tuned.json", "signature": sig, "best": cfg}
            log(f"[bench] tuned.json {sig}: {cfg}")
    if tuned is not None:
        prep_ms = ctx.prepare()                    # the items of the tuned split
        st = ctx.list_stats()
    n_list = st["npairs"]
    n_in = ctx.count_inside()
    log(f"[bench] prepared in {prep_ms:.1f} ms: rows={st['nrows']} items={st['nitems']} "
        f"N_list={n_list:.4e} N_in={n_in:.4e} ({n_in / max(n_list, 1):.3f})")

    def timed(form):
        """(ms per step, force ms per step, wall seconds) of `form`."""
        if args.graph > 0:
            reps = max(1, args.steps // args.graph)
            ctx.run_graph(form, nsteps=args.graph, nreplay=1)          # warm: capture + one replay
            w0 = time.perf_counter()
            ctx.run_graph(form, nsteps=args.graph, nreplay=reps)
            wl = time.perf_counter() - w0
            ms = ctx.last_kernel_ms() / (reps * args.graph)
            return ms, ms, wl
        if args.warmup > 0:
            ctx.time_steps(form, args.warmup)
        ctx.sync()
        w0 = time.perf_counter()
        tot, frc = ctx.time_steps(form, args.steps)
        ctx.sync()
        return tot / args.steps, frc / args.steps, time.perf_counter() - w0

    ms_step, force_step, wall = timed(exec_form)
    t_kernel = force_step * 1e-3
    plain = None
    if args.compare_plain:
        p_ms, p_force, p_wall = timed(pp.EXEC_PER_TARGET)
        plain = {"ms_per_step": p_ms, "value": n_list / (p_ms * 1e-3) / 1e9, "speedup_of_queue": p_ms / ms_step,
                 "form": "one launch per target leaf", "graph_steps": args.graph}
        log(f"[bench] per-target form: {p_ms:.3f} ms/step; queue form is {p_ms / ms_step:.2f}x faster")
    grid_form = None
    if args.compare_grid:
        g_ms, _, _ = timed(pp.EXEC_GRID)
        grid_form = {"ms_per_step": g_ms, "value": n_list / (g_ms * 1e-3) / 1e9, "speedup_of_queue": g_ms / ms_step,
                     "form": "one launch, one wave per item, list order", "graph_steps": args.graph}
        log(f"[bench] grid form: {g_ms:.3f} ms/step; queue form is {g_ms / ms_step:.2f}x faster")
    value = n_list / (ms_step * 1e-3) / 1e9
    flops = FLOPS_IN * n_in + FLOPS_OUT * (n_list - n_in)
    achieved = flops / t_kernel / 1e12
    log(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (force {t_kernel * 1e3:.3f} ms), "
        f"wall {wall:.3f}s, {value:.1f} Gpairs/s, {achieved:.1f} TFLOP/s "
        f"({100 * achieved / PEAK_FP32_TFLOPS:.1f}% of FP32 peak)")

    ctx_tuning = ctx.get_tuning()
    rel = None
    forms = None
    if not args.no_acc:
        acc = ctx.acc()
        mask = sample_rows(w, args.acc_rows, seed=20261202)
        (med, p99, mx), nchk = accuracy(w, acc, mask, threads, pos=pos_eval, act=act)
        rel = {"median": med, "p99": p99, "max": mx, "particles": int(nchk),
               "sample": f"{args.acc_rows} target leaves, FP64 oracle"}
        log(f"[bench] rel err vs FP64 oracle ({nchk} particles): median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
        if args.precision_forms:
            ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, law=law,
                                    frame=pp.FRAME_ABS32, box=w.box))
            ctx.set_particles(pos_eval, w.leaf_ip, w.leaf_np)
            if act is not None:
                ctx.set_active(act, 1)
            ctx.set_tasks(w.task_t, w.task_s)
            ctx.compute(exec_form)
            (m2, p2, x2), _ = accuracy(w, ctx.acc(), mask, threads, pos=pos_eval, act=act)
            forms = {"fp32_leaf_relative_fp64_offsets": rel,
                     "fp32_absolute": {"median": m2, "p99": p2, "max": x2},
                     "oracle": "FP64 direct sum, same lists and force law"}
            log(f"[bench] FP32 absolute coordinates: median {m2:.2e} p99 {p2:.2e} max {x2:.2e}")
    sweep = None
    if args.active_list is None and w.name.startswith("clustered") and args.active <= 0 and not args.no_active_sweep:
        args.active_list = "1,0.1,0.01"          # the active fractions of the 512^3 configuration
    if args.active_list:
        from photons_p2p import workloads as W
        n_grid = (w.meta["n1"] ** 3) // 2 if w.name.startswith("clustered") else None
        sweep = []
        for f in [float(x) for x in args.active_list.split(",")]:
            a_f, tm = W.active_subset(w, f, n_grid=n_grid)
            ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, law=law, box=w.box))
            ctx.set_particles(pos_eval, w.leaf_ip, w.leaf_np)
            ctx.set_active(a_f, 1)
            ctx.set_tasks(w.task_t[tm], w.task_s[tm])
            p_ms = ctx.prepare()
            sf = ctx.list_stats()
            f_ms, f_force, _ = timed(exec_form)
            nact = int((a_f == 1).sum())
            ent = {"fraction": f, "active_particles": nact, "ntask": int(tm.sum()), "n_list": int(sf["npairs"]),
                   "ms_per_step": f_ms, "force_ms": f_force, "prepare_ms": p_ms,
                   "active_per_s": nact / (f_ms * 1e-3), "ms_vs_all_active": f_ms / ms_step}
            if not args.no_acc:
                sub = w.task_t[tm]
                tg = np.unique(sub)
                pick = np.random.default_rng(7).choice(tg, size=min(100, tg.size), replace=False)
                mk = tm & np.isin(w.task_t, pick)
                (m3, p3, x3), nck = accuracy(w, ctx.acc(), mk, threads, pos=pos_eval, act=a_f)
                ent["rel_err"] = {"median": m3, "p99": p3, "max": x3, "particles": int(nck)}
            sweep.append(ent)
            log(f"[bench] active {f}: {nact} active, {ent['ntask']} tasks, N_list {sf['npairs']:.3e}, "
                f"{f_ms:.3f} ms/step ({f_ms / ms_step:.3f} of the all-active step), prepare {p_ms:.1f} ms"
                + (f", p99 {ent['rel_err']['p99']:.2e}" if "rel_err" in ent else ""))
    lists = None
    if "tree" in w.meta and args.active <= 0 and not args.no_device_walk:
        # the list built again on the device, checked against the host's:
        # (1) the walk of the host tree (uploaded), (2) tree and walk both on the
        # device from the input-order grid positions; best of 3 calls each
        m = w.meta
        ls = m["list_seconds"]

        def best_of(fn, k=3):
            dts = []
            for _ in range(k):
                t1 = time.perf_counter()
                r = fn()
                dts.append(time.perf_counter() - t1)
            return r, 1e3 * min(dts)

        nwalk, walk_ms = best_of(lambda: ctx.walk_tasks(m["tree"], m["active_leaf"], np.float32(m["open_angle"]),
                                                        np.float32(m["cutoff_radius"])))
        tt, ts = ctx.get_tasks()
        same_walk = bool(np.array_equal(tt, w.task_t) and np.array_equal(ts, w.task_s))
        perm = m["perm"]
        posi_in = np.empty_like(m["posi"])
        tag_in = np.empty_like(m["tag"])
        posi_in[perm] = m["posi"]
        tag_in[perm] = m["tag"]
        # the upload alone for scale: the same grid positions through set_particles_i32
        _, upload_ms = best_of(lambda: ctx.set_particles_i32(m["posi"], m["int2pos"], w.leaf_ip, w.leaf_np))
        _, tree_ms = best_of(lambda: ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"]))
        t = ctx.get_tree()
        same_tree = bool(np.array_equal(t["perm"], perm) and np.array_equal(t["leaf_ip"], w.leaf_ip)
                         and np.array_equal(t["leaf_np"], w.leaf_np) and np.array_equal(t["active"], m["active_leaf"])
                         and all(np.array_equal(t[k], m["tree"][k]) for k in m["tree"]))
        del t
        _, walk2_ms = best_of(lambda: ctx.walk_tree(np.float32(m["open_angle"]), np.float32(m["cutoff_radius"])))
        tt, ts = ctx.get_tasks()
        same_walk2 = bool(np.array_equal(tt, w.task_t) and np.array_equal(ts, w.task_s))
        del tt, ts

        # one whole P2P step from grid positions in input order, steady state:
        # (a) all on the device: tree, walk, CSR + work queue, force;
        # (b) host tree and walk (timed above), then positions and list uploaded,
        #     CSR + work queue, force
        def device_step():
            ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"])
            ctx.walk_tree(np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]))
            ctx.prepare()
            ctx.zero_acc()
            ctx.compute(exec_form)
            ctx.sync()

        def upload_step():
            ctx.set_particles_i32(m["posi"], m["int2pos"], w.leaf_ip, w.leaf_np)
            ctx.set_tasks(w.task_t, w.task_s)
            ctx.prepare()
            ctx.zero_acc()
            ctx.compute(exec_form)
            ctx.sync()

        _, dev_step_ms = best_of(device_step)
        _, up_step_ms = best_of(upload_step)
        lists = {"host_tree_s": ls["tree"], "host_walk_s": ls["walk"], "host_threads": threads,
                 "device_walk_ms": walk_ms, "device_walk_identical": same_walk,
                 "device_tree_ms": tree_ms, "device_tree_identical": same_tree, "set_particles_i32_ms": upload_ms,
                 "device_tree_walk_ms": walk2_ms, "device_tree_walk_identical": same_walk2, "ntask": int(nwalk),
                 "step_all_device_ms": dev_step_ms,
                 "step_host_lists_ms": 1e3 * (ls["tree"] + ls["walk"]) + up_step_ms,
                 "step_host_lists_upload_and_device_ms": up_step_ms}
        log(f"[bench] lists: host tree {ls['tree']:.2f}s + walk {ls['walk']:.2f}s on {threads} threads; "
            f"device walk of the host tree {walk_ms:.1f} ms (identical: {same_walk}); device tree {tree_ms:.1f} ms "
            f"(identical: {same_tree}; the position upload alone {upload_ms:.1f} ms) + walk {walk2_ms:.1f} ms "
            f"(identical: {same_walk2})")
        log(f"[bench] one step from grid positions: all on the device {dev_step_ms:.1f} ms; host tree + walk, "
            f"then upload + prepare + force {1e3 * (ls['tree'] + ls['walk']) + up_step_ms:.1f} ms "
            f"({up_step_ms:.1f} ms after the host lists)")
    m2l = None
    if "m2l" in w.meta and args.active <= 0 and not args.no_device_walk:
        # the short-range M2L task loop on the device over the walk's
        # well-separated pairs (cells: leaves, then nodes); synthetic
        # multipoles of realistic shape (the cost does not depend on them)
        m = w.meta
        mt, mst = m["m2l"]
        tree = m["tree"]
        center = np.concatenate([tree["leaf_center"], tree["node_center"]]).astype(np.float32)
        ncell = center.shape[0]
        g = np.random.default_rng(20261211)
        scale = np.array([1.0] + [0.3] * 3 + [0.1] * 6 + [0.03] * 10, np.float32)
        Mm = (g.normal(size=(ncell, 20)).astype(np.float32) * scale)
        Mm[:, 0] = np.abs(Mm[:, 0]) + 0.5
        kms, walls = [], []
        for _ in range(3):
            t1 = time.perf_counter()
            Lm, k_ms = ctx.m2l(center, Mm, mt, mst, w.rs)
            walls.append(time.perf_counter() - t1)
            kms.append(k_ms)
        # accuracy on a sample of target cells against the FP64 oracle
        import pyoracle
        rows = np.unique(mt)
        pick = g.choice(rows, size=min(300, rows.size), replace=False)
        sel = np.isin(mt, pick)
        ref = pyoracle.m2l_tasks_f64(center.astype(np.float64), Mm.astype(np.float64), mt[sel], mst[sel], w.rs)
        worst = 0.0
        for lo, hi in ((0, 1), (1, 4), (4, 10), (10, 20)):
            sc = np.abs(ref[pick, lo:hi]).max(axis=1)
            ok = sc > 0
            worst = max(worst, float((np.abs(Lm[pick][ok, lo:hi] - ref[pick][ok, lo:hi]).max(axis=1)
                                      / sc[ok]).max(initial=0.0)))
        # CPU: the reference's task loop in its FP32 arithmetic (the oracle's
        # restatement, one thread, the whole list), best of 2
        cts = []
        for _ in range(2):
            t1 = time.perf_counter()
            pyoracle.m2l_tasks_ref32(center, Mm, mt, mst, np.float32(w.rs))
            cts.append(time.perf_counter() - t1)
        cpu_rate = mt.size / min(cts)
        # the whole multipole pass on the device tree: P2M, M2M, M2L, L2L, L2P
        perm = m["perm"]
        posi_in = np.empty_like(m["posi"])
        tag_in = np.empty_like(m["tag"])
        posi_in[perm] = m["posi"]
        tag_in[perm] = m["tag"]
        ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"])
        fmm_ms = []
        for _ in range(3):
            ctx.zero_acc()
            t1 = time.perf_counter()
            dev_ms = ctx.fmm_short(mt, mst)
            fmm_ms.append((dev_ms, 1e3 * (time.perf_counter() - t1)))
        fmm_dev, fmm_wall = min(fmm_ms)

        # the whole short-range step from grid positions in input order, all on
        # the device: tree, walk (near pairs and well-separated pairs), P2P
        # force, multipole pass over the walk's own pairs; best of 3
        def short_step():
            ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"])
            ctx.walk_tree(np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]))
            ctx.prepare()
            ctx.zero_acc()
            ctx.compute(exec_form)
            ctx.fmm_short()
            ctx.sync()

        sts = []
        for _ in range(3):
            t1 = time.perf_counter()
            short_step()
            sts.append(1e3 * (time.perf_counter() - t1))
        dmt, dms = ctx.get_m2l_tasks()
        same_pairs = bool(dmt.size == mt.size and np.array_equal(np.sort(dmt.astype(np.int64) << 32 | dms),
                                                                 np.sort(mt.astype(np.int64) << 32 | mst)))
        del dmt, dms
        m2l = {"ntask": int(mt.size), "ncell": int(ncell), "nrows": int(rows.size), "kernel_ms": min(kms),
               "fmm_short_device_ms": fmm_dev, "fmm_short_call_ms": fmm_wall,
               "device_walk_pairs_identical": same_pairs, "short_step_all_device_ms": min(sts),
               "call_ms": 1e3 * min(walls), "tasks_per_s": mt.size / (min(kms) * 1e-3),
               "worst_rel_err": worst, "cpu_ref32_tasks_per_s": cpu_rate,
               "cpu_sample": f"all {mt.size} tasks, the reference's m2l task loop in FP32 (C restatement), 1 thread"}
        log(f"[bench] m2l: {mt.size} well-separated pairs on {rows.size} target cells: kernel {min(kms):.3f} ms "
            f"({mt.size / (min(kms) * 1e-3) / 1e9:.2f} G tasks/s), call {1e3 * min(walls):.1f} ms, "
            f"worst per-order error {worst:.1e}; CPU (reference arithmetic, 1 thread) {cpu_rate / 1e6:.2f} M tasks/s")
        log(f"[bench] multipole pass on the device tree (P2M, M2M, M2L, L2L, L2P): {fmm_dev:.2f} ms device, "
            f"{fmm_wall:.1f} ms call")
        log(f"[bench] short-range step all on the device (tree, walk, P2P, multipoles): {min(sts):.1f} ms; "
            f"device walk's well-separated pairs = host's: {same_pairs}")
    pm = None
    if "box_i" in w.meta and args.active <= 0 and not args.no_pm:
        # the long-range particle mesh of the same particles (local ones,
        # tag 1): mesh of the split (rs = 1.25 cells), whole box on the device
        m = w.meta
        box = float(m["box_i"] * m["int2pos"])
        nside = int(round(1.25 * box / w.rs))
        nside += nside & 1
        calls = []
        for _ in range(3):
            t1 = time.perf_counter()
            acc_pm, dev_ms = ctx.pm_forces(m["posi"], m["tag"], m["int2pos"], box, nside, w.rs, w.mass)
            calls.append((dev_ms, 1e3 * (time.perf_counter() - t1)))
        pm_dev, pm_call = min(calls)
        import oracle_pm
        loc = m["tag"] == 1
        rho = ctx.pm_mesh(0)
        mass_err = abs(rho.sum() / (loc.sum() * float(np.float32(w.mass))) - 1.0)
        a64 = acc_pm[loc].astype(np.float64)
        momentum = float(np.abs(a64.sum(0)).max() / np.abs(a64).sum())
        # the oracle's transform of the device mass mesh, and its gradient and
        # interpolation at 20,000 sampled particles of the device potential
        pot = ctx.pm_mesh(1)
        t1 = time.perf_counter()
        pot_ref = oracle_pm.potential(rho, np.float32(w.rs), np.float32(box))
        cpu_pot_s = time.perf_counter() - t1
        pot_err = float(np.abs(pot - pot_ref).max() / np.abs(pot_ref).max())
        del pot_ref, rho
        pick = np.random.default_rng(5).choice(np.nonzero(loc)[0], size=min(20000, int(loc.sum())), replace=False)
        t1 = time.perf_counter()
        a_ref = oracle_pm.accelerations(pot, m["posi"][pick], m["int2pos"], np.float32(box))
        cpu_interp_rate = pick.size / (time.perf_counter() - t1)
        acc_err = float(np.abs(acc_pm[pick] - a_ref).max() / np.abs(a_ref).max())
        del pot
        # the whole force of the split from grid positions in input order, all on
        # the device: tree, walk, P2P, multipole pass, particle mesh, into one
        # accumulator; best of 3
        grav_ms = None
        if "perm" in m and not args.no_device_walk:
            perm = m["perm"]
            posi_in = np.empty_like(m["posi"])
            tag_in = np.empty_like(m["tag"])
            posi_in[perm] = m["posi"]
            tag_in[perm] = m["tag"]

            def gravity_step():
                ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"])
                ctx.walk_tree(np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]))
                ctx.prepare()
                ctx.zero_acc()
                ctx.compute(exec_form)
                ctx.fmm_short()
                ctx.pm_add_tree(box, nside, w.rs, w.mass)
                ctx.sync()

            gts = []
            for _ in range(3):
                t1 = time.perf_counter()
                gravity_step()
                gts.append(1e3 * (time.perf_counter() - t1))
            grav_ms = min(gts)
            log(f"[bench] whole force of the split all on the device (tree, walk, P2P, multipoles, particle mesh): "
                f"{grav_ms:.1f} ms")
        pm = {"nside": nside, "box": box, "nlocal": int(loc.sum()), "step_all_forces_device_ms": grav_ms, "device_ms": pm_dev, "call_ms": pm_call,
              "mass_rel_err": mass_err, "momentum_rel": momentum, "potential_rel_err_vs_oracle_fft": pot_err,
              "acc_rel_err_vs_oracle_sample": acc_err,
              "cpu_oracle_potential_s": cpu_pot_s, "cpu_oracle_interp_particles_per_s": cpu_interp_rate,
              "cpu_sample": "numpy oracle, one process: FFT + Green's function + inverse FFT of the whole mesh; "
                            "gradient + interpolation at the 20,000 checked particles"}
        log(f"[bench] pm: {int(loc.sum())} particles on a {nside}^3 mesh: {pm_dev:.2f} ms device "
            f"(deposit, FFT, Green, inverse FFT, gradient + interpolation), {pm_call:.1f} ms per call with the "
            f"upload; mass {mass_err:.1e}, |sum a| / sum |a| {momentum:.1e}, potential vs oracle FFT {pot_err:.1e}, "
            f"acc vs oracle (20000 particles) {acc_err:.1e}; CPU (numpy oracle): potential {cpu_pot_s:.2f} s, "
            f"interpolation {cpu_interp_rate / 1e6:.2f} M particles/s")
    ctx.close()

    copy_cal = None
    dropin = None
    if "tree" in w.meta and args.active <= 0 and not args.no_dropin:
        # LEN_TASK of src/setparam.c:41 for this mesh on one rank
        len_task = int(80000000 * (w.meta["mesh"] / 512.0) ** 3)
        d = dropin_protocol(w, threads, max(len_task, 1 << 16))
        acc_d = d.pop("acc")
        if not args.no_acc:
            mask = sample_rows(w, args.acc_rows, seed=20261202)
            # the drop-in applies the reference's max(r, SoftenScale) softening
            # (src/gpu_kernel.impl.h:66-81) whatever the workload's
            (md, pd, xd), nd = accuracy(w, acc_d, mask, threads, pos=pos_eval, softening=0)
            d["rel_err"] = {"median": md, "p99": pd, "max": xd, "particles": int(nd)}
        del acc_d
        d["step_host_lists_ms_before"] = lists["step_host_lists_ms"] if lists else None
        dropin = d
        ov, sq = d["overlapped"], d["walk_then_flushes"]
        log(f"[bench] drop-in protocol step ({w.name}, LEN_TASK {d['len_task']}): host tree {d['host_tree_s']:.2f}s, "
            f"then walk streaming flushes into task_compute_p2p_gpu + collect: overlapped {ov['step_ms']:.1f} ms "
            f"(walk {ov['walk_and_flush_calls_ms']:.1f} ms, then waited {ov['wait_for_device_ms']:.1f} ms), "
            f"walk first {sq['step_ms']:.1f} ms; force kernels {ov['force_kernels_ms']:.1f} ms"
            + (f"; p99 {d['rel_err']['p99']:.2e}" if "rel_err" in d else ""))
    if args.copy_calibration:
        copy_cal = copy_calibration(args.copy_calibration)

    cpu = None
    if not args.no_cpu:
        rate, nrows, npairs, dt = cpu_baseline(w, threads, args.cpu_budget)
        cpu = {"value": rate / 1e9, "unit": "Gpairs/s", "cores": threads, "kind": "port",
               "sample": f"{nrows} target leaves ({npairs:.3e} pair-interactions) of the same list, "
                         f"FP32 restatement of P2P_kernel/P2P_core, best of 3 passes ({dt:.2f}s)",
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        log(f"[bench] cpu baseline: {rate / 1e9:.3f} Gpairs/s on {threads} threads ({dt:.1f}s)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "tuning": ctx_tuning,
        "autotune": tuned,
        "config": {"workload": w.name, "npart": int(w.npart), "nleaf": int(w.nleaf), "ntask": int(w.task_t.size),
                   "n_list": int(n_list), "n_in": int(n_in), "exec": args.exec,
                   "law": {v: k for k, v in law_ids.items()}[law],
                   "softening": "plummer" if w.softening == 1 else "ref", "eps": w.eps, "rs": w.rs},
        "pct_fp32_peak": 100 * achieved / PEAK_FP32_TFLOPS,
        "rel_err": rel,
        "prepare_ms": prep_ms,
        "graph_steps": args.graph,
        "active_fraction": args.active if args.active > 0 else None,
        "precision_forms": forms,
        "plain_form": plain,
        "grid_form": grid_form,
        "active_sweep": sweep,
        "copy_calibration": copy_cal,
        "lists": lists,
        "dropin_step": dropin,
        "m2l": m2l,
        "pm": pm,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_TRAFFIC.get(w.name, {}).get("bytes") if args.active <= 0 else None,
                     "traffic_source": MEASURED_TRAFFIC.get(w.name, {}).get("profile") if args.active <= 0 else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

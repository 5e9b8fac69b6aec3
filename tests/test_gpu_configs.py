"""GPU parity on the tree-built configurations (kd-tree + dual walk lists with
periodic ghost copies), scaled down so the FP64 oracle finishes in seconds,
plus size-independent checks at larger sizes."""
import numpy as np
import pytest

import photons_p2p as pp
import pyoracle
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu


def _ctx_run(ctx, w, pos=None, frame=pp.FRAME_LEAF, law=pp.LAW_TABLE, exec_form=pp.EXEC_QUEUE):
    print(f"case: {w.name} N={w.npart} tasks={w.task_t.size} frame={frame} law={law}", flush=True)
    ctx.set_tuning()
    ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, law=law, frame=frame,
                            box=w.box))
    ctx.set_particles(w.pos if pos is None else pos, w.leaf_ip, w.leaf_np)
    ctx.set_tasks(w.task_t, w.task_s)
    ctx.compute(exec_form)
    return ctx.acc()


def _f64(w, pos=None, mask=None):
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening, box=w.box)
    tt, ts = (w.task_t, w.task_s) if mask is None else (w.task_t[mask], w.task_s[mask])
    csr = pyoracle.build_csr(tt, ts, w.nleaf)
    return pyoracle.p2p_f64(prm, w.pos if pos is None else pos, w.leaf_ip, w.leaf_np, csr), csr


@pytest.mark.parametrize("softening", [0, 1])
def test_plain_config_whole_list(gpu_ctx, softening):
    w = W.config_plain(24, softening=softening)
    ref, _ = _f64(w)
    a = _ctx_run(gpu_ctx, w)
    med, p99, mx = pyoracle.rel_err(a, ref)
    print(f"plain24 soft={softening}: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4


@pytest.mark.parametrize("law", [pp.LAW_TABLE, pp.LAW_LINEAR])
def test_clustered_config_whole_list(gpu_ctx, law):
    w = W.config_clustered(32, box=50.0, n_spheres=40)
    ref, _ = _f64(w)
    a = _ctx_run(gpu_ctx, w, law=law)
    med, p99, mx = pyoracle.rel_err(a, ref)
    print(f"clustered32 law {law}: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4


def test_count_inside_matches_oracle(gpu_ctx):
    w = W.config_plain(20)
    prm = pyoracle.params(w.eps, w.rs, softening=w.softening)
    csr = pyoracle.build_csr(w.task_t, w.task_s, w.nleaf)
    n_list, n_in = pyoracle.count_pairs(prm, w.pos, w.leaf_ip, w.leaf_np, csr)
    _ctx_run(gpu_ctx, w)
    st = gpu_ctx.list_stats()
    assert st["npairs"] == n_list == w.n_list()
    got = gpu_ctx.count_inside()
    print(f"N_list {n_list} N_in oracle {n_in} gpu {got}")
    assert abs(got - n_in) <= max(2, 1e-6 * n_in)


def test_precision_forms(gpu_ctx):
    """Precision stress (NFW halos, box offset by 1000 box lengths): the
    leaf-relative form stays at the FP32 floor; FP32 absolute coordinates do not."""
    w = W.config_precision(24)
    off = w.meta["offset"]
    pos_off = w.pos + off
    ref, _ = _f64(w, pos=pos_off)
    a_leaf = _ctx_run(gpu_ctx, w, pos=pos_off)
    med, p99, mx = pyoracle.rel_err(a_leaf, ref)
    print(f"leaf frame: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4
    a_abs = _ctx_run(gpu_ctx, w, pos=pos_off, frame=pp.FRAME_ABS32)
    med_a, p99_a, mx_a = pyoracle.rel_err(a_abs, ref)
    print(f"abs32 frame: median {med_a:.2e} p99 {p99_a:.2e} max {mx_a:.2e}")
    assert p99_a > 100 * p99


def test_exec_forms_agree_on_tree_lists(gpu_ctx):
    w = W.config_plain(20)
    outs = [_ctx_run(gpu_ctx, w, exec_form=e) for e in (pp.EXEC_QUEUE, pp.EXEC_GRID, pp.EXEC_PER_TARGET)]
    # GRID and PER_TARGET run the same items; split targets meet in FP64
    # atomics whose order varies: equal to FP64 rounding
    np.testing.assert_allclose(outs[1], outs[2], rtol=1e-11, atol=1e-13 * np.abs(outs[1]).max())
    # the queue splits long source lists: FP32 partial sums regroup
    med, p99, mx = pyoracle.rel_err(outs[0], outs[1])
    print(f"queue vs grid: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 2e-6 and mx <= 5e-5


def test_full_size_sample_and_momentum(gpu_ctx):
    """64^3 plain config at full size: sampled rows against the FP64 oracle.
    Then the same particles with no periodic ghost copies, so that every pair
    of the symmetric list (src/fmm.c:666-684) appears both ways and the total
    force sum_i a_i vanishes in exact arithmetic: the device's residual
    |sum a| / sum |a| must stay at the FP32 level (<= 1e-7), and the FP64
    oracle's on the same list (the whole list, every particle) at FP64
    rounding (<= 1e-12).  The whole open-box list is also compared particle
    by particle with the oracle (p99 <= 1e-5)."""
    w = W.config_plain(64)
    a = _ctx_run(gpu_ctx, w)
    assert np.all(np.isfinite(a))
    g = np.random.default_rng(3)
    tg = np.unique(w.task_t)
    pick = g.choice(tg, size=60, replace=False)
    mask = np.isin(w.task_t, pick)
    ref, csr = _f64(w, mask=mask)
    idx = np.concatenate([np.arange(w.leaf_ip[t], w.leaf_ip[t] + w.leaf_np[t]) for t in csr[0]])
    med, p99, mx = pyoracle.rel_err(a[idx], ref[idx])
    print(f"plain64 sample: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4

    pos = W.zeldovich_grid(64, 100.0, 0.3, W.rng_for(20261202))
    wn = W.tree_workload("plain64-open", pos, 100.0, 64, 32, 0.5, 100.0 / 64 / 40, 1, periodic_ghosts=False)
    an = _ctx_run(gpu_ctx, wn)
    refn, _ = _f64(wn)
    res_gpu = np.linalg.norm(an.sum(axis=0)) / np.linalg.norm(an, axis=1).sum()
    res_f64 = np.linalg.norm(refn.sum(axis=0)) / np.linalg.norm(refn, axis=1).sum()
    med, p99, mx = pyoracle.rel_err(an, refn)
    print(f"plain64 open box, whole list: median {med:.2e} p99 {p99:.2e} max {mx:.2e}; "
          f"|sum a| / sum |a|: device {res_gpu:.2e}, oracle {res_f64:.2e}")
    # every one of the 262,144 particles: the maximum is set by the particle
    # whose pair forces cancel the most (|a| far below the sum of the pair
    # force magnitudes, so FP32 rounding of the pairs is amplified), hence a
    # looser bound on the max than on the sampled rows above
    assert p99 <= 1e-5 and mx <= 1e-3
    assert res_f64 <= 1e-12
    assert res_gpu <= 1e-7


def _ref32_sample(w, pos, nrows, seed):
    """The reference kernel restated in FP32 (oracle_p2p_ref32: P2P_kernel /
    P2P_core(_1) with the 512-node table and 4th-order Taylor, FP32 absolute
    coordinates relative to the domain centre, src/gpu_entry.cu:90-96) and the
    FP64 oracle, on every leaf pair of a seeded sample of targets."""
    tg = np.unique(w.task_t)
    pick = np.random.default_rng(seed).choice(tg, size=min(nrows, tg.size), replace=False)
    mask = np.isin(w.task_t, pick)
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening)
    csr = pyoracle.build_csr(w.task_t[mask], w.task_s[mask], w.nleaf)
    c = 0.5 * (pos.min(axis=0) + pos.max(axis=0))
    r32 = pyoracle.p2p_ref32(prm, np.ascontiguousarray((pos - c).T.astype(np.float32)), w.leaf_ip, w.leaf_np,
                             csr).T.astype(np.float64)
    f64 = pyoracle.p2p_f64(prm, pos, w.leaf_ip, w.leaf_np, csr)
    idx = np.concatenate([np.arange(w.leaf_ip[t], w.leaf_ip[t] + w.leaf_np[t]) for t in csr[0]])
    return idx, r32, f64


@pytest.fixture(scope="module")
def clustered256():
    """Config 2's workload (built once per module: about 10 s of host work)."""
    return W.config_clustered(256)


@pytest.mark.parametrize("cfg", ["plain64", "clustered256"])
def test_against_the_reference_kernel_restatement(gpu_ctx, cfg, request):
    """Configs 1 and 2 against the FP32 restatement of the reference's own
    kernel on identical inputs and lists.  The reference evaluates pairs on
    FP32 absolute coordinates, whose rounding (about 4e-6 at |x| = 50) limits
    its own accuracy: against the FP64 oracle it reaches p99 3e-5 to 8e-5 and
    max 1e-4 to 4e-4 on these configurations.  The stated tolerance is that
    level -- p99 <= 1e-4 of the device against the restatement -- and the
    device must be closer to the FP64 oracle than the restatement is."""
    w = W.config_plain(64) if cfg == "plain64" else request.getfixturevalue("clustered256")
    a = _ctx_run(gpu_ctx, w)
    idx, r32, f64 = _ref32_sample(w, w.pos, 150, seed=11)
    d_ref = pyoracle.rel_err(a[idx], r32[idx])
    d_gpu = pyoracle.rel_err(a[idx], f64[idx])
    d_r32 = pyoracle.rel_err(r32[idx], f64[idx])
    print(f"{cfg} ({idx.size} particles): device vs ref32 median {d_ref[0]:.2e} p99 {d_ref[1]:.2e} "
          f"max {d_ref[2]:.2e}; device vs f64 p99 {d_gpu[1]:.2e}; ref32 vs f64 p99 {d_r32[1]:.2e}")
    assert d_ref[1] <= 1e-4 and d_ref[2] <= 1e-3
    assert d_gpu[1] < d_r32[1] and d_gpu[1] <= 1e-5


def test_against_the_reference_kernel_restatement_exact_inputs(gpu_ctx):
    """The same comparison where FP32 absolute coordinates lose nothing:
    positions rounded to multiples of 2^-14 in a box of 100 (21 significant
    bits, so the reference's float positions and their differences are
    exact).  What remains of the reference's error is its table law and FP32
    sums, p99 about 1e-5 against the FP64 oracle on this list; the device
    must agree with it to twice that (p99 <= 2e-5, the two errors added) and
    be closer to the FP64 oracle than the restatement is."""
    w = W.config_plain(24)
    pos = np.round(w.pos * 2.0 ** 14) / 2.0 ** 14
    a = _ctx_run(gpu_ctx, w, pos=pos)
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening)
    csr = pyoracle.build_csr(w.task_t, w.task_s, w.nleaf)
    p32 = np.ascontiguousarray(pos.T.astype(np.float32))
    assert np.array_equal(p32.T.astype(np.float64), pos)            # exact in FP32
    r32 = pyoracle.p2p_ref32(prm, p32, w.leaf_ip, w.leaf_np, csr).T.astype(np.float64)
    f64 = pyoracle.p2p_f64(prm, pos, w.leaf_ip, w.leaf_np, csr)
    med, p99, mx = pyoracle.rel_err(a, r32)
    m2, p2, x2 = pyoracle.rel_err(r32, f64)
    m3, p3, x3 = pyoracle.rel_err(a, f64)
    print(f"exact inputs: device vs ref32 median {med:.2e} p99 {p99:.2e} max {mx:.2e}; "
          f"ref32 vs f64 p99 {p2:.2e} max {x2:.2e}; device vs f64 p99 {p3:.2e} max {x3:.2e}")
    assert p99 <= 2e-5 and mx <= 2e-4
    assert p3 < p2 and p3 <= 1e-5


@pytest.mark.parametrize("fraction", [0.01, 0.1, 1.0])
def test_active_subsets_of_the_clustered_mix(gpu_ctx, fraction):
    """Config 5's workload scaled to 32^3: the Zel'dovich + Plummer mix, active
    particles chosen first among Plummer members (workloads.active_subset),
    act == level (ACT_OPT, src/gpu_kernel.impl.h:392-396), targets the leaves
    holding an active particle (src/fmm.c:1303-1320).  The whole list against
    the FP64 oracle with the same mask; inactive particles get exactly zero."""
    w = W.config_clustered(32, box=62.5, n_spheres=40)
    n_grid = 32 ** 3 // 2
    act, tmask = W.active_subset(w, fraction, n_grid=n_grid)
    tt, ts = w.task_t[tmask], w.task_s[tmask]
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, box=w.box))
    gpu_ctx.set_particles(w.pos, w.leaf_ip, w.leaf_np)
    gpu_ctx.set_active(act, 1)
    gpu_ctx.set_tasks(tt, ts)
    print(f"case: clustered32 active {fraction}: {int((act == 1).sum())} active, {tt.size} tasks", flush=True)
    gpu_ctx.compute()
    a = gpu_ctx.acc()
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening)
    ref = pyoracle.p2p_f64(prm, w.pos, w.leaf_ip, w.leaf_np, pyoracle.build_csr(tt, ts, w.nleaf), act=act, level=1)
    sel = act == 1
    assert sel.any()
    assert np.all(a[~sel] == 0.0)
    med, p99, mx = pyoracle.rel_err(a[sel], ref[sel])
    print(f"active {fraction}: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4
    if fraction < 1.0:
        # Plummer members first: the active particles are sphere members
        local = w.meta["local_index"]
        assert np.all(local[sel] >= n_grid)


def test_active_subsets_at_256(gpu_ctx, clustered256):
    """The same selection on the 256^3 clustered list (size-independent checks):
    per fraction, inactive rows exactly zero, the listed pairs those of the
    active particles only, p99 <= 1e-5 on 100 sampled active targets."""
    w = clustered256
    n_grid = 256 ** 3 // 2
    gpu_ctx.set_tuning()
    for f in (1.0, 0.1, 0.01):
        act, tmask = W.active_subset(w, f, n_grid=n_grid)
        tt, ts = w.task_t[tmask], w.task_s[tmask]
        print(f"case: clustered256 active {f}: {tt.size} tasks", flush=True)
        gpu_ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, box=w.box))
        gpu_ctx.set_particles(w.pos, w.leaf_ip, w.leaf_np)
        gpu_ctx.set_active(act, 1)
        gpu_ctx.set_tasks(tt, ts)
        gpu_ctx.compute()
        a = gpu_ctx.acc()
        sel = act == 1
        assert np.all(a[~sel] == 0.0)
        n_act_t = np.bincount(np.repeat(np.arange(w.nleaf), w.leaf_np)[sel], minlength=w.nleaf)
        assert gpu_ctx.list_stats()["npairs"] == int(np.sum(n_act_t[tt].astype(np.int64) * w.leaf_np[ts]))
        tg = np.unique(tt)
        pick = np.random.default_rng(7).choice(tg, size=min(100, tg.size), replace=False)
        mk = np.isin(tt, pick)
        prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening)
        ref = pyoracle.p2p_f64(prm, w.pos, w.leaf_ip, w.leaf_np, pyoracle.build_csr(tt[mk], ts[mk], w.nleaf),
                               act=act, level=1)
        idx = np.concatenate([np.arange(w.leaf_ip[t], w.leaf_ip[t] + w.leaf_np[t]) for t in pick])
        idx = idx[sel[idx]]
        med, p99, mx = pyoracle.rel_err(a[idx], ref[idx])
        print(f"active {f}: {int(sel.sum())} active, median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
        assert p99 <= 1e-5 and mx <= 1e-4

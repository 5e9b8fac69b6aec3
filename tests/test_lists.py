"""The interaction-list producer (libphotons_host.so) against lists made by
the reference's own src/fmm.c (golden fixtures in tests/golden, and the live
reference build in oracle/_ref when it exists), plus structural invariants.
CPU only."""
import ctypes
import glob
import os

import numpy as np
import pytest

import pyoracle
from photons_p2p import workloads as W

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "lists_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_lists_identical_to_reference_fixture(path):
    f = np.load(path)
    mine = W.build_lists(f["posi"], f["tag"], float(f["int2pos"]), int(f["maxleaf"]), np.float32(f["open_angle"]),
                         np.float32(f["cutoff_radius"]), f["bdl"], f["bdr"])
    for k in ("perm", "leaf_ip", "leaf_np", "task_t", "task_s"):
        np.testing.assert_array_equal(mine[k], f[k], err_msg=k)


@pytest.mark.skipif(not pyoracle.reflist_available(), reason="reference list producer not built here")
@pytest.mark.parametrize("n1,maxleaf,theta", [(24, 16, 0.5), (20, 64, 0.4)])
def test_lists_identical_to_live_reference(n1, maxleaf, theta):
    g = np.random.default_rng(n1)
    box = 100.0
    d = box / n1
    gr = (np.arange(n1) + 0.5) * d
    pos = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3)
    # clustered: pull a third of the particles into a clump
    pos[: pos.shape[0] // 3] = 50.0 + g.normal(0, 2.0, (pos.shape[0] // 3, 3))
    pos = np.mod(pos + g.normal(0, 0.1 * d, pos.shape), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    mine = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr)
    ref = pyoracle.reflist_build(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr)
    for k in ("perm", "leaf_ip", "leaf_np", "task_t", "task_s"):
        np.testing.assert_array_equal(mine[k], ref[k], err_msg=k)


def test_list_invariants():
    w = W.config_plain(16)
    m = w.meta
    # leaves partition the particles in tree order, sizes within MAXLEAF
    assert w.leaf_ip[0] == 0 and np.all(w.leaf_ip[1:] == w.leaf_ip[:-1] + w.leaf_np[:-1])
    assert w.leaf_ip[-1] + w.leaf_np[-1] == w.npart and w.leaf_np.max() <= m["maxleaf"]
    # symmetric list (src/fmm.c:666-684): (a, b) present iff (b, a) present,
    # for pairs whose both leaves are targets
    act = m["active_leaf"] >= 0
    key = set(zip(w.task_t.tolist(), w.task_s.tolist()))
    for a, b in list(key)[:20000]:
        if act[b]:
            assert (b, a) in key
    # every target leaf holds a local particle, every self pair is listed once
    assert np.all(act[w.task_t])
    tag = m["tag"]
    for t in np.unique(w.task_t)[:500]:
        assert tag[w.leaf_ip[t]:w.leaf_ip[t] + w.leaf_np[t]].max() == 1
    selfs = w.task_t[w.task_t == w.task_s]
    assert np.unique(selfs).size == selfs.size == np.unique(w.task_t).size
    # no duplicate pairs
    assert len(key) == w.task_t.size


def test_ghost_images_cover_cutoff_shell():
    g = np.random.default_rng(0)
    box_i = 1000
    posi = g.integers(0, box_i, (2000, 3)).astype(np.int32)
    gp, src = W.ghost_images(posi, box_i, 80)
    assert np.all((gp > -80) & (gp < box_i + 80))
    assert not np.any(np.all((gp >= 0) & (gp < box_i), axis=1))
    # each ghost is an integer-box translate of its original
    d = gp.astype(np.int64) - posi[src]
    assert np.all(np.isin(d, [-box_i, 0, box_i]))
    # brute force count of images within the shell
    cnt = 0
    for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
        if not s.any():
            continue
        q = posi.astype(np.int64) + s * box_i
        cnt += int(np.all((q > -80) & (q < box_i + 80), axis=1).sum())
    assert cnt == gp.shape[0]


def test_generators_seeded_and_shaped():
    a = W.config_plain(12)
    b = W.config_plain(12)
    np.testing.assert_array_equal(a.pos, b.pos)
    np.testing.assert_array_equal(a.task_t, b.task_t)
    g = np.random.default_rng(1)
    z = W.zeldovich_grid(16, 100.0, 0.3, g)
    assert z.shape == (16 ** 3, 3) and np.all((z >= 0) & (z < 100.0))
    p = W.plummer_spheres(5000, 20, 100.0, 1.0, np.random.default_rng(2))
    assert p.shape == (5000, 3)
    h = W.nfw_halos(4000, 10, 100.0, 100000, np.random.default_rng(3))
    assert h.shape == (4000, 3)
    u = W.uneven_pairs(seed=1, lattice=16, n_pool=300, max_src=50, reach=3, total_pairs=2000)
    assert u.task_t.size == 2000 and u.leaf_np.min() >= 1 and u.leaf_np.max() <= 256


def _accept_np(wi, wj, dist, theta, rc):
    """Vectorised acceptance test (src/fmm.c:411-479) in per-operation FP32."""
    f32 = np.float32
    w = ((wi + wj).astype(np.float64) * 0.5).astype(f32)
    m = np.abs(dist) - w
    m[m <= 0] = f32(0)
    dd2 = (dist[:, 0] * dist[:, 0] + dist[:, 1] * dist[:, 1]) + dist[:, 2] * dist[:, 2]
    dm2 = (m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1]) + m[:, 2] * m[:, 2]
    wmax = np.abs(np.maximum(np.maximum(w[:, 0], w[:, 1]), w[:, 2])) * f32(2)
    r = np.where(wmax * wmax < (f32(theta) * f32(theta)) * dd2, 1, 0)
    r = np.where(dm2 > f32(rc) * f32(rc), -1, r)
    return np.where(((m[:, 0] + m[:, 1]) + m[:, 2]).astype(np.float64) < 1.0e-15, 0, r)


def _walk_levels(tree, active, theta, rc, m2l=None):
    """The walk run level by level (every frame replaced by its children in
    call order, leaf-leaf pairs carried) instead of recursively.  m2l: a list
    that receives the well-separated pairs (acceptance 1) of each level as
    cells (leaf id, or nleaf + node id)."""
    son, nc, nw = tree["node_son"], tree["node_center"], tree["node_width"]
    lc, lw = tree["leaf_center"], tree["leaf_width"]
    nleaf = lc.shape[0]
    im, jm = np.zeros(1, np.int64), np.zeros(1, np.int64)
    while True:
        li, lj = im < 0, jm < 0
        if np.all(li & lj):
            keep = active[-im - 1] >= 0
            return (-im[keep] - 1).astype(np.int32), (-jm[keep] - 1).astype(np.int32)
        k = np.zeros(im.size, np.int64)
        ch = np.zeros((im.size, 4, 2), np.int64)
        ll = li & lj
        k[ll] = (active[-im[ll] - 1] >= 0)
        ch[ll, 0] = np.stack([im[ll], jm[ll]], 1)
        sf = (im == jm) & ~li
        s = son[im[sf]]
        k[sf] = 4
        ch[sf] = np.stack([s[:, [0, 0]], s[:, [0, 1]], s[:, [1, 0]], s[:, [1, 1]]], 1)
        rest = ~ll & ~sf
        ci = np.where(li[:, None], lc[np.maximum(-im - 1, 0)], nc[np.maximum(im, 0)])
        cj = np.where(lj[:, None], lc[np.maximum(-jm - 1, 0)], nc[np.maximum(jm, 0)])
        wi = np.where(li[:, None], lw[np.maximum(-im - 1, 0)], nw[np.maximum(im, 0)])
        wj = np.where(lj[:, None], lw[np.maximum(-jm - 1, 0)], nw[np.maximum(jm, 0)])
        acc = _accept_np(wi, wj, ci - cj, theta, rc)
        opened = rest & (acc == 0)
        if m2l is not None:
            sep = rest & (acc == 1)
            cell = lambda v: np.where(v < 0, -v - 1, nleaf + v)  # noqa: E731
            m2l.extend(zip(cell(im[sep]).tolist(), cell(jm[sep]).tolist()))
        wsi = (wi[:, 0] + wi[:, 1]) + wi[:, 2]
        wsj = (wj[:, 0] + wj[:, 1]) + wj[:, 2]
        split_i = opened & ~li & (lj | (wsi > wsj))
        split_j = opened & ~split_i
        si, sj = son[np.maximum(im, 0)], son[np.maximum(jm, 0)]
        ch[split_i, 0] = np.stack([si[split_i, 0], jm[split_i]], 1)
        ch[split_i, 1] = np.stack([si[split_i, 1], jm[split_i]], 1)
        ch[split_j, 0] = np.stack([im[split_j], sj[split_j, 0]], 1)
        ch[split_j, 1] = np.stack([im[split_j], sj[split_j, 1]], 1)
        k[opened] = 2
        nxt = ch[np.arange(4)[None, :] < k[:, None]]
        im, jm = nxt[:, 0], nxt[:, 1]


@pytest.mark.parametrize("n1,maxleaf,theta", [(24, 16, 0.5), (20, 64, 0.4), (16, 32, 0.5)])
def test_level_synchronous_walk_matches_recursive(n1, maxleaf, theta):
    """The order argument of the device walk, on the CPU: expanding every
    frame into its children in call order, one level at a time, yields the
    recursive walk's list in its order."""
    g = np.random.default_rng(100 + n1)
    box = 100.0
    pos = np.mod(g.random((n1 ** 3, 3)) * box + 0.0, box)
    pos[: pos.shape[0] // 4] = np.mod(30.0 + g.normal(0, 3.0, (pos.shape[0] // 4, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    L = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr, with_tree=True)
    assert np.any(L["active"] < 0)                       # ghost-only leaves exist and are no targets
    tt, ts = _walk_levels(L, L["active"], theta, cr)
    np.testing.assert_array_equal(tt, L["task_t"])
    np.testing.assert_array_equal(ts, L["task_s"])


def test_tree_export_and_two_step_build():
    f = np.load(GOLDEN[0]) if GOLDEN else None
    w = W.config_plain(12)
    m = w.meta
    t = m["tree"]
    son = t["node_son"]
    nn = son.shape[0]
    # sons are created after their parent; every leaf and every node but the root is a son exactly once
    ids = np.arange(nn)[:, None]
    assert np.all(np.where(son >= 0, son > ids, True))
    assert np.array_equal(np.sort(son[son >= 0]), np.arange(1, nn))
    assert np.array_equal(np.sort(-son[son < 0] - 1), np.arange(w.nleaf))
    # leaf boxes hold their particles (grid positions, FP32 boxes)
    lo = t["leaf_center"] - 0.5 * t["leaf_width"]
    hi = t["leaf_center"] + 0.5 * t["leaf_width"]
    x = (m["posi"] * m["int2pos"]).astype(np.float64)
    leaf_of = np.repeat(np.arange(w.nleaf), w.leaf_np)
    tol = 1e-4 * np.abs(x).max()
    assert np.all(x >= lo[leaf_of] - tol) and np.all(x <= hi[leaf_of] + tol)
    # tree alone, then the walk: the same list as the one-step build
    if f is not None:
        args = (f["posi"], f["tag"], float(f["int2pos"]), int(f["maxleaf"]), np.float32(f["open_angle"]),
                np.float32(f["cutoff_radius"]), f["bdl"], f["bdr"])
        a = W.build_lists(*args, walk=False)
        assert a["task_t"].size == 0 and np.array_equal(a["perm"], f["perm"])


def test_m2l_pairs_are_the_well_separated_frames():
    """The host walk's M2L list (walk_task_m2l's tasks: the frames with
    acceptance 1, src/fmm.c:812-900) against the level-by-level walk's
    acceptance-1 frames, and each pair well separated by the acceptance test."""
    g = np.random.default_rng(7)
    box, n1, maxleaf, theta = 100.0, 20, 16, 0.5
    pos = g.random((n1 ** 3, 3)) * box
    pos[: pos.shape[0] // 3] = np.mod(50.0 + g.normal(0, 3.0, (pos.shape[0] // 3, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    L = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr, with_tree=True,
                      with_m2l=True)
    mine = list(zip(L["m2l_t"].tolist(), L["m2l_s"].tolist()))
    assert len(mine) > 1000 and len(set(mine)) == len(mine)
    ref = []
    _walk_levels(L, L["active"], theta, cr, m2l=ref)
    assert sorted(mine) == sorted(ref)
    nleaf = L["leaf_ip"].size
    cen = np.concatenate([L["leaf_center"], L["node_center"]])
    wid = np.concatenate([L["leaf_width"], L["node_width"]])
    t, s_ = L["m2l_t"], L["m2l_s"]
    assert np.all(_accept_np(wid[t], wid[s_], cen[t] - cen[s_], theta, cr) == 1)
    assert np.all((t >= nleaf) | (s_ >= nleaf))          # never two leaves


# the callback type of phost_walk_stream: task_compute_p2p_gpu's signature
FLUSH_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), ctypes.c_int)


@pytest.mark.parametrize("overlap,len_task", [(1, 997), (0, 997), (1, 50000)])
def test_streamed_walk_flushes_the_list_in_order(overlap, len_task):
    """phost_walk_stream hands the walk's list out in flushes of len_task
    tasks (the turn2compute_p2p protocol, src/fmm.c:500-515) while the walk
    runs: concatenated, the flushes are the list of phost_walk, with the
    leaf base added; every flush but the last is full and only the last has
    last_time = 1."""
    w = W.config_plain(16)
    m = w.meta
    lib = W._host_lib()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    posi = np.ascontiguousarray(m["posi"], np.int32)
    tag = np.ascontiguousarray(m["tag"], np.int32)
    h = lib.phost_build_tree(posi.shape[0], P(posi), P(tag), m["int2pos"], m["maxleaf"],
                             P(np.asarray(m["bdl"], np.float32)), P(np.asarray(m["bdr"], np.float32)), 4)
    assert h
    got, flags = [], []
    base = 1000

    def flush(n, tt, ts, last):
        got.append((np.ctypeslib.as_array(tt, (n,)).copy() if n else np.zeros(0, np.int32),
                    np.ctypeslib.as_array(ts, (n,)).copy() if n else np.zeros(0, np.int32)))
        flags.append(last)
        return None

    cb = FLUSH_FN(flush)
    bt = np.empty(len_task, np.int32)
    bs = np.empty(len_task, np.int32)
    try:
        assert lib.phost_walk_stream(h, np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]), 4, len_task,
                                     base, P(bt), P(bs), ctypes.cast(cb, ctypes.c_void_p), overlap) == 0
        nt = lib.phost_count(h, 3)
        tt = np.empty(nt, np.int32)
        ts = np.empty(nt, np.int32)
        lib.phost_get_tasks(h, P(tt), P(ts))
    finally:
        lib.phost_free(h)
    # the tree here is the workload's (same input, same build): the same list
    np.testing.assert_array_equal(tt, w.task_t)
    np.testing.assert_array_equal(ts, w.task_s)
    assert flags == [0] * (len(flags) - 1) + [1]
    assert all(g[0].size == len_task for g in got[:-1]) and got[-1][0].size <= len_task
    np.testing.assert_array_equal(np.concatenate([g[0] for g in got]), tt + base)
    np.testing.assert_array_equal(np.concatenate([g[1] for g in got]), ts + base)

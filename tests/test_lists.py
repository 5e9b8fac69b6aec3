"""The interaction-list producer (libphotons_host.so) against lists made by
the reference's own src/fmm.c (golden fixtures in tests/golden, and the live
reference build in oracle/_ref when it exists), plus structural invariants.
CPU only."""
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

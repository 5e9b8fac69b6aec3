"""The kd-tree built on the device (pp2p_build_tree) against the host builder,
which the list tests pin to the reference's src/fmm.c: the same particle
order, leaf ranges, target marks, node sons and FP32 boxes, bit for bit;
then the device walk over it against the host walk, and the forces of the
all-device pipeline against those of the host-built lists."""
import numpy as np
import pytest

import photons_p2p as pp
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu

TREE_KEYS = ("perm", "leaf_ip", "leaf_np", "active", "leaf_center", "leaf_width", "node_son", "node_center",
             "node_width")


def _grid_case(n1, maxleaf, seed, clump=True):
    g = np.random.default_rng(seed)
    box = 100.0
    pos = g.random((n1 ** 3, 3)) * box
    if clump:
        pos[: pos.shape[0] // 3] = np.mod(40.0 + g.normal(0, 2.5, (pos.shape[0] // 3, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    return allp, tag, int2pos, maxleaf, bdl, bdr, np.float32(cr)


def _device_tree(ctx, allp, tag, int2pos, maxleaf, bdl, bdr):
    ctx.set_tuning()
    ctx.set_law(pp.ForceLaw(eps=0.05, rs=1.0))
    nl, nn = ctx.build_tree(allp, tag, int2pos, maxleaf, bdl, bdr)
    t = ctx.get_tree()
    assert t["leaf_ip"].size == nl and t["node_son"].shape[0] == nn
    return t


def _assert_same_tree(dev, host):
    for k in TREE_KEYS:
        assert dev[k].shape == host[k].shape, k
        np.testing.assert_array_equal(dev[k], host[k], err_msg=k)


@pytest.mark.parametrize("n1,maxleaf", [(24, 16), (20, 64), (12, 8), (8, 1), (5, 2), (3, 64)])
def test_device_tree_equals_host_tree(gpu_ctx, n1, maxleaf):
    print(f"case: n1={n1} maxleaf={maxleaf}", flush=True)
    allp, tag, int2pos, maxleaf, bdl, bdr, cr = _grid_case(n1, maxleaf, seed=n1 * 3 + maxleaf)
    host = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(0.5), cr, bdl, bdr, walk=False, with_tree=True)
    dev = _device_tree(gpu_ctx, allp, tag, int2pos, maxleaf, bdl, bdr)
    print(f"N={allp.shape[0]} nodes={dev['node_son'].shape[0]} leaves={dev['leaf_ip'].size}")
    _assert_same_tree(dev, host)


@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_device_tree_tiny(gpu_ctx, n):
    """Roots of one to a few particles: an empty leaf beside a single one, the
    two-particle compare, and an all-equal coordinate (every particle at or
    below the mean: the left son gets len - 1)."""
    print(f"case: n={n}", flush=True)
    g = np.random.default_rng(n)
    posi = g.integers(1000, 2000, (n, 3)).astype(np.int32)
    if n == 7:
        posi[:, 0] = 1500
    tag = np.ones(n, np.int32)
    int2pos = 1e-3
    bdl = np.array([0.9, 0.9, 0.9], np.float32)
    bdr = np.array([2.1, 2.1, 2.1], np.float32)
    host = W.build_lists(posi, tag, int2pos, 1, np.float32(0.5), np.float32(1.0), bdl, bdr, walk=False,
                         with_tree=True)
    dev = _device_tree(gpu_ctx, posi, tag, int2pos, 1, bdl, bdr)
    _assert_same_tree(dev, host)


def test_device_tree_and_walk_plain64(gpu_ctx):
    """64^3 plain configuration, all on the device: tree, walk and forces,
    against the host-built tree and lists."""
    w = W.config_plain(64)
    m = w.meta
    print(f"case: plain64 N={w.npart} leaves={w.nleaf} tasks={w.task_t.size}", flush=True)
    law = pp.ForceLaw(eps=w.eps, rs=w.rs, softening=w.softening)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(law)
    gpu_ctx.set_particles(w.pos, w.leaf_ip, w.leaf_np)
    gpu_ctx.set_tasks(w.task_t, w.task_s)
    gpu_ctx.compute(pp.EXEC_QUEUE)
    a_host = gpu_ctx.acc()
    # the input order: tree position k holds input particle perm[k]
    perm = m["perm"]
    posi_in = np.empty_like(m["posi"])
    tag_in = np.empty_like(m["tag"])
    posi_in[perm] = m["posi"]
    tag_in[perm] = m["tag"]
    nl, nn = gpu_ctx.build_tree(posi_in, tag_in, m["int2pos"], m["maxleaf"], m["bdl"], m["bdr"])
    t = gpu_ctx.get_tree()
    np.testing.assert_array_equal(t["perm"], perm)
    np.testing.assert_array_equal(t["leaf_ip"], w.leaf_ip)
    np.testing.assert_array_equal(t["leaf_np"], w.leaf_np)
    np.testing.assert_array_equal(t["active"], m["active_leaf"])
    for k in ("node_son", "node_center", "node_width", "leaf_center", "leaf_width"):
        np.testing.assert_array_equal(t[k], m["tree"][k], err_msg=k)
    n = gpu_ctx.walk_tree(np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]))
    tt, ts = gpu_ctx.get_tasks()
    assert n == w.task_t.size
    np.testing.assert_array_equal(tt, w.task_t)
    np.testing.assert_array_equal(ts, w.task_s)
    gpu_ctx.compute(pp.EXEC_QUEUE)
    a_dev = gpu_ctx.acc()
    mag = np.linalg.norm(a_host, axis=1)
    assert np.all(np.linalg.norm(a_dev - a_host, axis=1) <= 1e-12 * np.maximum(mag, 1e-30))


def test_device_tree_rejects_bad_input_on_the_host(gpu_ctx):
    allp, tag, int2pos, maxleaf, bdl, bdr, cr = _grid_case(6, 8, seed=9)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05, rs=1.0))
    with pytest.raises(pp.P2PError):
        gpu_ctx.build_tree(allp, tag, int2pos, 0, bdl, bdr)
    with pytest.raises(pp.P2PError):
        gpu_ctx.build_tree(allp, tag, -1.0, maxleaf, bdl, bdr)
    with pytest.raises(pp.P2PError):
        gpu_ctx.build_tree(allp[:0], tag[:0], int2pos, maxleaf, bdl, bdr)
    gpu_ctx.set_particles(allp.astype(np.float64) * int2pos, np.array([0], np.int32),
                          np.array([allp.shape[0]], np.int32))
    with pytest.raises(pp.P2PError):
        gpu_ctx.walk_tree(0.5, cr)                   # no device tree since the particles were replaced
    host = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(0.5), cr, bdl, bdr, with_tree=True)
    dev = _device_tree(gpu_ctx, allp, tag, int2pos, maxleaf, bdl, bdr)
    _assert_same_tree(dev, host)
    assert gpu_ctx.walk_tree(np.float32(0.5), cr) == host["task_t"].size

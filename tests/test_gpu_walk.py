"""The device dual walk (pp2p_walk_tasks) against the host walk, which the
list tests pin to the reference's src/fmm.c: the same (target, source) pairs
in the same order, on trees with ghost-only (non-target) leaves, both
opening angles, and forces from the device-built list equal to those from
the host-built list."""
import numpy as np
import pytest

import photons_p2p as pp
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu


def _tree_case(n1, maxleaf, theta, seed):
    g = np.random.default_rng(seed)
    box = 100.0
    pos = g.random((n1 ** 3, 3)) * box
    pos[: pos.shape[0] // 3] = np.mod(40.0 + g.normal(0, 2.5, (pos.shape[0] // 3, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    L = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr, with_tree=True)
    x = allp[L["perm"]].astype(np.float64) * int2pos
    return L, x, np.float32(cr)


def _walk_on_device(ctx, L, x, theta, cr):
    ctx.set_tuning()
    ctx.set_law(pp.ForceLaw(eps=0.05, rs=float(cr) / 4.5))
    ctx.set_particles(x, L["leaf_ip"], L["leaf_np"])
    n = ctx.walk_tasks(L, L["active"], np.float32(theta), cr)
    tt, ts = ctx.get_tasks()
    assert tt.size == n
    return tt, ts


@pytest.mark.parametrize("n1,maxleaf,theta", [(24, 16, 0.5), (20, 64, 0.4), (32, 32, 0.5), (12, 8, 0.3)])
def test_device_walk_equals_host_walk(gpu_ctx, n1, maxleaf, theta):
    print(f"case: n1={n1} maxleaf={maxleaf} theta={theta}", flush=True)
    L, x, cr = _tree_case(n1, maxleaf, theta, seed=n1 * 7 + maxleaf)
    assert np.any(L["active"] < 0)
    tt, ts = _walk_on_device(gpu_ctx, L, x, theta, cr)
    print(f"tasks host {L['task_t'].size} device {tt.size}")
    np.testing.assert_array_equal(tt, L["task_t"])
    np.testing.assert_array_equal(ts, L["task_s"])


def test_device_walk_plain64_forces_identical(gpu_ctx):
    """64^3 plain configuration: the device list equals the host list, and the
    forces computed from either are equal up to the order of the FP64 accumulation."""
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
    gpu_ctx.zero_acc()
    n = gpu_ctx.walk_tasks(m["tree"], m["active_leaf"], np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]))
    tt, ts = gpu_ctx.get_tasks()
    assert n == w.task_t.size
    np.testing.assert_array_equal(tt, w.task_t)
    np.testing.assert_array_equal(ts, w.task_s)
    gpu_ctx.compute(pp.EXEC_QUEUE)
    a_dev = gpu_ctx.acc()
    # the same list; only the order of the FP64 atomic accumulation of the
    # work items' partial sums may differ between two launches
    mag = np.linalg.norm(a_host, axis=1)
    assert np.all(np.linalg.norm(a_dev - a_host, axis=1) <= 1e-12 * np.maximum(mag, 1e-30))


def test_device_walk_rejects_bad_tree_on_the_host(gpu_ctx):
    L, x, cr = _tree_case(10, 16, 0.5, seed=3)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05, rs=float(cr) / 4.5))
    gpu_ctx.set_particles(x, L["leaf_ip"], L["leaf_np"])
    bad = dict(L)
    son = L["node_son"].copy()
    son[1, 0] = 0                       # a son that is not after its parent
    bad["node_son"] = son
    with pytest.raises(pp.P2PError):
        gpu_ctx.walk_tasks(bad, L["active"], 0.5, cr)
    son = L["node_son"].copy()
    son[0, 1] = -(x.shape[0] + 5)       # a leaf that does not exist
    bad["node_son"] = son
    with pytest.raises(pp.P2PError):
        gpu_ctx.walk_tasks(bad, L["active"], 0.5, cr)
    with pytest.raises(ValueError):
        gpu_ctx.walk_tasks(L, L["active"][:-1], 0.5, cr)
    # the context still walks after the rejections
    n = gpu_ctx.walk_tasks(L, L["active"], np.float32(0.5), cr)
    assert n == L["task_t"].size

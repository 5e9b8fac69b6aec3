"""The multipole part of the short-range force on the device tree
(pp2p_fmm_short): P2M and M2M equal the oracle's restatement of the
reference's FP32 p2m / m2m (itself bit-equal to the reference) exactly; the
local expansions after M2L and L2L, and the L2P forces, agree with the FP64
oracle chain on the same cells and pairs."""
import numpy as np
import pytest

import photons_p2p as pp
import pyoracle as po
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu


def _case(n1, maxleaf, seed):
    g = np.random.default_rng(seed)
    box = 100.0
    pos = g.random((n1 ** 3, 3)) * box
    pos[: pos.shape[0] // 3] = np.mod(50.0 + g.normal(0, 3.0, (pos.shape[0] // 3, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    L = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(0.5), np.float32(cr), bdl, bdr, with_tree=True,
                      with_m2l=True)
    return allp, tag, int2pos, maxleaf, bdl, bdr, cr, L


@pytest.mark.parametrize("n1,maxleaf", [(20, 16), (16, 32)])
def test_fmm_short_against_oracle_chain(gpu_ctx, n1, maxleaf):
    allp, tag, int2pos, maxleaf, bdl, bdr, cr, T = _case(n1, maxleaf, seed=n1 + maxleaf)
    rs, mass = cr / 4.5, 0.37
    print(f"case: n1={n1} maxleaf={maxleaf} m2l pairs={T['m2l_t'].size}", flush=True)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05 * rs, rs=rs, mass=mass))
    nl, nn = gpu_ctx.build_tree(allp, tag, int2pos, maxleaf, bdl, bdr)
    gpu_ctx.zero_acc()
    ms = gpu_ctx.fmm_short(T["m2l_t"], T["m2l_s"])
    M_dev, L_dev = gpu_ctx.get_expansions()
    a_dev = gpu_ctx.acc()
    assert ms > 0 and nl == T["leaf_ip"].size and nn == T["node_son"].shape[0]

    # the reference's FP32 P2M / M2M (oracle restatement), cells: leaves, then nodes
    posi = allp[T["perm"]]
    lc, nc = T["leaf_center"], T["node_center"]
    center = np.concatenate([lc, nc])
    M = np.zeros((nl + nn, 20), np.float32)
    for l in range(nl):
        a, n = T["leaf_ip"][l], T["leaf_np"][l]
        M[l] = po.p2m_ref32(posi[a:a + n], int2pos, lc[l], mass)
    son = T["node_son"]
    cell = lambda sn: nl + sn if sn >= 0 else -sn - 1  # noqa: E731
    for k in range(nn - 1, -1, -1):                      # sons have larger ids than their node
        acc = np.zeros(20, np.float32)
        for sn in son[k]:
            acc = po.m2m_ref32(center[nl + k] - center[cell(sn)], M[cell(sn)], acc)
        M[nl + k] = acc
    assert np.array_equal(M_dev, M), np.abs(M_dev - M).max()

    # M2L and L2L in FP64 from the same multipoles
    Lf = po.m2l_tasks_f64(center.astype(np.float64), M.astype(np.float64), T["m2l_t"], T["m2l_s"], rs)
    for k in range(nn):                                  # parents before sons
        for sn in son[k]:
            Lf[cell(sn)] = po.l2l_f64(center[cell(sn)].astype(np.float64) - center[nl + k].astype(np.float64),
                                      Lf[nl + k], Lf[cell(sn)])
    worst = 0.0
    for lo, hi in ((0, 1), (1, 4), (4, 10), (10, 20)):
        sc = np.abs(Lf[:, lo:hi]).max(axis=1)
        ok = sc > 0
        worst = max(worst, float((np.abs(L_dev[ok, lo:hi] - Lf[ok, lo:hi]).max(axis=1) / sc[ok]).max()))
    print(f"L: worst per-order error {worst:.2e}")
    assert worst <= 5e-5

    # L2P at the particles
    x = posi.astype(np.float64) * int2pos
    F = np.zeros((x.shape[0], 3))
    for l in range(nl):
        a, n = T["leaf_ip"][l], T["leaf_np"][l]
        for i in range(a, a + n):
            F[i] = po.l2p_f64(x[i] - lc[l].astype(np.float64), Lf[l])
    mag = np.linalg.norm(F, axis=1)
    ok = mag > 1e-6 * mag.max()
    err = np.linalg.norm(a_dev[ok] - F[ok], axis=1) / mag[ok]
    print(f"forces: {ok.sum()} particles, p99 {np.quantile(err, 0.99):.2e}, max {err.max():.2e}")
    assert np.quantile(err, 0.99) <= 1e-4


def test_fmm_short_needs_a_device_tree(gpu_ctx):
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05, rs=1.0))
    gpu_ctx.set_particles(np.zeros((4, 3)), np.array([0], np.int32), np.array([4], np.int32))
    with pytest.raises(pp.P2PError):
        gpu_ctx.fmm_short(np.zeros(0, np.int32), np.zeros(0, np.int32))


def test_device_walk_m2l_pairs_and_all_device_multipoles(gpu_ctx):
    """The device walk records the same well-separated pairs as the host walk
    (round by round instead of depth first), and the multipole pass over them
    gives the host-list result up to the order of the FP64 sums."""
    allp, tag, int2pos, maxleaf, bdl, bdr, cr, T = _case(20, 16, seed=99)
    rs = cr / 4.5
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05 * rs, rs=rs, mass=0.5))
    gpu_ctx.build_tree(allp, tag, int2pos, maxleaf, bdl, bdr)
    n = gpu_ctx.walk_tree(np.float32(0.5), np.float32(cr))
    assert n == T["task_t"].size
    mt, ms_ = gpu_ctx.get_m2l_tasks()
    print(f"case: m2l pairs device {mt.size} host {T['m2l_t'].size}", flush=True)
    assert sorted(zip(mt.tolist(), ms_.tolist())) == sorted(zip(T["m2l_t"].tolist(), T["m2l_s"].tolist()))
    gpu_ctx.zero_acc()
    gpu_ctx.fmm_short()                                  # the walk's own pairs
    a_dev = gpu_ctx.acc()
    _, L_dev = gpu_ctx.get_expansions()
    gpu_ctx.zero_acc()
    gpu_ctx.fmm_short(T["m2l_t"], T["m2l_s"])
    a_host = gpu_ctx.acc()
    _, L_host = gpu_ctx.get_expansions()
    scale = np.abs(L_host).max(axis=0) + 1e-30
    assert np.all(np.abs(L_dev - L_host) <= 1e-6 * scale)
    mag = np.linalg.norm(a_host, axis=1).max()
    assert np.abs(a_dev - a_host).max() <= 1e-6 * mag


def test_device_walk_without_well_separated_pairs(gpu_ctx):
    """A tree too small for any well-separated pair: the device walk records
    none, and the multipole pass over them equals the pass over an empty list."""
    allp, tag, int2pos, maxleaf, bdl, bdr, cr, T = _case(6, 256, seed=5)      # 32 leaves, no M2L pair
    rs = cr / 4.5
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05 * rs, rs=rs, mass=0.5))
    gpu_ctx.build_tree(allp, tag, int2pos, maxleaf, bdl, bdr)
    gpu_ctx.walk_tree(np.float32(0.5), np.float32(cr))
    mt, ms_ = gpu_ctx.get_m2l_tasks()
    print(f"case: n1=6 maxleaf=256 m2l pairs device {mt.size} host {T['m2l_t'].size}", flush=True)
    assert mt.size == T["m2l_t"].size == 0
    gpu_ctx.zero_acc()
    gpu_ctx.fmm_short()
    a_dev = gpu_ctx.acc()
    gpu_ctx.zero_acc()
    gpu_ctx.fmm_short(T["m2l_t"], T["m2l_s"])
    a_host = gpu_ctx.acc()
    mag = max(np.linalg.norm(a_host, axis=1).max(), 1e-30)
    assert np.abs(a_dev - a_host).max() <= 1e-6 * mag

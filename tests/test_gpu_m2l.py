"""The device M2L task loop (pp2p_m2l) and L2P (pp2p_l2p) against the FP64
multipole oracle (pinned to the reference's own m2l / l2p and to finite
differences of the erfc kernel): random cells and tasks, the r^2 < 1 skip,
the well-separated pairs of a real tree, and the forces of local expansions
at the particles of real leaves."""
import numpy as np
import pytest

import photons_p2p as pp
import pyoracle as po
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu

ORDERS = ((0, 1), (1, 4), (4, 10), (10, 20))


def _check_L(L, ref, tol=2e-5):
    """Per cell and order: error relative to the largest coefficient of that
    order in the cell."""
    worst = 0.0
    for lo, hi in ORDERS:
        s = np.abs(ref[:, lo:hi]).max(axis=1)
        ok = s > 0
        e = np.abs(L[ok, lo:hi] - ref[ok, lo:hi]).max(axis=1) / s[ok]
        worst = max(worst, float(e.max(initial=0.0)))
    print(f"worst per-order relative error {worst:.2e}")
    assert worst <= tol


def _random_cells(g, n, box=30.0):
    center = g.uniform(0, box, (n, 3)).astype(np.float32)
    scale = np.array([1.0] + [0.3] * 3 + [0.1] * 6 + [0.03] * 10)
    M = (g.normal(size=(n, 20)) * scale).astype(np.float32)
    M[:, 0] = np.abs(M[:, 0]) + 0.5
    return center, M


@pytest.mark.parametrize("rs", [1.0, 2.5])
def test_m2l_random_tasks(gpu_ctx, rs):
    g = np.random.default_rng(int(rs * 100))
    n, nt = 600, 20000
    center, M = _random_cells(g, n)
    center[1] = center[0] + np.float32(0.4)            # a pair closer than 1: adds nothing
    tt = g.integers(0, n, nt).astype(np.int32)
    ts = g.integers(0, n, nt).astype(np.int32)
    tt[:3], ts[:3] = [0, 1, 5], [1, 0, 5]                # close pair both ways, and a self pair (r = 0)
    L0 = (g.normal(size=(n, 20)) * 1e-3).astype(np.float32)
    print(f"case: rs={rs} cells={n} tasks={nt}", flush=True)
    L, ms = gpu_ctx.m2l(center, M, tt, ts, rs, L0)
    ref = po.m2l_tasks_f64(center.astype(np.float64), M.astype(np.float64), tt, ts, rs, L0.astype(np.float64))
    _check_L(L, ref)
    assert ms > 0


def test_m2l_skip_and_empty(gpu_ctx):
    g = np.random.default_rng(3)
    center, M = _random_cells(g, 4)
    center[1] = center[0] + np.float32(0.5)
    L0 = np.ones((4, 20), np.float32)
    L, _ = gpu_ctx.m2l(center, M, np.array([0, 1], np.int32), np.array([1, 0], np.int32), 1.0, L0)
    assert np.array_equal(L, L0)                     # r^2 < 1 both ways: L unchanged
    L, _ = gpu_ctx.m2l(center, M, np.zeros(0, np.int32), np.zeros(0, np.int32), 1.0, L0)
    assert np.array_equal(L, L0)
    with pytest.raises(pp.P2PError):
        gpu_ctx.m2l(center, M, np.array([0], np.int32), np.array([4], np.int32), 1.0, L0)   # cell 4 does not exist


def test_m2l_tree_pairs_and_l2p(gpu_ctx):
    """The well-separated pairs of a clustered tree (cells: leaves, then
    nodes), multipoles from the oracle's P2M about the box centres; then L2P
    of the leaves' expansions at their particles."""
    g = np.random.default_rng(11)
    box, n1, maxleaf, theta = 100.0, 24, 16, 0.5
    pos = g.random((n1 ** 3, 3)) * box
    pos[: pos.shape[0] // 3] = np.mod(50.0 + g.normal(0, 3.0, (pos.shape[0] // 3, 3)), box)
    cr = 4.5 * 1.25 * box / n1
    rs = cr / 4.5
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    Lst = W.build_lists(allp, tag, int2pos, maxleaf, np.float32(theta), np.float32(cr), bdl, bdr, with_tree=True,
                        with_m2l=True)
    x = allp[Lst["perm"]].astype(np.float64) * int2pos
    nleaf = Lst["leaf_ip"].size
    center = np.concatenate([Lst["leaf_center"], Lst["node_center"]]).astype(np.float32)
    ncell = center.shape[0]
    # leaf multipoles from their particles; node multipoles random (the
    # translation is linear in M)
    M = np.zeros((ncell, 20))
    for l in range(nleaf):
        a, n = Lst["leaf_ip"][l], Lst["leaf_np"][l]
        M[l] = po.p2m_f64(x[a:a + n], 0.5, center[l].astype(np.float64))
    M[nleaf:] = g.normal(size=(ncell - nleaf, 20)) * 0.1
    M[nleaf:, 0] = np.abs(M[nleaf:, 0]) + 1.0
    M = M.astype(np.float32)
    tt, ts = Lst["m2l_t"], Lst["m2l_s"]
    print(f"case: tree n1={n1} cells={ncell} m2l tasks={tt.size}", flush=True)
    L, _ = gpu_ctx.m2l(center, M, tt, ts, rs)
    ref = po.m2l_tasks_f64(center.astype(np.float64), M.astype(np.float64), tt, ts, rs)
    _check_L(L, ref)
    # L2P at the leaves' particles
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05, rs=rs))
    gpu_ctx.set_particles(x, Lst["leaf_ip"], Lst["leaf_np"])
    gpu_ctx.zero_acc()
    Lleaf = L[:nleaf]
    gpu_ctx.l2p(Lleaf, center[:nleaf])
    a = gpu_ctx.acc()
    F = np.zeros_like(a)
    for l in range(nleaf):
        s, n = Lst["leaf_ip"][l], Lst["leaf_np"][l]
        c = center[l].astype(np.float64)
        for i in range(s, s + n):
            F[i] = po.l2p_f64(x[i] - c, Lleaf[l].astype(np.float64))
    mag = np.linalg.norm(F, axis=1)
    ok = mag > 0
    err = np.linalg.norm(a[ok] - F[ok], axis=1) / mag[ok]
    print(f"l2p: {ok.sum()} particles, p99 {np.quantile(err, 0.99):.2e} max {err.max():.2e}")
    assert np.quantile(err, 0.99) <= 1e-5 and err.max() <= 1e-4

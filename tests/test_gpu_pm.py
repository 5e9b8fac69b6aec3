"""The device particle mesh (pp2p_pm_forces) against the PM oracle
(oracle/oracle_pm.py, pinned in test_oracle_pm.py): the deposited mass mesh,
the potential and the accelerations on random and clustered particles, with
and without tags, the vanishing total force at a larger size, and the
host-side argument checks."""
import os
import sys

import numpy as np
import pytest

import photons_p2p as pp

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_pm as op  # noqa: E402

pytestmark = pytest.mark.gpu

BITS = 2 ** 30


def _particles(g, n, box, clustered):
    x = g.uniform(0, box, (n, 3))
    if clustered:
        k = n // 2
        x[:k] = (0.37 * box + g.normal(size=(k, 3)) * 0.02 * box) % box
    return np.round(x / (box / BITS)).astype(np.int64).clip(0, BITS - 1).astype(np.int32)


@pytest.mark.parametrize("N,box,clustered,tagged", [(16, 100.0, False, False), (32, 50.0, True, False),
                                                     (32, 400.0, True, True), (48, 100.0, False, True)])
def test_pm_matches_the_oracle(gpu_ctx, N, box, clustered, tagged):
    g = np.random.default_rng(N + int(box))
    n = 3000
    posi = _particles(g, n, box, clustered)
    tag = (g.uniform(size=n) < 0.7).astype(np.int32) if tagged else None
    rs, mass, int2pos = 1.25 * box / N, 0.37, box / BITS
    print(f"case: N={N} box={box} clustered={clustered} tagged={tagged}", flush=True)
    acc, ms = gpu_ctx.pm_forces(posi, tag, int2pos, box, N, rs, mass)
    ref, rho_ref, pot_ref = op.pm_forces(posi, tag, int2pos, np.float32(box), N, np.float32(rs), np.float32(mass))
    rho = gpu_ctx.pm_mesh(0)
    pot = gpu_ctx.pm_mesh(1)
    # the mass mesh: the same FP32 products summed in FP64 in another order
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-12, atol=1e-12 * rho_ref.max())
    # the potential: FP64 transforms on both sides
    assert np.abs(pot - pot_ref).max() <= 1e-10 * np.abs(pot_ref).max()
    # accelerations: FP64 on both sides, FP32 output
    scale = np.abs(ref).max()
    assert np.abs(acc - ref).max() <= 2e-6 * scale
    if tagged:
        assert np.all(acc[tag != 1] == 0.0)
    assert ms > 0.0


def test_pm_total_force_vanishes_at_size(gpu_ctx):
    """128^3 mesh, 2^20 particles, half in a clump: CIC + antisymmetric
    gradient + symmetric Green's function conserve momentum."""
    g = np.random.default_rng(77)
    box, N = 200.0, 128
    posi = _particles(g, 1 << 20, box, True)
    acc, ms = gpu_ctx.pm_forces(posi, None, box / BITS, box, N, 1.25 * box / N, 1.0)
    tot = acc.astype(np.float64).sum(0)
    print(f"sum {tot}, sum |a| {np.abs(acc).sum():.3e}, {ms:.2f} ms", flush=True)
    assert np.abs(tot).max() <= 1e-6 * np.abs(acc.astype(np.float64)).sum()


def test_pm_rejects_bad_arguments_on_the_host(gpu_ctx):
    posi = np.zeros((4, 3), np.int32)
    for N in (7, 6, 2048):
        with pytest.raises(pp.P2PError):
            gpu_ctx.pm_forces(posi, None, 1e-6, 100.0, N, 1.0, 1.0)
    with pytest.raises(pp.P2PError):
        gpu_ctx.pm_forces(posi, None, 1e-6, -1.0, 16, 1.0, 1.0)
    with pytest.raises(pp.P2PError):
        gpu_ctx.pm_forces(posi, None, 0.0, 100.0, 16, 1.0, 1.0)


def test_pm_on_the_device_tree_adds_to_the_accumulator(gpu_ctx):
    """pp2p_pm_add_tree on the particles of pp2p_build_tree (tree order, tags
    through the permutation) adds, in FP64, what pp2p_pm_forces gives for the
    same particles; ghosts get nothing; a second call adds it again."""
    from photons_p2p import workloads as W
    g = np.random.default_rng(31)
    box, n1, N = 100.0, 20, 32
    pos = g.random((n1 ** 3, 3)) * box
    pos[: pos.shape[0] // 3] = np.mod(50.0 + g.normal(0, 3.0, (pos.shape[0] // 3, 3)), box)
    rs = 1.25 * box / N
    cr = 4.5 * rs
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=0.05 * rs, rs=rs, mass=0.37))
    gpu_ctx.build_tree(allp, tag, int2pos, 16, bdl, bdr)
    perm = gpu_ctx.get_tree()["perm"]
    boxf = float(box_i * int2pos)
    ref, _ = gpu_ctx.pm_forces(allp[perm], tag[perm], int2pos, boxf, N, rs, 0.37)
    gpu_ctx.zero_acc()
    ms = gpu_ctx.pm_add_tree(boxf, N, rs, 0.37)
    a1 = gpu_ctx.acc()
    gpu_ctx.pm_add_tree(boxf, N, rs, 0.37)
    a2 = gpu_ctx.acc()
    print(f"case: {allp.shape[0]} particles ({len(gp)} ghosts), {ms:.3f} ms", flush=True)
    scale = np.abs(ref).max()
    assert np.abs(a1 - ref).max() <= 1e-6 * scale          # FP32 rounding of ref
    assert np.all(a1[tag[perm] != 1] == 0.0)
    np.testing.assert_allclose(a2, 2 * a1, rtol=1e-12, atol=1e-12 * scale)
    # the oracle on the local particles
    loc = tag[perm] == 1
    o, _, _ = op.pm_forces(allp[perm][loc], None, int2pos, np.float32(boxf), N, np.float32(rs), np.float32(0.37))
    assert np.abs(a1[loc] - o).max() <= 1e-9 * np.abs(o).max()


def test_pm_edge_positions(gpu_ctx):
    """Particles on the first and last grid points of the box, on cell faces
    and at cell centres (both CIC neighbour branches, the periodic wrap on
    both sides)."""
    box, N = 64.0, 16
    h = BITS // N                                 # grid points per cell
    c = np.array([0, 1, h // 2, h // 2 + 1, h - 1, h, BITS - h // 2, BITS - 1], np.int64)
    g = np.random.default_rng(8)
    posi = np.stack([g.choice(c, 200), g.choice(c, 200), g.choice(c, 200)], axis=1).astype(np.int32)
    acc, _ = gpu_ctx.pm_forces(posi, None, box / BITS, box, N, 1.25 * box / N, 1.0)
    ref, rho_ref, _ = op.pm_forces(posi, None, box / BITS, np.float32(box), N, np.float32(1.25 * box / N),
                                   np.float32(1.0))
    np.testing.assert_allclose(gpu_ctx.pm_mesh(0), rho_ref, rtol=1e-12, atol=1e-12 * rho_ref.max())
    assert np.abs(acc - ref).max() <= 2e-6 * np.abs(ref).max()

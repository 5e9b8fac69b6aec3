"""Edge cases of the P2P path on the GPU, each against the FP64 oracle on the
same lists: empty and ragged inputs, leaves from 1 particle up to several
work items, self pairs, repeated leaf pairs, periodic wrap, and the host-side
rejection of bad input (no kernel runs on it)."""
import glob
import os

import numpy as np
import pytest

import photons_p2p as pp
import pyoracle

pytestmark = pytest.mark.gpu

EPS, RS, BOX = 0.05, 1.0, 0.0


def _oracle(pos, ip, npl, tt, ts, softening=1, box=BOX):
    prm = pyoracle.params(EPS, RS, softening=softening, box=box)
    csr = pyoracle.build_csr(tt, ts, ip.size)
    return pyoracle.p2p_f64(prm, pos, ip, npl, csr)


def _run(ctx, pos, ip, npl, tt, ts, softening=1, box=BOX, exec_form=pp.EXEC_QUEUE):
    print(f"case: N={pos.shape[0]} leaves={ip.size} tasks={tt.size} soft={softening} box={box} "
          f"exec={exec_form}", flush=True)
    ctx.set_tuning()
    ctx.set_law(pp.ForceLaw(eps=EPS, rs=RS, softening=softening, box=box))
    ctx.set_particles(pos, ip, npl)
    ctx.set_tasks(tt, ts)
    ctx.compute(exec_form)
    return ctx.acc()


def _leaves(sizes, seed, spread=2.0, width=1.5):
    """Leaves of the given sizes, each a cube of side `width` at a random
    centre in a box of side spread * cutoff radius (2 rs ucut = 5.625)."""
    g = np.random.default_rng(seed)
    rc = 2 * RS * 2.8125
    centres = g.random((len(sizes), 3)) * spread * rc
    pos = np.concatenate([c + (g.random((n, 3)) - 0.5) * width for c, n in zip(centres, sizes)])
    npl = np.asarray(sizes, np.int32)
    ip = np.concatenate([[0], np.cumsum(npl)[:-1]]).astype(np.int32)
    return pos, ip, npl


def _check(a, ref, p99=1e-5, mx=1e-4):
    med, q, m = pyoracle.rel_err(a, ref)
    print(f"rel err median {med:.2e} p99 {q:.2e} max {m:.2e}")
    assert q <= p99 and m <= mx


def test_empty_list_leaves_zero(gpu_ctx):
    pos, ip, npl = _leaves([5, 7], seed=1)
    a = _run(gpu_ctx, pos, ip, npl, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert np.all(a == 0.0)


@pytest.mark.parametrize("exec_form", [pp.EXEC_QUEUE, pp.EXEC_GRID, pp.EXEC_PER_TARGET])
def test_ragged_leaf_sizes(gpu_ctx, exec_form):
    """1 .. 700 particles per leaf: every lane shape, i-chunks beyond one item,
    all pairs of leaves listed both ways."""
    sizes = [1, 2, 3, 7, 20, 31, 33, 48, 63, 64, 65, 96, 127, 191, 192, 193, 257, 700]
    pos, ip, npl = _leaves(sizes, seed=2)
    n = len(sizes)
    tt, ts = np.meshgrid(np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32), indexing="ij")
    tt, ts = tt.ravel(), ts.ravel()
    ref = _oracle(pos, ip, npl, tt, ts)
    a = _run(gpu_ctx, pos, ip, npl, tt, ts, exec_form=exec_form)
    _check(a, ref)


def test_self_pairs_only(gpu_ctx):
    sizes = [1, 9, 64, 150]
    pos, ip, npl = _leaves(sizes, seed=3, width=0.8)
    tt = np.arange(len(sizes), dtype=np.int32)
    ref = _oracle(pos, ip, npl, tt, tt)
    a = _run(gpu_ctx, pos, ip, npl, tt, tt)
    assert np.all(a[0] == 0.0)                    # a lone particle feels nothing from itself
    _check(a[1:], ref[1:])


def test_repeated_leaf_pairs_count_twice(gpu_ctx):
    """A leaf pair listed twice contributes twice, as the reference's loop over
    its task list does."""
    pos, ip, npl = _leaves([40, 25, 60], seed=4, spread=0.3)        # all within the cutoff
    tt = np.array([0, 0, 1, 2, 0], np.int32)
    ts = np.array([1, 1, 2, 0, 2], np.int32)
    ref = _oracle(pos, ip, npl, tt, ts)
    a = _run(gpu_ctx, pos, ip, npl, tt, ts)
    _check(a, ref)
    once = _oracle(pos, ip, npl, tt[1:], ts[1:])
    assert not np.allclose(ref[ip[0]:ip[0] + npl[0]], once[ip[0]:ip[0] + npl[0]])


@pytest.mark.parametrize("softening", [0, 1])
def test_periodic_wrap_across_the_box(gpu_ctx, softening):
    """Two leaves on opposite faces of a periodic box, within the cutoff of each
    other only through the minimum image."""
    box = 40.0
    g = np.random.default_rng(5)
    a0 = np.array([0.5, 20.0, 20.0]) + (g.random((30, 3)) - 0.5) * 0.8
    a1 = np.array([box - 0.8, 20.3, 19.6]) + (g.random((45, 3)) - 0.5) * 0.8
    pos = np.concatenate([a0, a1])
    ip = np.array([0, 30], np.int32)
    npl = np.array([30, 45], np.int32)
    tt = np.array([0, 1, 0, 1], np.int32)
    ts = np.array([1, 0, 0, 1], np.int32)
    ref = _oracle(pos, ip, npl, tt, ts, softening=softening, box=box)
    a = _run(gpu_ctx, pos, ip, npl, tt, ts, softening=softening, box=box)
    _check(a, ref)
    # the leaves are ~39 apart inside the box, far beyond the cutoff: the cross
    # pairs act only through the minimum image
    cross = _oracle(pos, ip, npl, tt[:2], ts[:2], softening=softening, box=box)
    assert np.all(np.linalg.norm(cross, axis=1) > 0)


def test_bad_input_is_rejected_on_the_host(gpu_ctx):
    pos, ip, npl = _leaves([4, 4], seed=6)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=EPS, rs=RS))
    gpu_ctx.set_particles(pos, ip, npl)
    with pytest.raises(pp.P2PError):
        gpu_ctx.set_tasks(np.array([0, 2], np.int32), np.array([1, 0], np.int32))      # leaf 2 does not exist
    with pytest.raises(pp.P2PError):
        gpu_ctx.set_tasks(np.array([-1], np.int32), np.array([0], np.int32))
    with pytest.raises(pp.P2PError):
        gpu_ctx.set_particles(pos, np.array([0, 6], np.int32), npl)                       # leaf runs past N
    with pytest.raises(pp.P2PError):
        gpu_ctx.set_law(pp.ForceLaw(eps=1e-9, rs=RS))                                     # below the table range
    with pytest.raises(pp.P2PError):
        gpu_ctx.set_law(pp.ForceLaw(eps=20.0, rs=RS))                                     # softening beyond the cutoff
    # the context still works after the rejections
    tt = np.array([0, 1], np.int32)
    ts = np.array([1, 0], np.int32)
    ref = _oracle(pos, ip, npl, tt, ts)
    a = _run(gpu_ctx, pos, ip, npl, tt, ts)
    _check(a, ref)


def test_overlapping_leaves_are_rejected(gpu_ctx):
    """Leaves own their particles: nested or partly overlapping non-empty
    ranges are refused on the host (the frames and i-order lists are written
    per particle); empty leaves may sit anywhere; the context stays usable."""
    pos, ip, npl = _leaves([20, 20], seed=8)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=EPS, rs=RS))
    for bad_ip, bad_np in [([0, 10], [20, 20]),            # partial overlap
                           ([0, 5], [30, 10]),             # nested
                           ([7, 0], [10, 40])]:            # overlap, out of order
        with pytest.raises(pp.P2PError, match="share particles"):
            gpu_ctx.set_particles(pos, np.array(bad_ip, np.int32), np.array(bad_np, np.int32))
    # empty leaves inside another leaf's range, in any order, are accepted
    ip3 = np.array([20, 0, 5, 40], np.int32)
    np3 = np.array([20, 20, 0, 0], np.int32)
    tt = np.array([0, 1, 0, 1, 2, 1], np.int32)
    ts = np.array([1, 0, 0, 1, 1, 2], np.int32)
    ref = _oracle(pos, ip3, np3, tt, ts)
    a = _run(gpu_ctx, pos, ip3, np3, tt, ts)
    _check(a, ref)


def test_trailing_empty_source_at_particle_zero(gpu_ctx):
    """An empty source leaf starting at particle 0, listed last after a source
    of more than 64 particles: lanes past the end of the staged sources meet
    it while prefetching the next window of 64, and must stay in bounds
    (index clamped at 0).  Results equal the oracle's; the empty leaf adds
    nothing."""
    pos, ip, npl = _leaves([100, 30], seed=9, spread=0.4)
    ip2 = np.array([0, 0, 100], np.int32)                     # leaf 0 empty at particle 0
    np2 = np.array([0, 100, 30], np.int32)
    # host-side bounds check of the prefetch: lanes 64..127 of the second
    # window locate source 1 (the empty leaf 0) once past 100 particles, whose
    # base is ip - (prefix before it) = 0 - 100; J2 clamps to total - 1 = 99
    base_empty = 0 - 100
    assert max(base_empty + 99, 0) == 0
    tt = np.array([2, 2, 1, 1], np.int32)
    ts = np.array([1, 0, 2, 0], np.int32)
    ref = _oracle(pos, ip2, np2, tt, ts)
    for exec_form in (pp.EXEC_QUEUE, pp.EXEC_GRID):
        a = _run(gpu_ctx, pos, ip2, np2, tt, ts, exec_form=exec_form)
        _check(a, ref)


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lists_*.npz")))


@pytest.mark.parametrize("path", GOLDEN + ["uneven"], ids=[os.path.basename(p) for p in GOLDEN] + ["uneven"])
def test_device_csr_equals_the_stable_sort(gpu_ctx, path):
    """The target CSR the device builds (rocPRIM radix sort + row kernels) is
    the one of the reference's thrust sort_by_key -> reduce_by_key ->
    exclusive_scan (src/gpu_entry.cu:132-138), bit for bit: a stable sort of
    the list by target (oracle_build_csr, a counting sort), the distinct
    targets in order and their offsets.  Lists: the reference's own (golden
    fixtures from src/fmm.c) and an uneven random list in random order."""
    from photons_p2p import workloads as W
    if path == "uneven":
        w = W.uneven_pairs(seed=21, lattice=24, n_pool=3000, max_src=120, reach=4, total_pairs=40000)
        g = np.random.default_rng(1)
        order = g.permutation(w.task_t.size)
        pos, ip, npl, tt, ts = w.pos, w.leaf_ip, w.leaf_np, w.task_t[order], w.task_s[order]
    else:
        f = np.load(path)
        pos = f["posi"][f["perm"]].astype(np.float64) * float(f["int2pos"])
        ip, npl, tt, ts = f["leaf_ip"], f["leaf_np"], f["task_t"], f["task_s"]
    print(f"case: csr {path}: {tt.size} tasks, {ip.size} leaves", flush=True)
    gpu_ctx.set_tuning()
    gpu_ctx.set_law(pp.ForceLaw(eps=EPS, rs=RS))
    gpu_ctx.set_particles(pos, ip, npl)
    gpu_ctx.set_tasks(tt, ts)
    c = gpu_ctx.csr()
    tgt, off, src = pyoracle.build_csr(tt, ts, ip.size)
    np.testing.assert_array_equal(c["key_sorted"], np.repeat(tgt, np.diff(off)))
    np.testing.assert_array_equal(c["src"], src)
    np.testing.assert_array_equal(c["row_tgt"], tgt)
    np.testing.assert_array_equal(c["row_off"], off)
    # and the same as numpy's stable argsort of the targets
    o = np.argsort(tt, kind="stable")
    np.testing.assert_array_equal(c["src"], ts[o])


def _block_edge_case():
    """A target whose CSR sources fall in 64-source blocks of 0, 64, 65, 128
    and (partial block) 129 particles, with empty leaves at particle 0, in the
    middle and at the end of the particle array; a second target whose
    sources are empty leaves only (and itself)."""
    g = np.random.default_rng(17)
    rc = 2 * RS * 2.8125
    sizes = [20, 5] + [1] * 64 + [1] * 63 + [2] + [64, 64] + [129]
    pos = np.concatenate([(g.random(3) * 0.8 * rc) + (g.random((n, 3)) - 0.5) * 1.0 for n in sizes])
    npl = np.asarray(sizes, np.int32)
    ip = np.concatenate([[0], np.cumsum(npl)[:-1]]).astype(np.int32)
    npart = int(npl.sum())
    full = ip.size
    # empty leaves: 64 + 62 + 10 of them, at particle 0, mid-array and the end
    e_at = np.array([0, npart // 2, npart] * 46, np.int32)[:136]
    ip = np.concatenate([ip, e_at])
    npl = np.concatenate([npl, np.zeros(e_at.size, np.int32)])
    E = list(range(full, full + e_at.size))
    singles = list(range(2, 66))
    b2 = list(range(66, 129)) + [129]                         # 63 singles + the 2-particle leaf
    b3 = [130, 131] + E[64:126]                               # two 64-particle leaves + 62 empty
    b4 = [132, E[126], E[0]]                                  # 129 particles, then two empty leaves
    t1 = E[:64] + singles + b2 + b3 + b4 + [0]
    t2 = E[126:136] + E[:64] + [1]
    tt = np.array([0] * len(t1) + [1] * len(t2), np.int32)
    ts = np.array(t1 + t2, np.int32)
    # per-block particle totals of target 0's sources, as the kernel sees them
    tot = [int(npl[ts[:len(t1)][k:k + 64]].sum()) for k in range(0, len(t1), 64)]
    assert tot[:4] == [0, 64, 65, 128] and tot[4] == 129 + 20
    # a window load past a block's end clamps its index to [0, npart - 1]:
    # from an empty leaf at ip the index is max(ip - 1, 0)
    assert all(0 <= max(int(x) - 1, 0) <= npart - 1 for x in e_at)
    return pos, ip, npl, tt, ts


@pytest.mark.parametrize("law", [pp.LAW_TABLE, pp.LAW_LINEAR], ids=["table", "linear"])
def test_source_blocks_of_zero_and_window_multiples(gpu_ctx, law):
    """Source blocks of 64 leaves holding no particle at all, exactly one or
    two windows, one window and one particle, and an odd number of windows
    with trailing empty leaves: the window loop stages two windows per pass
    and refills past a block's end from clamped indices.  Forces equal the
    oracle's; the target whose sources are all empty gets only its self
    pairs."""
    pos, ip, npl, tt, ts = _block_edge_case()
    ref = _oracle(pos, ip, npl, tt, ts)
    for exec_form in (pp.EXEC_QUEUE, pp.EXEC_GRID):
        print(f"case: source blocks law={law} exec={exec_form}", flush=True)
        gpu_ctx.set_tuning()
        gpu_ctx.set_law(pp.ForceLaw(eps=EPS, rs=RS, softening=1, box=BOX, law=law))
        gpu_ctx.set_particles(pos, ip, npl)
        gpu_ctx.set_tasks(tt, ts)
        gpu_ctx.compute(exec_form)
        a = gpu_ctx.acc()
        _check(a, ref)

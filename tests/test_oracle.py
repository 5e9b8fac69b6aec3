"""Pins of the CPU oracle (no GPU).

The reference ships no tests, fixtures or golden vectors for the P2P path
(its CPU P2P is a stub, src/fmm.c:484-491), so the oracle is pinned by the
mathematics: closed-form values of g(u), the table identity of the
reference's prologue, the Taylor coefficients of its pair function, brute
force on tiny inputs written independently here, physical invariants, and
the CSR build against an independent stable sort.  Each check is chosen so
that a plausible slip (dropped term, wrong sign, transposed index, wrong
softening branch) fails it.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

import pyoracle

C = 2.0 / math.sqrt(math.pi)


def g_ref(u):
    return math.erfc(u) + C * u * math.exp(-u * u)


# ---------------------------------------------------------------- g(u)

@pytest.mark.parametrize("u,val", [(0.0, 1.0), (0.5, 0.9188914117), (1.0, 0.5724067045), (2.0, 4.601170569e-2),
                                   (2.25, 1.753293439e-2), (2.8125, 1.234340844e-3), (3.0, 4.398496528e-4)])
def test_g_closed_form_values(u, val):
    assert pyoracle.g(u) == pytest.approx(val, rel=2e-9)


def test_g_derivative_identity():
    # g'(u) = -(4/sqrt(pi)) u^2 exp(-u^2): the erfc and Gaussian terms of the
    # PM split combine so that the short-range force derivative is this
    for u in np.linspace(0.05, 3.0, 40):
        h = 1e-5
        num = (pyoracle.g(u + h) - pyoracle.g(u - h)) / (2 * h)
        assert num == pytest.approx(-2 * C * u * u * math.exp(-u * u), rel=1e-6, abs=1e-12)


def test_g_split_complements_newton():
    # erfc-part + erf-part of a point mass = Newton: g(u) + [erf(u) - C u e^{-u^2}] = 1
    for u in (0.1, 0.7, 1.9, 2.8125):
        long_part = math.erf(u) - C * u * math.exp(-u * u)
        assert pyoracle.g(u) + long_part == pytest.approx(1.0, abs=1e-14)


# ---------------------------------------------------------------- reference tables

def test_ref_tables_match_prologue_definition():
    t0, t1 = pyoracle.ref_tables()
    x = np.arange(512) * 3.0 / 512
    np.testing.assert_allclose(t0, (erfc(x) + C * x * np.exp(-x * x)).astype(np.float32), rtol=0, atol=0)
    np.testing.assert_allclose(t1, (-2 * C * np.exp(-x * x)).astype(np.float32), rtol=0, atol=0)
    # g'(x) = x^2 tab1(x): the identity the T4 Taylor expansion is built on
    for k in (10, 100, 300, 500):
        xx = x[k]
        h = 1e-6
        d = (g_ref(xx + h) - g_ref(xx - h)) / (2 * h)
        assert d == pytest.approx(xx * xx * float(t1[k]), rel=1e-5, abs=1e-9)


def test_ref32_pair_taylor_against_exact_law():
    """One pair of the FP32 restatement against the exact force law: the table
    + T4 Taylor reproduces g to FP32 accuracy over the whole cutoff range, the
    sign is attractive (towards j) and the cutoff edge is at u = 2.8125."""
    t0, t1 = pyoracle.ref_tables()
    lib = pyoracle.load()
    rs = 1.0
    eps = 1e-3
    for u in np.linspace(0.01, 2.8, 57):
        r = 2 * rs * u
        xi = np.zeros(3, np.float32)
        xj = np.array([r, 0, 0], np.float32)
        acc = np.zeros(3, np.float32)
        lib.oracle_pair_ref32(0, np.float32(eps), np.float32(1.0), np.float32(1.0 / rs), pyoracle._p(t0),
                              pyoracle._p(t1), pyoracle._p(xi), pyoracle._p(xj), pyoracle._p(acc))
        r32 = float(xj[0])
        exact = g_ref(0.5 * r32 / rs) / (r32 * r32)
        assert acc[0] > 0 and acc[1] == 0 and acc[2] == 0
        assert float(acc[0]) == pytest.approx(exact, rel=3e-6)
    # beyond the cutoff: nothing
    acc = np.zeros(3, np.float32)
    xj = np.array([2 * rs * 2.82, 0, 0], np.float32)
    lib.oracle_pair_ref32(0, np.float32(eps), np.float32(1.0), np.float32(1.0), pyoracle._p(t0), pyoracle._p(t1),
                          pyoracle._p(np.zeros(3, np.float32)), pyoracle._p(xj), pyoracle._p(acc))
    assert np.all(acc == 0)


# ---------------------------------------------------------------- FP64 direct sum

def brute_force(pos, pairs, eps, rs, ucut, mass, softening, box=0.0):
    """Independent pair-by-pair sum over (i, j) particle pairs."""
    acc = np.zeros_like(pos)
    for i, j in pairs:
        d = pos[j] - pos[i]
        if box > 0:
            d = d - box * np.round(d / box)
        r2 = float(d @ d)
        if softening == 0:
            reff = math.sqrt(r2) if r2 > eps * eps else eps
        else:
            reff = math.sqrt(r2 + eps * eps)
        u = 0.5 * reff / rs
        if u > ucut:
            continue
        acc[i] += mass * d * g_ref(u) / reff ** 3
    return acc


def tiny_lists(seed=0, nleaf=6, box=0.0):
    g = np.random.default_rng(seed)
    npl = g.integers(1, 6, nleaf).astype(np.int32)
    ip = np.concatenate([[0], np.cumsum(npl)[:-1]]).astype(np.int32)
    n = int(npl.sum())
    pos = g.random((n, 3)) * 3.0
    tt, ts = [], []
    for t in range(nleaf):
        for s in range(nleaf):
            if g.random() < 0.7:
                tt.append(t)
                ts.append(s)
    return pos, ip, npl, np.array(tt, np.int32), np.array(ts, np.int32)


@pytest.mark.parametrize("softening", [0, 1])
def test_f64_equals_brute_force(softening):
    pos, ip, npl, tt, ts = tiny_lists(seed=softening)
    eps, rs, ucut, mass = 0.05, 0.4, 2.8125, 0.7
    prm = pyoracle.params(eps, rs, mass=mass, ucut=ucut, softening=softening)
    acc = pyoracle.p2p_f64(prm, pos, ip, npl, pyoracle.build_csr(tt, ts, ip.size))
    pairs = [(i, j) for t, s in zip(tt, ts) for i in range(ip[t], ip[t] + npl[t]) for j in range(ip[s], ip[s] + npl[s])]
    ref = brute_force(pos, pairs, eps, rs, ucut, mass, softening)
    np.testing.assert_allclose(acc, ref, rtol=1e-12, atol=1e-12)


def test_f64_periodic_minimum_image():
    pos, ip, npl, tt, ts = tiny_lists(seed=4)
    box = 3.2
    prm = pyoracle.params(0.05, 0.2, softening=1, box=box)
    acc = pyoracle.p2p_f64(prm, pos, ip, npl, pyoracle.build_csr(tt, ts, ip.size))
    pairs = [(i, j) for t, s in zip(tt, ts) for i in range(ip[t], ip[t] + npl[t]) for j in range(ip[s], ip[s] + npl[s])]
    np.testing.assert_allclose(acc, brute_force(pos, pairs, 0.05, 0.2, 2.8125, 1.0, 1, box), rtol=1e-12, atol=1e-12)


def test_newtonian_limit_and_softening_branches():
    # r << rs: g -> 1, a = m d / r^3; r < eps (REF): r_eff = eps exactly
    pos = np.array([[0.0, 0, 0], [1e-3, 0, 0], [0, 3e-5, 0]])
    ip = np.array([0, 1, 2], np.int32)
    npl = np.array([1, 1, 1], np.int32)
    csr = pyoracle.build_csr(np.array([0, 0], np.int32), np.array([1, 2], np.int32), 3)
    rs = 10.0
    eps = 1e-4
    a = pyoracle.p2p_f64(pyoracle.params(eps, rs, softening=0), pos, ip, npl, csr)
    ex = 1e-3 / 1e-9 * g_ref(0.5e-3 / rs) + 0.0
    assert a[0, 0] == pytest.approx(ex, rel=1e-12)
    assert a[0, 1] == pytest.approx(3e-5 / eps ** 3 * g_ref(0.5 * eps / rs), rel=1e-12)
    assert a[0, 0] == pytest.approx(1e6, rel=1e-6)


def test_self_pair_and_cutoff_edge():
    rs = 1.0
    ucut = 2.8125
    r_in = 2 * rs * ucut * (1 - 1e-9)
    r_out = 2 * rs * ucut * (1 + 1e-9)
    pos = np.array([[0.0, 0, 0], [r_in, 0, 0], [0, r_out, 0]])
    ip = np.array([0, 1, 2], np.int32)
    npl = np.array([1, 1, 1], np.int32)
    csr = pyoracle.build_csr(np.array([0, 0, 0], np.int32), np.array([0, 1, 2], np.int32), 3)
    a = pyoracle.p2p_f64(pyoracle.params(0.01, rs, softening=0), pos, ip, npl, csr)
    assert a[0, 1] == 0.0                       # beyond the cutoff
    assert a[0, 0] == pytest.approx(g_ref(ucut) / r_in ** 2, rel=1e-6)   # self pair adds nothing
    n_list, n_in = pyoracle.count_pairs(pyoracle.params(0.01, rs, softening=0), pos, ip, npl, csr)
    assert (n_list, n_in) == (3, 2)


def test_momentum_conservation_symmetric_lists():
    # the walk's lists are symmetric ((a,b) and (b,a)); without ghosts the pair
    # forces cancel: sum_i m a_i = 0
    pos, ip, npl, _, _ = tiny_lists(seed=7, nleaf=8)
    L = ip.size
    tt = np.repeat(np.arange(L), L).astype(np.int32)
    ts = np.tile(np.arange(L), L).astype(np.int32)
    a = pyoracle.p2p_f64(pyoracle.params(0.05, 0.5, softening=1), pos, ip, npl, pyoracle.build_csr(tt, ts, L))
    scale = np.abs(a).sum()
    assert np.abs(a.sum(axis=0)).max() <= 1e-13 * scale


def test_translation_invariance():
    pos, ip, npl, tt, ts = tiny_lists(seed=9)
    prm = pyoracle.params(0.05, 0.4, softening=0)
    csr = pyoracle.build_csr(tt, ts, ip.size)
    a = pyoracle.p2p_f64(prm, pos, ip, npl, csr)
    b = pyoracle.p2p_f64(prm, pos + np.array([123.25, -7.5, 1e3]), ip, npl, csr)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


def test_active_mask_semantics():
    pos, ip, npl, tt, ts = tiny_lists(seed=11)
    prm = pyoracle.params(0.05, 0.4, softening=1)
    csr = pyoracle.build_csr(tt, ts, ip.size)
    full = pyoracle.p2p_f64(prm, pos, ip, npl, csr)
    act = (np.arange(pos.shape[0]) % 3 == 0).astype(np.uint8) * 5
    part = pyoracle.p2p_f64(prm, pos, ip, npl, csr, act=act, level=5)
    sel = act == 5
    np.testing.assert_array_equal(part[sel], full[sel])
    assert np.all(part[~sel] == 0)


def test_ref32_close_to_f64_on_small_lists():
    pos, ip, npl, tt, ts = tiny_lists(seed=13, nleaf=10)
    prm = pyoracle.params(0.02, 0.4, softening=0)
    csr = pyoracle.build_csr(tt, ts, ip.size)
    a64 = pyoracle.p2p_f64(prm, pos, ip, npl, csr)
    a32 = pyoracle.p2p_ref32(prm, pos.T.astype(np.float32), ip, npl, csr)
    med, p99, mx = pyoracle.rel_err(a32.T, a64)
    assert p99 < 2e-5


# ---------------------------------------------------------------- CSR

def test_csr_is_stable_sort_by_target():
    g = np.random.default_rng(1)
    tt = g.integers(0, 50, 5000).astype(np.int32)
    ts = g.integers(0, 50, 5000).astype(np.int32)
    tgt, off, src = pyoracle.build_csr(tt, ts, 50)
    order = np.argsort(tt, kind="stable")
    np.testing.assert_array_equal(src, ts[order])
    u, counts = np.unique(tt, return_counts=True)
    np.testing.assert_array_equal(tgt, u)
    np.testing.assert_array_equal(np.diff(off), counts)
    assert off[0] == 0 and off[-1] == tt.size


def test_csr_rejects_bad_ids():
    with pytest.raises(ValueError):
        pyoracle.build_csr(np.array([0, 5], np.int32), np.array([0, 1], np.int32), 5)

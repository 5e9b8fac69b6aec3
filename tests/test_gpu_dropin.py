"""The drop-in libp2p_hip.so driven through the reference's call protocol.

A test host library defines the host-program globals (part, leaf, packlist,
acc, ...); this test fills them the way the reference's gravity/fmm code does
(Body records with integer positions, leaf Pack records offset by
first_leaf = NPART, flushes of LEN_TASK tasks with absolute leaf ids) and
calls the reference entry points in the order of src/fmm.c:1243-1292 and
src/gravity.c:441-447.  The accelerations added into acc[] are compared with
the FP64 oracle."""
import ctypes
import os

import numpy as np
import pytest

import photons_p2p as pp
import pyoracle
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "support", "libhostglobals.so")


class Body(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int), ("act", ctypes.c_int), ("posi", ctypes.c_int * 3),
                ("vel", ctypes.c_float * 3), ("id", ctypes.c_ulong)]


class Pack(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("npart", ctypes.c_int), ("ipart", ctypes.c_int),
                ("width", ctypes.c_float * 3), ("center", ctypes.c_float * 3), ("M", ctypes.c_float * 20),
                ("L", ctypes.c_float * 20), ("acc_pm", ctypes.c_float * 3)]


class P2PPack(ctypes.Structure):
    _fields_ = [("rankid", ctypes.c_int), ("ipart", ctypes.c_int), ("npart", ctypes.c_int),
                ("ileaf", ctypes.c_int), ("nleaf", ctypes.c_int), ("ntask", ctypes.c_int),
                ("task_t", ctypes.POINTER(ctypes.c_int)), ("task_s", ctypes.POINTER(ctypes.c_int))]


def test_abi_record_sizes():
    # inc/typedef.h under INTXYZ / FP32: 40-byte Body, 208-byte Pack, 40-byte P2PPack
    assert ctypes.sizeof(Body) == 40 and ctypes.sizeof(Pack) == 208 and ctypes.sizeof(P2PPack) == 40


@pytest.fixture(scope="module")
def libs():
    host = ctypes.CDLL(HOST, mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(pp.DROPIN_LIB)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    lib.task_compute_p2p_gpu.argtypes = [i32, ip, ip, i32]
    lib.task_compute_p2p_gpu.restype = vp
    lib.task_compute_p2p_act_gpu.argtypes = [i32, ip, ip, i32, i32]
    lib.task_compute_p2p_act_gpu.restype = vp
    lib.p2ppack_compute_gpu.argtypes = [ctypes.POINTER(P2PPack), i32]
    lib.p2ppack_compute_gpu.restype = vp
    lib.hostmalloc_task.argtypes = [i32]
    lib.hostmalloc_out.argtypes = [i32]
    lib.devMemcpyDH_updateAcc.argtypes = [ctypes.POINTER(P2PPack), i32, i32]
    lib.get_last_kntime.argtypes = [i32]
    lib.get_last_kntime.restype = ctypes.c_double
    lib.p2p_hip_status.restype = i32
    lib.p2p_hip_error.restype = ctypes.c_char_p
    return host, lib


class Host:
    """Fills the host globals from a tree workload (leaf ids offset by NPART)."""

    def __init__(self, host, w, mass=1.0):
        self.h = host
        n = w.npart
        self.body = (Body * n)()
        posi = w.meta["posi"]
        arr = np.ctypeslib.as_array(self.body)
        arr["tag"] = w.meta["tag"]
        arr["posi"] = posi
        arr["id"] = np.arange(n)
        nl = w.nleaf
        self.first_leaf = n
        self.leaves = (Pack * nl)()
        la = np.ctypeslib.as_array(self.leaves)
        la["ipart"] = w.leaf_ip
        la["npart"] = w.leaf_np
        self.acc = [np.zeros(n, np.float32) for _ in range(3)]
        self.packs = (P2PPack * 9)()
        g = lambda name, t: t.in_dll(host, name)  # noqa: E731
        g("NPART", ctypes.c_int).value = n
        g("first_leaf", ctypes.c_int).value = n
        g("last_leaf", ctypes.c_int).value = n + nl
        g("NPACK", ctypes.c_int).value = 0
        g("MAXNPACK", ctypes.c_int).value = 9
        g("SoftenScale", ctypes.c_float).value = w.eps
        g("splitRadius", ctypes.c_float).value = w.rs
        g("MASSPART", ctypes.c_float).value = mass
        g("INT2POS", ctypes.c_double).value = w.meta["int2pos"]
        ctypes.c_void_p.in_dll(host, "part").value = ctypes.addressof(self.body)
        # leaf[first_leaf + l] is the l-th leaf record (src/fmm.c:393)
        ctypes.c_void_p.in_dll(host, "leaf").value = ctypes.addressof(self.leaves) - n * ctypes.sizeof(Pack)
        ctypes.c_void_p.in_dll(host, "packlist").value = ctypes.addressof(self.packs)
        accp = (ctypes.c_void_p * 3).in_dll(host, "acc")
        for d in range(3):
            accp[d] = self.acc[d].ctypes.data

    def tasks(self):
        tt = ctypes.c_void_p.in_dll(self.h, "task_t").value
        ts = ctypes.c_void_p.in_dll(self.h, "task_s").value
        return ctypes.cast(tt, ctypes.POINTER(ctypes.c_int)), ctypes.cast(ts, ctypes.POINTER(ctypes.c_int))

    def npack(self):
        return ctypes.c_int.in_dll(self.h, "NPACK").value

    def acc_aos(self):
        return np.stack(self.acc, axis=1).astype(np.float64)


def _oracle(w, mass=1.0, act=None, level=0):
    prm = pyoracle.params(w.eps, w.rs, mass=mass, softening=0)
    csr = pyoracle.build_csr(w.task_t, w.task_s, w.nleaf)
    return pyoracle.p2p_f64(prm, w.pos, w.leaf_ip, w.leaf_np, csr, act=act, level=level)


def test_flushed_step_matches_oracle(libs):
    host, lib = libs
    w = W.config_plain(20, softening=0)
    H = Host(host, w, mass=0.5)
    len_task = 5000                                   # several flushes per step
    lib.hostmalloc_task(len_task)
    tt, ts = H.tasks()
    nt = w.task_t.size
    abs_t = w.task_t + H.first_leaf
    abs_s = w.task_s + H.first_leaf
    k = 0
    while nt - k > len_task:                          # turn2compute_p2p, src/fmm.c:500-515
        for j in range(len_task):
            tt[j], ts[j] = int(abs_t[k + j]), int(abs_s[k + j])
        lib.task_compute_p2p_gpu(len_task, tt, ts, 0)
        assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
        assert tt[0] == abs_t[k]                      # caller's buffer not rebased
        k += len_task
    rest = nt - k
    for j in range(rest):
        tt[j], ts[j] = int(abs_t[k + j]), int(abs_s[k + j])
    lib.task_compute_p2p_gpu(rest, tt, ts, 1)
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    lib.hostfree_task()
    t = lib.get_last_kntime(0)
    assert t > 0.0
    lib.hostmalloc_out(w.npart)
    lib.devMemcpyDH_updateAcc(H.packs, H.npack(), -1)
    lib.hostfree_out()
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    ref = _oracle(w, mass=0.5)
    med, p99, mx = pyoracle.rel_err(H.acc_aos(), ref)
    print(f"drop-in: {nt} tasks in {nt // len_task + 1} flushes, kernel {t * 1e3:.3f} ms, "
          f"median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4


def test_active_step_filters_levels(libs):
    host, lib = libs
    w = W.config_plain(16, softening=0)
    H = Host(host, w)
    act = (np.random.default_rng(2).random(w.npart) < 0.2).astype(np.int32) * 2
    np.ctypeslib.as_array(H.body)["act"] = act
    lib.hostmalloc_task(w.task_t.size)
    tt, ts = H.tasks()
    for j in range(w.task_t.size):
        tt[j], ts[j] = int(w.task_t[j] + H.first_leaf), int(w.task_s[j] + H.first_leaf)
    lib.task_compute_p2p_act_gpu(w.task_t.size, tt, ts, 1, 2)
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    lib.hostfree_task()
    a = H.acc_aos()
    sel = act == 2
    assert np.all(a[~sel] == 0.0)
    ref = _oracle(w, act=act.astype(np.uint8), level=2)
    med, p99, mx = pyoracle.rel_err(a[sel], ref[sel])
    assert p99 <= 1e-5 and mx <= 1e-4


def test_packs_equal_single_list(libs):
    host, lib = libs
    w = W.config_plain(16, softening=0)
    H = Host(host, w)
    # two packs splitting the targets by particle range, absolute leaf ids
    half = w.npart // 2
    tgt_first = w.leaf_ip[w.task_t]
    sel = [tgt_first < half, tgt_first >= half]
    bufs = []
    for n, m in enumerate(sel):
        t = np.ascontiguousarray(w.task_t[m] + H.first_leaf, np.int32)
        s = np.ascontiguousarray(w.task_s[m] + H.first_leaf, np.int32)
        bufs.append((t, s))
        pk = H.packs[n]
        lo = 0 if n == 0 else int(w.leaf_ip[w.task_t[m]].min())
        hi = half if n == 0 else w.npart
        if n == 0:
            hi = int((w.leaf_ip[w.task_t[m]] + w.leaf_np[w.task_t[m]]).max())
        pk.rankid, pk.ipart, pk.npart, pk.ntask = 0, lo, hi - lo, t.size
        pk.ileaf, pk.nleaf = H.first_leaf, w.nleaf
        pk.task_t = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        pk.task_s = s.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    lib.p2ppack_compute_gpu(H.packs, 2)
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    lib.hostmalloc_out(w.npart)
    lib.devMemcpyDH_updateAcc(H.packs, 2, -1)
    lib.hostfree_out()
    ref = _oracle(w)
    med, p99, mx = pyoracle.rel_err(H.acc_aos(), ref)
    assert p99 <= 1e-5 and mx <= 1e-4


def test_pack_active_step_filters_levels(libs):
    """p2ppack_compute_act_gpu (src/gpu_active.impl.h:155-321): two packs
    splitting the targets by particle range, every pack computed, then the
    results of particles whose act equals the level added into acc[] (the
    filter of updateAcc_fromHostBuff); level 0 adds nothing."""
    host, lib = libs
    lib.p2ppack_compute_act_gpu.argtypes = [ctypes.POINTER(P2PPack), ctypes.c_int, ctypes.c_int]
    lib.p2ppack_compute_act_gpu.restype = ctypes.c_void_p
    w = W.config_plain(16, softening=0)
    H = Host(host, w)
    act = (np.random.default_rng(4).random(w.npart) < 0.3).astype(np.int32) * 3
    np.ctypeslib.as_array(H.body)["act"] = act
    half = w.npart // 2
    tgt_first = w.leaf_ip[w.task_t]
    bufs = []
    for n, m in enumerate([tgt_first < half, tgt_first >= half]):
        t = np.ascontiguousarray(w.task_t[m] + H.first_leaf, np.int32)
        s = np.ascontiguousarray(w.task_s[m] + H.first_leaf, np.int32)
        bufs.append((t, s))
        lo = int(w.leaf_ip[w.task_t[m]].min())
        hi = int((w.leaf_ip[w.task_t[m]] + w.leaf_np[w.task_t[m]]).max())
        pk = H.packs[n]
        pk.rankid, pk.ipart, pk.npart, pk.ntask = 0, lo, hi - lo, t.size
        pk.ileaf, pk.nleaf = H.first_leaf, w.nleaf
        pk.task_t = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        pk.task_s = s.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    lib.p2ppack_compute_act_gpu(H.packs, 2, 3)
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    a = H.acc_aos()
    sel = act == 3
    assert np.all(a[~sel] == 0.0)
    ref = _oracle(w, act=act.astype(np.uint8), level=3)
    med, p99, mx = pyoracle.rel_err(a[sel], ref[sel])
    print(f"pack active: {int(sel.sum())} active, median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4
    before = a.copy()
    lib.p2ppack_compute_act_gpu(H.packs, 2, 0)                 # level 0: the filter adds nothing
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    assert np.array_equal(H.acc_aos(), before)


def test_walk_streams_flushes_into_the_dropin(libs):
    """The host walk hands its list to the drop-in while it runs
    (phost_walk_stream with task_compute_p2p_gpu itself as the flush call, the
    protocol of src/fmm.c:500-515, flushes of LEN_TASK tasks in the host
    program's task_t / task_s).  The drop-in computes flushes on its worker
    as they arrive; the collected step equals the FP64 oracle on the whole
    list, and get_last_kntime sums the launches' device time."""
    host, lib = libs
    w = W.config_plain(24, softening=0)
    H = Host(host, w)
    m = w.meta
    hl = W._host_lib()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    # the host program builds its tree from the particles in input order (the
    # partition's outcome depends on the order), as the workload's was built
    posi = np.empty_like(np.asarray(m["posi"], np.int32))
    posi[m["perm"]] = m["posi"]
    tag = np.empty_like(np.asarray(m["tag"], np.int32))
    tag[m["perm"]] = m["tag"]
    h = hl.phost_build_tree(posi.shape[0], P(posi), P(tag), m["int2pos"], m["maxleaf"],
                            P(np.asarray(m["bdl"], np.float32)), P(np.asarray(m["bdr"], np.float32)), 0)
    assert h
    try:
        ip = np.empty(w.nleaf, np.int32)
        npl = np.empty(w.nleaf, np.int32)
        assert hl.phost_count(h, 1) == w.nleaf
        hl.phost_get_leaves(h, P(ip), P(npl), None, None, None)
        assert np.array_equal(ip, w.leaf_ip) and np.array_equal(npl, w.leaf_np)
        hperm = np.empty(w.npart, np.int32)
        hl.phost_get_perm(h, P(hperm))
        assert np.array_equal(hperm, m["perm"])
        len_task = 3000
        lib.hostmalloc_task(len_task)
        tt, ts = H.tasks()
        fn = ctypes.cast(lib.task_compute_p2p_gpu, ctypes.c_void_p)
        print(f"case: streamed walk, {w.task_t.size} tasks in flushes of {len_task}", flush=True)
        assert hl.phost_walk_stream(h, np.float32(m["open_angle"]), np.float32(m["cutoff_radius"]), 0, len_task,
                                    H.first_leaf, ctypes.cast(tt, ctypes.c_void_p), ctypes.cast(ts, ctypes.c_void_p),
                                    fn, 1) == 0
        assert hl.phost_count(h, 3) == w.task_t.size
    finally:
        hl.phost_free(h)
    lib.hostfree_task()
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    t = lib.get_last_kntime(0)
    assert t > 0.0
    lib.hostmalloc_out(w.npart)
    lib.devMemcpyDH_updateAcc(H.packs, H.npack(), -1)
    lib.hostfree_out()
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    ref = _oracle(w)
    med, p99, mx = pyoracle.rel_err(H.acc_aos(), ref)
    print(f"streamed drop-in: kernels {t * 1e3:.3f} ms, median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4


def test_tiny_softening_falls_back_to_the_polynomial_law(libs):
    """SoftenScale far below the table law's floor (eps / rc ~ 6e-5, where its
    cubic coefficients leave the FP32 range): the drop-in takes the polynomial
    law instead of failing, and the step matches the FP64 oracle with that
    softening."""
    host, lib = libs
    w = W.config_plain(16, softening=0)
    rc = 2 * w.rs * 2.8125
    w.eps = 1e-6 * rc
    p = pp.ForceLaw(eps=w.eps, rs=w.rs, softening=pp.SOFT_REF)
    with pytest.raises(pp.P2PError):
        pp.law_eval(p, np.ones(1, np.float32))                      # the table law refuses it
    H = Host(host, w)
    lib.hostmalloc_task(w.task_t.size)
    tt, ts = H.tasks()
    for j in range(w.task_t.size):
        tt[j], ts[j] = int(w.task_t[j] + H.first_leaf), int(w.task_s[j] + H.first_leaf)
    lib.task_compute_p2p_gpu(w.task_t.size, tt, ts, 1)
    lib.hostfree_task()
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    lib.hostmalloc_out(w.npart)
    lib.devMemcpyDH_updateAcc(H.packs, H.npack(), -1)
    lib.hostfree_out()
    assert lib.p2p_hip_status() == 0, lib.p2p_hip_error()
    ref = _oracle(w)
    med, p99, mx = pyoracle.rel_err(H.acc_aos(), ref)
    print(f"eps/rc 1e-6 (polynomial law): median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4

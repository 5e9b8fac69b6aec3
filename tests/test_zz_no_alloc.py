"""pp2p_compute_noalloc (the entry point that allocates nothing: the list's
CSR, work items, sort storage and queue head live in the library's static
device scratch array) against pp2p_compute on the same inputs, on two calls
in a row, the second with a smaller list; and its refusals.  The device
arrays of the test itself come from the HIP runtime through ctypes."""
import ctypes

import numpy as np
import pytest

import photons_p2p as pp
import pyoracle
from photons_p2p import workloads

pytestmark = pytest.mark.gpu

_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        pp.load()                                   # brings in the runtime the library links
        # that same runtime (not another copy that a name lookup might find)
        with open("/proc/self/maps") as f:
            mapped = [ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln]
        _hip = ctypes.CDLL(mapped[0] if mapped else "libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """A device copy of a numpy array (test plumbing)."""

    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        h = _hip_lib()
        assert h.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 16)) == 0
        if self.host.nbytes:
            assert h.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p), self.host.nbytes, 1) == 0

    @property
    def addr(self) -> int:
        return self.ptr.value

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert _hip_lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            _hip_lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _setup(ctx, w):
    ctx.set_tuning()
    ctx.set_law(pp.ForceLaw(eps=w.eps, rs=w.rs, mass=w.mass, softening=w.softening, box=w.box))
    ctx.set_particles(w.pos, w.leaf_ip, w.leaf_np)


def _existing(ctx, tt, ts):
    ctx.set_tasks(tt, ts)
    ctx.zero_acc()
    ctx.compute(pp.EXEC_QUEUE)
    return ctx.acc()


def _noalloc(ctx, w, tt, ts):
    dt, ds = DevArray(tt.astype(np.int32)), DevArray(ts.astype(np.int32))
    dacc = DevArray(np.zeros((3, w.npart), np.float64))
    try:
        print(f"case: compute_noalloc N={w.npart} tasks={tt.size}", flush=True)
        ctx.compute_noalloc(tt.size, dt.addr, ds.addr, dacc.addr)
        ctx.sync()
        return dacc.get().T.copy()
    finally:
        for d in (dt, ds, dacc):
            d.free()


def test_noalloc_matches_compute_on_two_calls(gpu_ctx):
    w = workloads.uneven_pairs(seed=31, lattice=24, n_pool=2000, max_src=80, reach=4, total_pairs=6000)
    _setup(gpu_ctx, w)
    prm = pyoracle.params(w.eps, w.rs, mass=w.mass, softening=w.softening, box=w.box)
    half = w.task_t.size // 3
    for tt, ts in ((w.task_t, w.task_s), (w.task_t[:half], w.task_s[:half])):
        ref = _existing(gpu_ctx, tt, ts)
        got = _noalloc(gpu_ctx, w, tt, ts)
        # same CSR, items and kernels; only the order of the FP64 atomic adds differs
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-12 * scale
        ora = pyoracle.p2p_f64(prm, w.pos, w.leaf_ip, w.leaf_np, pyoracle.build_csr(tt, ts, w.nleaf))
        med, p99, mx = pyoracle.rel_err(got, ora)
        print(f"noalloc ({tt.size} tasks): vs compute max {np.abs(got - ref).max() / scale:.1e}; "
              f"vs oracle median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
        assert p99 <= 1e-5 and mx <= 1e-4


def test_noalloc_refuses_bad_leaves_and_oversized_lists(gpu_ctx):
    w = workloads.uneven_pairs(seed=32, lattice=16, n_pool=400, max_src=40, reach=3, total_pairs=1500)
    _setup(gpu_ctx, w)
    tt = w.task_t.astype(np.int32).copy()
    tt[7] = w.nleaf                                  # outside [0, nleaf): checked on the device, refused
    dt, ds = DevArray(tt), DevArray(w.task_s.astype(np.int32))
    dacc = DevArray(np.zeros((3, w.npart), np.float64))
    try:
        with pytest.raises(pp.P2PError, match="outside"):
            gpu_ctx.compute_noalloc(tt.size, dt.addr, ds.addr, dacc.addr)
    finally:
        for d in (dt, ds, dacc):
            d.free()
    # a list whose CSR alone outgrows the scratch: refused before any launch
    n = 3_000_000
    g = np.random.default_rng(5)
    big_t = g.integers(0, w.nleaf, n).astype(np.int32)
    big_s = g.integers(0, w.nleaf, n).astype(np.int32)
    dt, ds = DevArray(big_t), DevArray(big_s)
    dacc = DevArray(np.zeros((3, w.npart), np.float64))
    try:
        with pytest.raises(pp.P2PError, match="scratch"):
            gpu_ctx.compute_noalloc(n, dt.addr, ds.addr, dacc.addr)
    finally:
        for d in (dt, ds, dacc):
            d.free()
    # the context still works afterwards
    got = _noalloc(gpu_ctx, w, w.task_t, w.task_s)
    ref = _existing(gpu_ctx, w.task_t, w.task_s)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_noalloc_takes_a_list_of_a_million_pairs(gpu_ctx):
    """1.2 M leaf pairs (random pairs of the small workload's leaves) fit the
    64 MiB scratch and give what pp2p_compute gives."""
    w = workloads.uneven_pairs(seed=33, lattice=16, n_pool=400, max_src=40, reach=3, total_pairs=1500)
    _setup(gpu_ctx, w)
    g = np.random.default_rng(9)
    n = 1_200_000
    tt = g.integers(0, w.nleaf, n).astype(np.int32)
    ts = g.integers(0, w.nleaf, n).astype(np.int32)
    ref = _existing(gpu_ctx, tt, ts)
    got = _noalloc(gpu_ctx, w, tt, ts)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()

"""Python binding of libphotons_p2p.so, the gfx950 short-range P2P library.

The binding is a thin ctypes layer over include/photons_p2p.h: numpy arrays
in, numpy arrays out.  All computation happens in the HIP library; when the
library is missing or no GPU is visible the calls raise, they never fall back
to a CPU path.

Reference: the leaf-pair P2P path of PhotoNs, task_compute_p2p_gpu
(src/gpu_entry.cu:32-209) over P2P_kernel / P2P_core
(src/gpu_kernel.impl.h:50-157, :364-437).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(ROOT_DIR, "lib")
CORE_LIB = os.path.join(LIB_DIR, "libphotons_p2p.so")
DROPIN_LIB = os.path.join(LIB_DIR, "libp2p_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libphotons_host.so")

# pp2p_status
OK, ERR_ARG, ERR_DEVICE, ERR_ALLOC, ERR_STATE, ERR_NOGPU = range(6)
SOFT_REF, SOFT_PLUMMER = 0, 1
LAW_TABLE, LAW_POLY, LAW_LINEAR = 0, 1, 2
FRAME_LEAF, FRAME_ABS32 = 0, 1
EXEC_QUEUE, EXEC_GRID, EXEC_PER_TARGET = 0, 1, 2

UCUT_REF = 2.8125   # src/gpu_kernel.impl.h:86


class P2PError(RuntimeError):
    """A pp2p_* call returned a non-zero status."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"pp2p status {status}: {msg}")
        self.status = status


class _Params(ctypes.Structure):
    _fields_ = [("softening", ctypes.c_int), ("law", ctypes.c_int), ("frame", ctypes.c_int),
                ("reserved", ctypes.c_int), ("eps", ctypes.c_double), ("rs", ctypes.c_double),
                ("ucut", ctypes.c_double), ("mass", ctypes.c_double), ("box", ctypes.c_double)]


class _Tuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int), ("split_pairs", ctypes.c_int),
                ("segs_per_octave", ctypes.c_int), ("no_cull", ctypes.c_int),
                ("chunk_model", ctypes.c_int), ("ipl_max", ctypes.c_int), ("reserved", ctypes.c_int * 2)]


# Every symbol include/photons_p2p.h declares.
CORE_SYMBOLS = (
    "pp2p_version", "pp2p_device_count", "pp2p_last_error", "pp2p_create", "pp2p_destroy",
    "pp2p_set_params", "pp2p_set_tuning", "pp2p_get_tuning", "pp2p_set_particles_f64",
    "pp2p_set_particles_i32", "pp2p_set_active", "pp2p_set_tasks", "pp2p_add_tasks",
    "pp2p_walk_tasks", "pp2p_get_tasks", "pp2p_build_tree", "pp2p_walk_tree", "pp2p_get_tree",
    "pp2p_m2l", "pp2p_l2p", "pp2p_fmm_short", "pp2p_get_expansions", "pp2p_get_m2l_tasks",
    "pp2p_pm_forces", "pp2p_pm_add_tree", "pp2p_pm_get_mesh",
    "pp2p_prepare", "pp2p_compute", "pp2p_compute_noalloc", "pp2p_zero_acc", "pp2p_run_graph", "pp2p_sync",
    "pp2p_get_acc_f64", "pp2p_get_acc_f32_soa", "pp2p_list_stats", "pp2p_last_kernel_ms",
    "pp2p_last_prepare_ms", "pp2p_count_inside", "pp2p_time_steps", "pp2p_law_eval", "pp2p_get_csr",
)

_lib = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Loads libphotons_p2p.so (raises if it has not been built).  The
    environment variable PP2P_CORE_LIB selects another build of it."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("PP2P_CORE_LIB") or CORE_LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built: run `make -C {ROOT_DIR}` or __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    sig = {
        "pp2p_version": (ctypes.c_char_p, []),
        "pp2p_device_count": (i32, []),
        "pp2p_last_error": (ctypes.c_char_p, []),
        "pp2p_create": (vp, [i32, ctypes.POINTER(i32)]),
        "pp2p_destroy": (None, [vp]),
        "pp2p_set_params": (i32, [vp, ctypes.POINTER(_Params)]),
        "pp2p_set_tuning": (i32, [vp, ctypes.POINTER(_Tuning)]),
        "pp2p_get_tuning": (i32, [vp, ctypes.POINTER(_Tuning)]),
        "pp2p_set_particles_f64": (i32, [vp, i64, vp, i32, vp, vp]),
        "pp2p_set_particles_i32": (i32, [vp, i64, vp, ctypes.c_size_t, dbl, i32, vp, vp]),
        "pp2p_set_active": (i32, [vp, vp, i32]),
        "pp2p_set_tasks": (i32, [vp, i64, vp, vp]),
        "pp2p_add_tasks": (i32, [vp, i64, vp, vp]),
        "pp2p_walk_tasks": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                  ctypes.POINTER(i64)]),
        "pp2p_get_tasks": (i32, [vp, ctypes.POINTER(i64), vp, vp]),
        "pp2p_build_tree": (i32, [vp, i64, vp, vp, dbl, i32, vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "pp2p_walk_tree": (i32, [vp, ctypes.c_float, ctypes.c_float, ctypes.POINTER(i64)]),
        "pp2p_get_tree": (i32, [vp] + [vp] * 9),
        "pp2p_m2l": (i32, [vp, i32, vp, vp, i64, vp, vp, ctypes.c_float, vp, ctypes.POINTER(dbl)]),
        "pp2p_l2p": (i32, [vp, vp, vp]),
        "pp2p_fmm_short": (i32, [vp, i64, vp, vp, ctypes.POINTER(dbl)]),
        "pp2p_get_expansions": (i32, [vp, vp, vp]),
        "pp2p_get_m2l_tasks": (i32, [vp, ctypes.POINTER(i64), vp, vp]),
        "pp2p_pm_forces": (i32, [vp, i64, vp, vp, dbl, ctypes.c_float, i32, ctypes.c_float, ctypes.c_float, vp,
                                 ctypes.POINTER(dbl)]),
        "pp2p_pm_add_tree": (i32, [vp, ctypes.c_float, i32, ctypes.c_float, ctypes.c_float, ctypes.POINTER(dbl)]),
        "pp2p_pm_get_mesh": (i32, [vp, i32, vp]),
        "pp2p_prepare": (i32, [vp]),
        "pp2p_compute": (i32, [vp, i32]),
        "pp2p_compute_noalloc": (i32, [vp, i64, vp, vp, vp]),
        "pp2p_zero_acc": (i32, [vp]),
        "pp2p_run_graph": (i32, [vp, i32, i32, i32]),
        "pp2p_sync": (i32, [vp]),
        "pp2p_get_acc_f64": (i32, [vp, vp]),
        "pp2p_get_acc_f32_soa": (i32, [vp, vp]),
        "pp2p_list_stats": (i32, [vp] + [ctypes.POINTER(i64)] * 4),
        "pp2p_last_kernel_ms": (dbl, [vp]),
        "pp2p_last_prepare_ms": (dbl, [vp]),
        "pp2p_count_inside": (i32, [vp, ctypes.POINTER(i64)]),
        "pp2p_get_csr": (i32, [vp, ctypes.POINTER(i64), vp, vp, vp, vp]),
        "pp2p_law_eval": (i32, [ctypes.POINTER(_Params), i32, i64, vp, vp, ctypes.POINTER(i32)]),
        "pp2p_time_steps": (i32, [vp, i32, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class ForceLaw:
    """Force-law parameters (pp2p_params).

    eps: softening length (SoftenScale, src/setparam.c:23); rs: split radius
    (splitRadius, src/setparam.c:29); ucut: cutoff on u = r_eff / (2 rs)
    (2.8125, src/gpu_kernel.impl.h:86); mass: particle mass (MASSPART);
    box: periodic box for minimum-image offsets (0 = open boundaries).
    """
    eps: float
    rs: float
    mass: float = 1.0
    ucut: float = UCUT_REF
    softening: int = SOFT_REF
    law: int = LAW_TABLE
    frame: int = FRAME_LEAF
    box: float = 0.0

    @property
    def rc(self) -> float:
        return 2.0 * self.rs * self.ucut


class Context:
    """One device context of the P2P library (pp2p_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        st = ctypes.c_int(0)
        self._h = self._lib.pp2p_create(device, ctypes.byref(st))
        if not self._h:
            raise P2PError(st.value, self._lib.pp2p_last_error().decode())
        self.npart = 0
        self.nleaf = 0
        self.nnode = 0

    def _chk(self, status: int):
        if status != OK:
            raise P2PError(status, self._lib.pp2p_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pp2p_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration -------------------------------------------------------
    def set_law(self, law: ForceLaw):
        p = _Params(law.softening, law.law, law.frame, 0, law.eps, law.rs, law.ucut, law.mass, law.box)
        self._chk(self._lib.pp2p_set_params(self._h, ctypes.byref(p)))

    def set_tuning(self, waves_per_cu: int = 0, split_pairs: int = 0, segs_per_octave: int = 0,
                   no_cull: int = 0, chunk_model: int = 0, ipl_max: int = 0):
        t = _Tuning(waves_per_cu, split_pairs, segs_per_octave, no_cull, chunk_model, ipl_max)
        self._chk(self._lib.pp2p_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = _Tuning()
        self._chk(self._lib.pp2p_get_tuning(self._h, ctypes.byref(t)))
        return {"waves_per_cu": t.waves_per_cu, "split_pairs": t.split_pairs,
                "segs_per_octave": t.segs_per_octave, "no_cull": t.no_cull, "chunk_model": t.chunk_model,
                "ipl_max": t.ipl_max}

    # -- data ----------------------------------------------------------------
    def set_particles(self, pos: np.ndarray, leaf_ip: np.ndarray, leaf_np: np.ndarray):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("pos must be (N, 3)")
        ip = np.ascontiguousarray(leaf_ip, dtype=np.int32)
        npl = np.ascontiguousarray(leaf_np, dtype=np.int32)
        if ip.shape != npl.shape:
            raise ValueError("leaf_ip and leaf_np differ in length")
        self._chk(self._lib.pp2p_set_particles_f64(self._h, pos.shape[0], _ptr(pos), ip.size, _ptr(ip), _ptr(npl)))
        self.npart, self.nleaf = pos.shape[0], ip.size

    def set_particles_i32(self, posi: np.ndarray, int2pos: float, leaf_ip: np.ndarray, leaf_np: np.ndarray):
        """posi: (N, 3) int32, or a structured array whose first 12 bytes are posi."""
        posi = np.ascontiguousarray(posi)
        stride = posi.strides[0]
        n = posi.shape[0]
        ip = np.ascontiguousarray(leaf_ip, dtype=np.int32)
        npl = np.ascontiguousarray(leaf_np, dtype=np.int32)
        self._chk(self._lib.pp2p_set_particles_i32(self._h, n, _ptr(posi), stride, float(int2pos), ip.size,
                                                   _ptr(ip), _ptr(npl)))
        self.npart, self.nleaf = n, ip.size

    def set_active(self, act: np.ndarray | None, level: int = 0):
        if act is None:
            self._chk(self._lib.pp2p_set_active(self._h, None, 0))
            return
        a = np.ascontiguousarray(act, dtype=np.uint8)
        if a.size != self.npart:
            raise ValueError("act must have one entry per particle")
        self._chk(self._lib.pp2p_set_active(self._h, _ptr(a), level))

    def set_tasks(self, task_t: np.ndarray, task_s: np.ndarray):
        tt = np.ascontiguousarray(task_t, dtype=np.int32)
        ts = np.ascontiguousarray(task_s, dtype=np.int32)
        if tt.shape != ts.shape:
            raise ValueError("task_t and task_s differ in length")
        self._chk(self._lib.pp2p_set_tasks(self._h, tt.size, _ptr(tt), _ptr(ts)))

    def add_tasks(self, task_t: np.ndarray, task_s: np.ndarray):
        tt = np.ascontiguousarray(task_t, dtype=np.int32)
        ts = np.ascontiguousarray(task_s, dtype=np.int32)
        if tt.shape != ts.shape:
            raise ValueError("task_t and task_s differ in length")
        self._chk(self._lib.pp2p_add_tasks(self._h, tt.size, _ptr(tt), _ptr(ts)))

    def walk_tasks(self, tree: dict, active: np.ndarray, open_angle: float, cutoff_radius: float) -> int:
        """Builds the list on the device by the dual walk of `tree` (node_son,
        node_center, node_width, leaf_center, leaf_width as build_lists
        returns them) and makes it the context's list.  Returns its length."""
        son = np.ascontiguousarray(tree["node_son"], np.int32)
        nc = np.ascontiguousarray(tree["node_center"], np.float32)
        nw = np.ascontiguousarray(tree["node_width"], np.float32)
        lc = np.ascontiguousarray(tree["leaf_center"], np.float32)
        lw = np.ascontiguousarray(tree["leaf_width"], np.float32)
        act = np.ascontiguousarray(active, np.int32)
        if lc.shape != (self.nleaf, 3) or lw.shape != lc.shape or act.shape != (self.nleaf,):
            raise ValueError("leaf arrays do not match the context's leaves")
        if son.ndim != 2 or son.shape[1] != 2 or nc.shape != (son.shape[0], 3) or nw.shape != nc.shape:
            raise ValueError("node arrays must be nnode x 2 and nnode x 3")
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_walk_tasks(self._h, son.shape[0], _ptr(son), _ptr(nc), _ptr(nw), _ptr(lc),
                                            _ptr(lw), _ptr(act), float(open_angle), float(cutoff_radius),
                                            ctypes.byref(n)))
        return n.value

    def build_tree(self, posi: np.ndarray, tag: np.ndarray, int2pos: float, maxleaf: int, bdl: np.ndarray,
                   bdr: np.ndarray) -> tuple[int, int]:
        """Builds the kd-tree of n x 3 int32 grid positions on the device and
        makes the particles (in tree order) and leaves the context's.  Returns
        (nleaf, nnode)."""
        p = np.ascontiguousarray(posi, np.int32)
        t = np.ascontiguousarray(tag, np.int32)
        lo = np.ascontiguousarray(bdl, np.float32)
        hi = np.ascontiguousarray(bdr, np.float32)
        if p.ndim != 2 or p.shape[1] != 3 or t.shape != (p.shape[0],) or lo.shape != (3,) or hi.shape != (3,):
            raise ValueError("posi must be n x 3, tag n, bdl and bdr 3")
        nl, nn = ctypes.c_int(0), ctypes.c_int(0)
        self._chk(self._lib.pp2p_build_tree(self._h, p.shape[0], _ptr(p), _ptr(t), float(int2pos), int(maxleaf),
                                            _ptr(lo), _ptr(hi), ctypes.byref(nl), ctypes.byref(nn)))
        self.npart, self.nleaf, self.nnode = p.shape[0], nl.value, nn.value
        return nl.value, nn.value

    def get_tree(self) -> dict:
        """The tree of the last build_tree, in the layout build_lists returns."""
        nl, nn = self.nleaf, self.nnode
        out = {"perm": np.empty(self.npart, np.int32), "leaf_ip": np.empty(nl, np.int32),
               "leaf_np": np.empty(nl, np.int32), "active": np.empty(nl, np.int32),
               "leaf_center": np.empty((nl, 3), np.float32), "leaf_width": np.empty((nl, 3), np.float32),
               "node_son": np.empty((nn, 2), np.int32), "node_center": np.empty((nn, 3), np.float32),
               "node_width": np.empty((nn, 3), np.float32)}
        self._chk(self._lib.pp2p_get_tree(self._h, *[_ptr(out[k]) for k in (
            "perm", "leaf_ip", "leaf_np", "active", "leaf_center", "leaf_width", "node_son", "node_center",
            "node_width")]))
        return out

    def walk_tree(self, open_angle: float, cutoff_radius: float) -> int:
        """The device walk over the tree of the last build_tree; returns the
        list length."""
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_walk_tree(self._h, float(open_angle), float(cutoff_radius), ctypes.byref(n)))
        return n.value

    def m2l(self, center: np.ndarray, M: np.ndarray, task_t: np.ndarray, task_s: np.ndarray, rs: float,
            L: np.ndarray | None = None) -> tuple[np.ndarray, float]:
        """The M2L task loop on the device: returns (L + translations, kernel
        ms).  center ncell x 3, M and L ncell x 20 (FP32)."""
        c = np.ascontiguousarray(center, np.float32)
        m = np.ascontiguousarray(M, np.float32)
        out = np.zeros_like(m) if L is None else np.array(L, np.float32, copy=True)
        tt = np.ascontiguousarray(task_t, np.int32)
        ts = np.ascontiguousarray(task_s, np.int32)
        if c.ndim != 2 or c.shape[1] != 3 or m.shape != (c.shape[0], 20) or out.shape != m.shape or tt.shape != ts.shape:
            raise ValueError("center must be n x 3, M and L n x 20, task arrays of one length")
        ms = ctypes.c_double(0.0)
        self._chk(self._lib.pp2p_m2l(self._h, c.shape[0], _ptr(c), _ptr(m), tt.size, _ptr(tt), _ptr(ts), float(rs),
                                     _ptr(out), ctypes.byref(ms)))
        return out, ms.value

    def l2p(self, L: np.ndarray, center: np.ndarray):
        """Adds the force of each leaf's local expansion at its particles to
        the accumulator."""
        l = np.ascontiguousarray(L, np.float32)
        c = np.ascontiguousarray(center, np.float32)
        if l.shape != (self.nleaf, 20) or c.shape != (self.nleaf, 3):
            raise ValueError("L must be nleaf x 20 and center nleaf x 3")
        self._chk(self._lib.pp2p_l2p(self._h, _ptr(l), _ptr(c)))

    def fmm_short(self, m2l_t: np.ndarray | None = None, m2l_s: np.ndarray | None = None) -> float:
        """P2M, M2M, M2L over the given cell pairs (None: the pairs of the last
        walk_tree), L2L and L2P on the device tree; the forces are added to
        the accumulator.  Returns device ms."""
        ms = ctypes.c_double(0.0)
        if m2l_t is None:
            self._chk(self._lib.pp2p_fmm_short(self._h, -1, None, None, ctypes.byref(ms)))
            return ms.value
        tt = np.ascontiguousarray(m2l_t, np.int32)
        ts = np.ascontiguousarray(m2l_s, np.int32)
        if tt.shape != ts.shape:
            raise ValueError("m2l_t and m2l_s differ in length")
        self._chk(self._lib.pp2p_fmm_short(self._h, tt.size, _ptr(tt), _ptr(ts), ctypes.byref(ms)))
        return ms.value

    def pm_forces(self, posi: np.ndarray, tag: np.ndarray | None, int2pos: float, box: float, nside: int,
                  split_radius: float, mass: float) -> tuple[np.ndarray, float]:
        """Long-range PM accelerations (n, 3) FP32 of the int32 grid positions
        posi (n, 3); tag None: all local, else local iff tag == 1.  Returns
        (acc, device ms)."""
        p = np.ascontiguousarray(posi, np.int32).reshape(-1, 3)
        n = p.shape[0]
        t = None if tag is None else np.ascontiguousarray(tag, np.int32)
        if t is not None and t.size != n:
            raise ValueError("tag and posi differ in length")
        acc = np.empty((3, n), np.float32)
        ms = ctypes.c_double(0.0)
        self._chk(self._lib.pp2p_pm_forces(self._h, n, _ptr(p), None if t is None else _ptr(t), float(int2pos),
                                           float(box), int(nside), float(split_radius), float(mass), _ptr(acc),
                                           ctypes.byref(ms)))
        self._pm_nside = int(nside)
        return acc.T, ms.value                    # (n, 3) view of the SoA result

    def pm_add_tree(self, box: float, nside: int, split_radius: float, mass: float) -> float:
        """Add the PM accelerations of the device tree's local particles to the
        accumulator; returns device ms."""
        ms = ctypes.c_double(0.0)
        self._chk(self._lib.pp2p_pm_add_tree(self._h, float(box), int(nside), float(split_radius), float(mass),
                                             ctypes.byref(ms)))
        self._pm_nside = int(nside)
        return ms.value

    def pm_mesh(self, which: int) -> np.ndarray:
        """The last pm_forces call's mass (0) or potential (1) mesh."""
        nside = getattr(self, "_pm_nside", 0)
        if nside <= 0:
            raise P2PError("pm_forces before pm_mesh")
        out = np.empty((nside, nside, nside), np.float64)
        self._chk(self._lib.pp2p_pm_get_mesh(self._h, int(which), _ptr(out)))
        return out

    def get_m2l_tasks(self) -> tuple[np.ndarray, np.ndarray]:
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_get_m2l_tasks(self._h, ctypes.byref(n), None, None))
        tt = np.empty(n.value, np.int32)
        ts = np.empty(n.value, np.int32)
        self._chk(self._lib.pp2p_get_m2l_tasks(self._h, ctypes.byref(n), _ptr(tt), _ptr(ts)))
        return tt, ts

    def get_expansions(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.nleaf + self.nnode
        M = np.empty((n, 20), np.float32)
        L = np.empty((n, 20), np.float32)
        self._chk(self._lib.pp2p_get_expansions(self._h, _ptr(M), _ptr(L)))
        return M, L

    def get_tasks(self) -> tuple[np.ndarray, np.ndarray]:
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_get_tasks(self._h, ctypes.byref(n), None, None))
        tt = np.empty(n.value, np.int32)
        ts = np.empty(n.value, np.int32)
        self._chk(self._lib.pp2p_get_tasks(self._h, ctypes.byref(n), _ptr(tt), _ptr(ts)))
        return tt, ts

    # -- execution -----------------------------------------------------------
    def prepare(self) -> float:
        self._chk(self._lib.pp2p_prepare(self._h))
        return self._lib.pp2p_last_prepare_ms(self._h)

    def compute(self, exec_form: int = EXEC_QUEUE):
        self._chk(self._lib.pp2p_compute(self._h, exec_form))

    def compute_noalloc(self, ntask: int, task_t_dev: int, task_s_dev: int, acc_dev: int):
        """pp2p_compute_noalloc: the list as device arrays (addresses of ntask
        int32 each), accumulating into a device array of 3 x npart doubles;
        the library allocates nothing for the call."""
        self._chk(self._lib.pp2p_compute_noalloc(self._h, ntask, ctypes.c_void_p(task_t_dev),
                                                 ctypes.c_void_p(task_s_dev), ctypes.c_void_p(acc_dev)))

    def zero_acc(self):
        self._chk(self._lib.pp2p_zero_acc(self._h))

    def run_graph(self, exec_form: int = EXEC_QUEUE, nsteps: int = 1, nreplay: int = 1):
        self._chk(self._lib.pp2p_run_graph(self._h, exec_form, nsteps, nreplay))

    def sync(self):
        self._chk(self._lib.pp2p_sync(self._h))

    def acc(self) -> np.ndarray:
        out = np.empty((self.npart, 3), dtype=np.float64)
        self._chk(self._lib.pp2p_get_acc_f64(self._h, _ptr(out)))
        return out

    def acc_f32_soa(self) -> np.ndarray:
        out = np.empty((3, self.npart), dtype=np.float32)
        self._chk(self._lib.pp2p_get_acc_f32_soa(self._h, _ptr(out)))
        return out

    def list_stats(self) -> dict:
        v = [ctypes.c_longlong(0) for _ in range(4)]
        self._chk(self._lib.pp2p_list_stats(self._h, *[ctypes.byref(x) for x in v]))
        return {"ntask": v[0].value, "nrows": v[1].value, "nitems": v[2].value, "npairs": v[3].value}

    def csr(self) -> dict:
        """The device's target CSR of the current list: key_sorted, src
        (ntask each), row_tgt (nrows), row_off (nrows + 1)."""
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_get_csr(self._h, ctypes.byref(n), None, None, None, None))
        tt, ts = self.get_tasks()
        out = {"key_sorted": np.empty(tt.size, np.int32), "src": np.empty(tt.size, np.int32),
               "row_tgt": np.empty(n.value, np.int32), "row_off": np.empty(n.value + 1, np.int32)}
        self._chk(self._lib.pp2p_get_csr(self._h, ctypes.byref(n), _ptr(out["key_sorted"]), _ptr(out["src"]),
                                         _ptr(out["row_tgt"]), _ptr(out["row_off"])))
        return out

    def count_inside(self) -> int:
        n = ctypes.c_longlong(0)
        self._chk(self._lib.pp2p_count_inside(self._h, ctypes.byref(n)))
        return n.value

    def time_steps(self, exec_form: int = EXEC_QUEUE, nsteps: int = 1) -> tuple[float, float]:
        """(total, force-only) device milliseconds of nsteps x (zero + compute)."""
        ms = ctypes.c_double(0.0)
        fms = ctypes.c_double(0.0)
        self._chk(self._lib.pp2p_time_steps(self._h, exec_form, nsteps, ctypes.byref(ms), ctypes.byref(fms)))
        return ms.value, fms.value

    def last_kernel_ms(self) -> float:
        return self._lib.pp2p_last_kernel_ms(self._h)


def law_eval(law: ForceLaw, s: np.ndarray, segs_per_octave: int = 32) -> tuple[np.ndarray, int]:
    """Host evaluation of the device table law at s = r_eff^2 / rc^2 (kernel
    units, rc = 1); returns (F, table entries).  No device is used."""
    lib = load()
    s = np.ascontiguousarray(s, np.float32)
    out = np.empty_like(s)
    n = ctypes.c_int(0)
    p = _Params(law.softening, law.law, law.frame, 0, law.eps, law.rs, law.ucut, law.mass, law.box)
    st = lib.pp2p_law_eval(ctypes.byref(p), segs_per_octave, s.size, _ptr(s), _ptr(out), ctypes.byref(n))
    if st != OK:
        raise P2PError(st, lib.pp2p_last_error().decode())
    return out, n.value


def device_count() -> int:
    return load().pp2p_device_count()

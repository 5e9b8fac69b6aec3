"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, as the checker.  The product library never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

SOFT_REF, SOFT_PLUMMER = 0, 1


class Params(ctypes.Structure):
    _fields_ = [("softening", ctypes.c_int), ("pad_", ctypes.c_int), ("eps", ctypes.c_double),
                ("rs", ctypes.c_double), ("ucut", ctypes.c_double), ("mass", ctypes.c_double),
                ("box", ctypes.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    lib.oracle_g.restype = dbl
    lib.oracle_g.argtypes = [dbl]
    lib.oracle_ref_tables.restype = None
    lib.oracle_ref_tables.argtypes = [vp, vp]
    lib.oracle_pair_ref32.restype = None
    lib.oracle_pair_ref32.argtypes = [i32, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp, vp, vp, vp]
    common = [ctypes.POINTER(Params), vp, vp, vp, i32, vp, vp, vp, vp, i32]
    lib.oracle_p2p_f64.restype = None
    lib.oracle_p2p_f64.argtypes = common + [vp, i32]
    lib.oracle_p2p_ref32.restype = None
    lib.oracle_p2p_ref32.argtypes = [ctypes.POINTER(Params), vp, i32, vp, vp, i32, vp, vp, vp, vp, i32, vp, i32]
    lib.oracle_build_csr.restype = i32
    lib.oracle_build_csr.argtypes = [i64, vp, vp, i32, vp, vp, vp]
    lib.oracle_count_pairs.restype = None
    lib.oracle_count_pairs.argtypes = common + [ctypes.POINTER(i64), ctypes.POINTER(i64), i32]
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def params(eps, rs, mass=1.0, ucut=2.8125, softening=SOFT_REF, box=0.0) -> Params:
    return Params(softening, 0, eps, rs, ucut, mass, box)


def g(u: float) -> float:
    return load().oracle_g(u)


def ref_tables():
    t0 = np.empty(512, np.float32)
    t1 = np.empty(512, np.float32)
    load().oracle_ref_tables(_p(t0), _p(t1))
    return t0, t1


def build_csr(task_t, task_s, nleaf):
    """Stable target CSR: (tgt_leaf[ntgt], tgt_off[ntgt+1] int64, src_leaf[ntask])."""
    tt = np.ascontiguousarray(task_t, np.int32)
    ts = np.ascontiguousarray(task_s, np.int32)
    tgt = np.empty(max(nleaf, 1), np.int32)
    off = np.empty(nleaf + 1, np.int64)
    src = np.empty(max(tt.size, 1), np.int32)
    n = load().oracle_build_csr(tt.size, _p(tt), _p(ts), nleaf, _p(tgt), _p(off), _p(src))
    if n < 0:
        raise ValueError("task id outside [0, nleaf)")
    return tgt[:n].copy(), off[:n + 1].copy(), src[:tt.size].copy()


def p2p_f64(prm: Params, pos, leaf_ip, leaf_np, csr, act=None, level=0, nthreads=0):
    """FP64 direct sum over the CSR; returns (N, 3) accelerations."""
    pos = np.ascontiguousarray(pos, np.float64)
    ip = np.ascontiguousarray(leaf_ip, np.int32)
    npl = np.ascontiguousarray(leaf_np, np.int32)
    tgt, off, src = csr
    a = None if act is None else np.ascontiguousarray(act, np.uint8)
    acc = np.zeros_like(pos)
    load().oracle_p2p_f64(ctypes.byref(prm), _p(pos), _p(ip), _p(npl), tgt.size, _p(tgt), _p(off), _p(src),
                          _p(a), level, _p(acc), nthreads)
    return acc


def p2p_ref32(prm: Params, pos_soa_f32, leaf_ip, leaf_np, csr, act=None, level=0, nthreads=0):
    """FP32 restatement of the reference kernel; pos as (3, N) float32; returns (3, N) float32."""
    ps = np.ascontiguousarray(pos_soa_f32, np.float32)
    n = ps.shape[1]
    ip = np.ascontiguousarray(leaf_ip, np.int32)
    npl = np.ascontiguousarray(leaf_np, np.int32)
    tgt, off, src = csr
    a = None if act is None else np.ascontiguousarray(act, np.uint8)
    acc = np.zeros((3, n), np.float32)
    load().oracle_p2p_ref32(ctypes.byref(prm), _p(ps), n, _p(ip), _p(npl), tgt.size, _p(tgt), _p(off), _p(src),
                            _p(a), level, _p(acc), nthreads)
    return acc


def count_pairs(prm: Params, pos, leaf_ip, leaf_np, csr, act=None, level=0, nthreads=0):
    """(N_list, N_in): listed pairs and those with u_eff <= ucut."""
    pos = np.ascontiguousarray(pos, np.float64)
    ip = np.ascontiguousarray(leaf_ip, np.int32)
    npl = np.ascontiguousarray(leaf_np, np.int32)
    tgt, off, src = csr
    a = None if act is None else np.ascontiguousarray(act, np.uint8)
    nl = ctypes.c_longlong(0)
    ni = ctypes.c_longlong(0)
    load().oracle_count_pairs(ctypes.byref(prm), _p(pos), _p(ip), _p(npl), tgt.size, _p(tgt), _p(off), _p(src),
                              _p(a), level, ctypes.byref(nl), ctypes.byref(ni), nthreads)
    return nl.value, ni.value


def rel_err(a, ref):
    """Per-particle |a - ref| / |ref| over particles with |ref| > 0: (median, p99, max)."""
    num = np.linalg.norm(np.asarray(a, np.float64) - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    m = den > 0
    e = num[m] / den[m]
    if e.size == 0:
        return 0.0, 0.0, 0.0
    return float(np.median(e)), float(np.percentile(e, 99)), float(e.max())

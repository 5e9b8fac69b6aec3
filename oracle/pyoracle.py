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
    flt = ctypes.c_float
    for name, args in (("oracle_m2l_ref32", [flt, flt, flt, vp, flt, vp]),
                       ("oracle_l2p_ref32", [i32, vp, dbl, vp, vp, vp, vp, vp]),
                       ("oracle_radial_f64", [dbl, dbl, vp]),
                       ("oracle_m2l_f64", [dbl, dbl, dbl, vp, dbl, vp]),
                       ("oracle_l2p_f64", [dbl, dbl, dbl, vp, vp]),
                       ("oracle_p2m_f64", [i32, vp, dbl, vp, vp]),
                       ("oracle_m2l_tasks_f64", [i32, vp, vp, i64, vp, vp, dbl, vp]),
                       ("oracle_m2l_tasks_ref32", [i32, vp, vp, i64, vp, vp, flt, vp]),
                       ("oracle_p2m_ref32", [i32, vp, dbl, vp, flt, vp]),
                       ("oracle_m2m_ref32", [flt, flt, flt, vp, vp]),
                       ("oracle_l2l_ref32", [flt, flt, flt, vp, vp]),
                       ("oracle_m2m_f64", [dbl, dbl, dbl, vp, vp]),
                       ("oracle_l2l_f64", [dbl, dbl, dbl, vp, vp])):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = args
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


# --------------------------------------------------------------------------
# Reference list producer (oracle/_ref/libreflist.so, built from the
# reference's own src/fmm.c by oracle/ref.mk in the build container)

REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libreflist.so")


def reflist_available() -> bool:
    return os.path.exists(REF_LIB)


def reflist_build(posi, tag, int2pos, maxleaf, open_angle, cutoff_radius, bdl, bdr) -> dict:
    """Runs the reference's build_localtree + walk_task_p2p on the given
    particles; returns perm, leaf_ip, leaf_np, task_t, task_s (0-based)."""
    lib = ctypes.CDLL(REF_LIB)
    vp = ctypes.c_void_p
    ip_ = ctypes.POINTER(ctypes.c_int)
    lib.reflist_build.restype = ctypes.c_int
    lib.reflist_build.argtypes = [ctypes.c_longlong, vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_float, vp, vp, ctypes.POINTER(ip_), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ip_), ctypes.POINTER(ip_), ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ip_), ctypes.POINTER(ip_)]
    lib.reflist_free.argtypes = [vp]
    posi = np.ascontiguousarray(posi, np.int32)
    tag = np.ascontiguousarray(tag, np.int32)
    bdl = np.ascontiguousarray(bdl, np.float32)
    bdr = np.ascontiguousarray(bdr, np.float32)
    perm, lip, lnp, tt, ts = ip_(), ip_(), ip_(), ip_(), ip_()
    nleaf = ctypes.c_int(0)
    ntask = ctypes.c_longlong(0)
    rc = lib.reflist_build(posi.shape[0], _p(posi), _p(tag), int2pos, maxleaf, open_angle, cutoff_radius,
                           _p(bdl), _p(bdr), ctypes.byref(perm), ctypes.byref(nleaf), ctypes.byref(lip),
                           ctypes.byref(lnp), ctypes.byref(ntask), ctypes.byref(tt), ctypes.byref(ts))
    if rc != 0:
        raise RuntimeError("reflist_build failed")
    n, nl, nt = posi.shape[0], nleaf.value, ntask.value
    out = {"perm": np.ctypeslib.as_array(perm, (n,)).copy(),
           "leaf_ip": np.ctypeslib.as_array(lip, (nl,)).copy(),
           "leaf_np": np.ctypeslib.as_array(lnp, (nl,)).copy(),
           "task_t": np.ctypeslib.as_array(tt, (nt,)).copy() if nt else np.zeros(0, np.int32),
           "task_s": np.ctypeslib.as_array(ts, (nt,)).copy() if nt else np.zeros(0, np.int32)}
    for p in (perm, lip, lnp, tt, ts):
        lib.reflist_free(ctypes.cast(p, vp))
    return out


# ---------------------------------------------------------------- multipoles

NMULTI = 20


def m2l_ref32(d, M, rs, L0):
    """The reference's FP32 M2L (src/operator.c:331-474) for each row of
    d (n x 3), M (n x 20), rs (n), added to L0 (n x 20)."""
    lib = load()
    d = np.asarray(d, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    L = np.array(L0, np.float32, copy=True)
    rs = np.broadcast_to(np.asarray(rs, np.float32), (d.shape[0],))
    for k in range(d.shape[0]):
        lib.oracle_m2l_ref32(float(d[k, 0]), float(d[k, 1]), float(d[k, 2]), _p(M[k]), float(rs[k]), _p(L[k]))
    return L


def l2p_ref32(posi, int2pos, center, L, acc0):
    """The reference's FP32 L2P (src/operator.c:203-262) of one leaf: acc0
    (3 x n) plus the forces at the n particles posi (n x 3 int32)."""
    lib = load()
    posi = np.ascontiguousarray(posi, np.int32)
    acc = [np.array(acc0[d], np.float32, copy=True) for d in range(3)]
    lib.oracle_l2p_ref32(posi.shape[0], _p(posi), float(int2pos), _p(np.ascontiguousarray(center, np.float32)),
                         _p(np.ascontiguousarray(L, np.float32)), _p(acc[0]), _p(acc[1]), _p(acc[2]))
    return np.stack(acc)


def radial_f64(r, rs):
    f = np.zeros(4)
    load().oracle_radial_f64(float(r), float(rs), _p(f))
    return f


def m2l_f64(d, M, rs, L0=None):
    """FP64 M2L of one separation d (3) and multipole M (20), added to L0."""
    L = np.zeros(NMULTI) if L0 is None else np.array(L0, np.float64, copy=True)
    load().oracle_m2l_f64(float(d[0]), float(d[1]), float(d[2]), _p(np.ascontiguousarray(M, np.float64)),
                          float(rs), _p(L))
    return L


def l2p_f64(d, L):
    F = np.zeros(3)
    load().oracle_l2p_f64(float(d[0]), float(d[1]), float(d[2]), _p(np.ascontiguousarray(L, np.float64)), _p(F))
    return F


def p2m_f64(pos, mass, center):
    pos = np.ascontiguousarray(pos, np.float64)
    M = np.zeros(NMULTI)
    load().oracle_p2m_f64(pos.shape[0], _p(pos), float(mass), _p(np.ascontiguousarray(center, np.float64)), _p(M))
    return M


def m2l_tasks_f64(center, M, task_t, task_s, rs, L0=None):
    """The M2L task loop (src/fmm.c:1188-1217) in FP64 over cells with
    centres center (n x 3) and multipoles M (n x 20)."""
    center = np.ascontiguousarray(center, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    L = np.zeros_like(M) if L0 is None else np.array(L0, np.float64, copy=True)
    tt = np.ascontiguousarray(task_t, np.int32)
    ts = np.ascontiguousarray(task_s, np.int32)
    load().oracle_m2l_tasks_f64(center.shape[0], _p(center), _p(M), tt.size, _p(tt), _p(ts), float(rs), _p(L))
    return L


def m2l_tasks_ref32(center, M, task_t, task_s, rs, L0=None):
    """The reference's task loop in its FP32 arithmetic (the CPU baseline)."""
    center = np.ascontiguousarray(center, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    L = np.zeros_like(M) if L0 is None else np.array(L0, np.float32, copy=True)
    tt = np.ascontiguousarray(task_t, np.int32)
    ts = np.ascontiguousarray(task_s, np.int32)
    load().oracle_m2l_tasks_ref32(center.shape[0], _p(center), _p(M), tt.size, _p(tt), _p(ts), float(rs), _p(L))
    return L


def p2m_ref32(posi, int2pos, center, mass):
    posi = np.ascontiguousarray(posi, np.int32)
    M = np.zeros(NMULTI, np.float32)
    load().oracle_p2m_ref32(posi.shape[0], _p(posi), float(int2pos), _p(np.ascontiguousarray(center, np.float32)),
                            float(mass), _p(M))
    return M


def _shift(fn, dtype, s, A, T0):
    T = np.array(T0, dtype, copy=True)
    getattr(load(), fn)(float(s[0]), float(s[1]), float(s[2]), _p(np.ascontiguousarray(A, dtype)), _p(T))
    return T


def m2m_ref32(s, M, T0):
    return _shift("oracle_m2m_ref32", np.float32, s, M, T0)


def l2l_ref32(s, L, T0):
    return _shift("oracle_l2l_ref32", np.float32, s, L, T0)


def m2m_f64(s, M, T0=None):
    return _shift("oracle_m2m_f64", np.float64, s, M, np.zeros(NMULTI) if T0 is None else T0)


def l2l_f64(s, L, T0=None):
    return _shift("oracle_l2l_f64", np.float64, s, L, np.zeros(NMULTI) if T0 is None else T0)

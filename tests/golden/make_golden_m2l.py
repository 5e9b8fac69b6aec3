"""Writes tests/golden/m2l_ref.npz: inputs and outputs of the reference's own
short-range multipole operators, called from oracle/_ref/libreflist.so (the
reference's src/operator.c compiled where it lies, flags INTXYZ + LONGSHORT,
QUADRUPOLE + OCTUPOLE as inc/typedef.h:12-13 set them):

  m2l  src/operator.c:331-474: toL += M2L(dx, dy, dz, M), FP32 cpuType, with
       the erfc-modified kernel of the long/short split
  l2p  src/operator.c:203-262: acc[d][n] += force of the leaf's local
       expansion L at each particle of the leaf
  m2m  src/operator.c:100-165: tM += M shifted by (dx, dy, dz)
  l2l  src/operator.c:477-580: toL += L shifted by (dx, dy, dz)
  p2m  src/operator.c:14-98: M of a leaf's particles about a centre

Run in the build container (where oracle/_ref exists):
    python tests/golden/make_golden_m2l.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libreflist.so")
NMULTI = 20


class Body(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int), ("act", ctypes.c_int), ("posi", ctypes.c_int * 3),
                ("vel", ctypes.c_float * 3), ("id", ctypes.c_ulong)]


class Pack(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("npart", ctypes.c_int), ("ipart", ctypes.c_int),
                ("width", ctypes.c_float * 3), ("center", ctypes.c_float * 3), ("M", ctypes.c_float * NMULTI),
                ("L", ctypes.c_float * NMULTI), ("acc_pm", ctypes.c_float * 3)]


def main():
    lib = ctypes.CDLL(REF_LIB)
    vp = ctypes.c_void_p
    lib.m2l.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp]
    lib.m2l.restype = None
    lib.l2p.argtypes = [ctypes.c_int]
    lib.l2p.restype = None
    g = np.random.default_rng(20261210)
    out = {}

    # m2l: random separations (a few with r^2 < 1), multipoles of realistic
    # shape (M[0] > 0, moments shrinking with order), local expansions to add to
    rs_list = np.array([1.0, 1.953125, 0.7], np.float32)
    n = 400
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rsv = rs_list[g.integers(0, rs_list.size, n)]
    # separations up to the cutoff radius 5.625 rs, a few with r^2 < 1
    r = np.concatenate([np.maximum(rsv[:n - 20] * g.uniform(0.3, 5.625, n - 20), 1.0 + 1e-3),
                        g.uniform(0.2, 0.99, 20)])
    dxyz = (d * r[:, None]).astype(np.float32)
    scale = np.array([1.0] + [0.5] * 3 + [0.25] * 6 + [0.125] * 10, np.float32)
    M = (g.normal(size=(n, NMULTI)) * scale).astype(np.float32)
    M[:, 0] = np.abs(M[:, 0]) + 0.5
    L0 = (g.normal(size=(n, NMULTI)) * 0.01).astype(np.float32)
    L = L0.copy()
    sr = ctypes.c_float.in_dll(lib, "splitRadius")
    for k in range(n):
        sr.value = float(rsv[k])
        Mk = np.ascontiguousarray(M[k])
        Lk = np.ascontiguousarray(L[k])
        lib.m2l(float(dxyz[k, 0]), float(dxyz[k, 1]), float(dxyz[k, 2]), Mk.ctypes.data_as(vp),
                Lk.ctypes.data_as(vp))
        L[k] = Lk
    out.update(m2l_d=dxyz, m2l_M=M, m2l_L0=L0, m2l_rs=rsv, m2l_L=L)

    # l2p: leaves of 1..40 particles on the integer grid, random expansions
    int2pos = 1.0 / 4096.0
    nleaf = 60
    sizes = g.integers(1, 41, nleaf)
    ip = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    npart = int(sizes.sum())
    centers = g.uniform(10.0, 90.0, (nleaf, 3)).astype(np.float32)
    posi = np.empty((npart, 3), np.int32)
    for l in range(nleaf):
        p = centers[l] + g.uniform(-1.5, 1.5, (sizes[l], 3))
        posi[ip[l]:ip[l] + sizes[l]] = np.round(p / int2pos).astype(np.int32)
    Lleaf = (g.normal(size=(nleaf, NMULTI)) * scale).astype(np.float32)
    bodies = (Body * npart)()
    for i in range(npart):
        bodies[i].posi[:] = [int(v) for v in posi[i]]
    packs = (Pack * nleaf)()
    for l in range(nleaf):
        packs[l].npart = int(sizes[l])
        packs[l].ipart = int(ip[l])
        packs[l].center[:] = [float(v) for v in centers[l]]
        packs[l].L[:] = [float(v) for v in Lleaf[l]]
    acc0 = (g.normal(size=(3, npart)) * 0.01).astype(np.float32)
    accs = [acc0[d].copy() for d in range(3)]
    ctypes.c_void_p.in_dll(lib, "part").value = ctypes.addressof(bodies)
    ctypes.c_void_p.in_dll(lib, "leaf").value = ctypes.addressof(packs)
    ctypes.c_double.in_dll(lib, "INT2POS").value = int2pos
    acc_ptrs = (ctypes.c_void_p * 3).in_dll(lib, "acc")
    for d in range(3):
        acc_ptrs[d] = accs[d].ctypes.data
    for l in range(nleaf):
        lib.l2p(l)
    out.update(l2p_posi=posi, l2p_ip=ip, l2p_np=sizes.astype(np.int32), l2p_center=centers, l2p_L=Lleaf,
               l2p_int2pos=np.float64(int2pos), l2p_acc0=acc0, l2p_acc=np.stack(accs))
    # m2m and l2l: random shifts of random expansions, added to random targets
    lib.m2m.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp]
    lib.m2m.restype = None
    lib.l2l.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp]
    lib.l2l.restype = None
    k = 200
    sh = g.uniform(-3.0, 3.0, (k, 3)).astype(np.float32)
    A = (g.normal(size=(k, NMULTI)) * scale).astype(np.float32)
    T0 = (g.normal(size=(k, NMULTI)) * 0.1).astype(np.float32)
    Tm, Tl = T0.copy(), T0.copy()
    for i in range(k):
        Ai = np.ascontiguousarray(A[i])
        tm, tl = np.ascontiguousarray(Tm[i]), np.ascontiguousarray(Tl[i])
        lib.m2m(float(sh[i, 0]), float(sh[i, 1]), float(sh[i, 2]), Ai.ctypes.data_as(vp), tm.ctypes.data_as(vp))
        lib.l2l(float(sh[i, 0]), float(sh[i, 1]), float(sh[i, 2]), Ai.ctypes.data_as(vp), tl.ctypes.data_as(vp))
        Tm[i], Tl[i] = tm, tl
    out.update(shift_d=sh, shift_A=A, shift_T0=T0, m2m_T=Tm, l2l_T=Tl)

    # p2m: the l2p leaves' particles about their centres, mass MASSPART
    lib.p2m.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
    lib.p2m.restype = None
    mass = 0.37
    ctypes.c_float.in_dll(lib, "MASSPART").value = mass
    Mp = np.zeros((nleaf, NMULTI), np.float32)
    for l in range(nleaf):
        c = np.ascontiguousarray(centers[l])
        ml = np.zeros(NMULTI, np.float32)
        lib.p2m(int(ip[l]), int(sizes[l]), c.ctypes.data_as(vp), ml.ctypes.data_as(vp))
        Mp[l] = ml
    out.update(p2m_M=Mp, p2m_mass=np.float32(mass))
    np.savez_compressed(os.path.join(HERE, "m2l_ref.npz"), **out)
    print(f"m2l: {n} cases, l2p: {nleaf} leaves / {npart} particles")


if __name__ == "__main__":
    main()

"""Seeded synthetic workloads for the P2P path.

Every generator draws from numpy.random.Generator(PCG64(seed)) and returns
FP64 master coordinates; particles are grouped by leaf (leaf l owns the
contiguous range [leaf_ip[l], leaf_ip[l] + leaf_np[l])).

Force-law scale conventions follow the reference's setup_param
(src/setparam.c:16-30): split radius rs = 1.25 mesh cells, tree cutoff
cutoffRadius = 4.5 rs, kernel cutoff u = r / (2 rs) <= 2.8125.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

UCUT = 2.8125


@dataclass
class Workload:
    name: str
    pos: np.ndarray            # (N, 3) float64
    leaf_ip: np.ndarray        # (L,) int32
    leaf_np: np.ndarray        # (L,) int32
    task_t: np.ndarray         # (T,) int32 target leaf of each leaf pair
    task_s: np.ndarray         # (T,) int32 source leaf
    box: float                 # periodic box length (0: open)
    rs: float                  # split radius
    eps: float                 # softening length
    softening: int = 0         # 0 REF max(r, eps), 1 PLUMMER
    mass: float = 1.0
    meta: dict = field(default_factory=dict)

    @property
    def npart(self) -> int:
        return self.pos.shape[0]

    @property
    def nleaf(self) -> int:
        return self.leaf_ip.size

    def n_list(self) -> int:
        """Sum over tasks of n_t * n_s (src/gpu_entry.cu:50-61)."""
        return int(np.sum(self.leaf_np[self.task_t].astype(np.int64) * self.leaf_np[self.task_s]))


def rng_for(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uneven_pairs(seed: int = 20261205, lattice: int = 64, box: float = 100.0, n_pool: int = 65536,
                 max_occ: int = 256, max_src: int = 400, reach: int = 7, total_pairs: int = 1_000_000,
                 eps: float | None = None, softening: int = 1) -> Workload:
    """Uneven batching workload: leaf cubes on distinct lattice sites, occupancy
    log-uniform in [1, max_occ], particles uniform in each cube; targets drawn
    from the pool, each with s ~ U{1..max_src} distinct sources within `reach`
    lattice cells (periodic), until sum(s) = total_pairs leaf pairs.  Force law
    of the 64^3 plain config: mesh cell = box / lattice, rs = 1.25 cells,
    Plummer eps = box / lattice / 40."""
    g = rng_for(seed)
    cell = box / lattice
    n_pool = min(n_pool, lattice ** 3)
    sites = g.choice(lattice ** 3, size=n_pool, replace=False)
    ijk = np.stack(np.unravel_index(sites, (lattice,) * 3), axis=1)
    occ = np.clip(np.floor(max_occ ** g.random(n_pool)).astype(np.int64), 1, max_occ)
    leaf_np = occ.astype(np.int32)
    leaf_ip = np.zeros(n_pool, np.int64)
    leaf_ip[1:] = np.cumsum(occ)[:-1]
    n = int(occ.sum())
    owner = np.repeat(np.arange(n_pool), occ)
    pos = (ijk[owner] + g.random((n, 3))) * cell
    # neighbour lookup on the lattice: site -> pool index
    site_of = -np.ones(lattice ** 3, np.int64)
    site_of[sites] = np.arange(n_pool)
    offs = np.arange(-reach, reach + 1)
    tt, ts = [], []
    done = 0
    while done < total_pairs:
        t = int(g.integers(n_pool))
        s = int(g.integers(1, max_src + 1))
        s = min(s, total_pairs - done)
        c = ijk[t]
        # candidate sites within `reach` cells (periodic cube)
        cand = np.stack(np.meshgrid(offs, offs, offs, indexing="ij"), -1).reshape(-1, 3)
        nb = (c[None, :] + cand) % lattice
        idx = site_of[np.ravel_multi_index(nb.T, (lattice,) * 3)]
        idx = idx[idx >= 0]
        s = min(s, idx.size)
        pick = g.choice(idx, size=s, replace=False)
        tt.append(np.full(s, t, np.int32))
        ts.append(pick.astype(np.int32))
        done += s
    task_t = np.concatenate(tt)
    task_s = np.concatenate(ts)
    rs = 1.25 * cell
    if eps is None:
        eps = cell / 40.0 if softening == 1 else 0.03 * cell
    return Workload("uneven", pos, leaf_ip.astype(np.int32), leaf_np, task_t, task_s, box, rs, eps,
                    softening=softening, meta={"lattice": lattice, "n_pool": n_pool, "total_pairs": total_pairs})


# --------------------------------------------------------------------------
# Reference integer grid, ghost images and the list producer

BITWIDTH = 1 << 30     # inc/photoNs.h:237


def int_grid(box: float, cutoff_radius: float) -> float:
    """INT2POS of the reference: (BOXSIZE + 1.1 cutoffRadius) / 2^30 (src/setparam.c:32-35)."""
    return (box + cutoff_radius * 1.1) / BITWIDTH


def to_grid(pos: np.ndarray, box: float, int2pos: float) -> tuple[np.ndarray, int]:
    """Periodic int32 grid positions of local particles and the box length in
    grid units (the reference shifts ghost copies by an integer box,
    src/boundary.c:647-687)."""
    box_i = int(round(box / int2pos))
    posi = np.floor(pos / int2pos).astype(np.int64) % box_i
    return posi.astype(np.int32), box_i


def ghost_images(posi: np.ndarray, box_i: int, width_i: int) -> tuple[np.ndarray, np.ndarray]:
    """Periodic copies of local particles lying within width_i outside the
    box [0, box_i): what a single domain covering the whole box imports as
    boundary particles.  Returns (ghost posi, index of the local original)."""
    out, src = [], []
    p = posi.astype(np.int64)
    near_lo = p < width_i
    near_hi = p >= box_i - width_i
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if sx == sy == sz == 0:
                    continue
                m = np.ones(p.shape[0], bool)
                for d, s in enumerate((sx, sy, sz)):
                    if s == 1:
                        m &= near_lo[:, d]      # shifted up by one box
                    elif s == -1:
                        m &= near_hi[:, d]
                idx = np.nonzero(m)[0]
                if idx.size == 0:
                    continue
                g = p[idx] + np.array([sx, sy, sz], np.int64) * box_i
                keep = np.all((g > -width_i) & (g < box_i + width_i), axis=1)
                out.append(g[keep])
                src.append(idx[keep])
    if not out:
        return np.zeros((0, 3), np.int32), np.zeros(0, np.int64)
    return np.concatenate(out).astype(np.int32), np.concatenate(src)


def combine_bounds(posi_all: np.ndarray, int2pos: float, box: float) -> tuple[np.ndarray, np.ndarray]:
    """Root box of the tree as particle_combine computes it (src/gravity.c:15-140):
    FP32 min/max of float(posi * INT2POS) widened by 1e-6 BOXSIZE."""
    f = (posi_all.astype(np.float64) * int2pos).astype(np.float32)
    box32 = np.float32(box)
    lo = np.minimum(f.min(axis=0), box32)
    hi = np.maximum(f.max(axis=0), np.float32(0.0))
    widen = 0.000001 * np.float64(box32)
    bdl = (lo.astype(np.float64) - widen).astype(np.float32)
    bdr = (hi.astype(np.float64) + widen).astype(np.float32)
    return bdl, bdr


def _host_lib():
    import ctypes
    import os
    from . import HOST_LIB
    if not os.path.exists(HOST_LIB):
        raise FileNotFoundError(f"{HOST_LIB} not built")
    lib = ctypes.CDLL(HOST_LIB)
    vp = ctypes.c_void_p
    lib.phost_build.restype = vp
    lib.phost_build.argtypes = [ctypes.c_longlong, vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_float,
                                ctypes.c_float, vp, vp, ctypes.c_int]
    lib.phost_count.restype = ctypes.c_longlong
    lib.phost_count.argtypes = [vp, ctypes.c_int]
    lib.phost_get_perm.argtypes = [vp, vp]
    lib.phost_get_leaves.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.phost_get_tasks.argtypes = [vp, vp, vp]
    lib.phost_free.argtypes = [vp]
    return lib


def build_lists(posi: np.ndarray, tag: np.ndarray, int2pos: float, maxleaf: int, open_angle: float,
                cutoff_radius: float, bdl: np.ndarray, bdr: np.ndarray, nthreads: int = 0) -> dict:
    """kd-tree + dual walk (libphotons_host.so).  Returns perm (tree order ->
    input index), leaf_ip, leaf_np, active, task_t, task_s (0-based leaf ids)."""
    lib = _host_lib()
    P = lambda a: a.ctypes.data_as(__import__("ctypes").c_void_p)  # noqa: E731
    posi = np.ascontiguousarray(posi, np.int32)
    tag = np.ascontiguousarray(tag, np.int32)
    bdl = np.ascontiguousarray(bdl, np.float32)
    bdr = np.ascontiguousarray(bdr, np.float32)
    h = lib.phost_build(posi.shape[0], P(posi), P(tag), int2pos, maxleaf, open_angle, cutoff_radius,
                        P(bdl), P(bdr), nthreads)
    if not h:
        raise ValueError("phost_build rejected its input")
    try:
        n = lib.phost_count(h, 0)
        nl = lib.phost_count(h, 1)
        nt = lib.phost_count(h, 3)
        perm = np.empty(n, np.int32)
        ip = np.empty(nl, np.int32)
        npl = np.empty(nl, np.int32)
        act = np.empty(nl, np.int32)
        tt = np.empty(nt, np.int32)
        ts = np.empty(nt, np.int32)
        lib.phost_get_perm(h, P(perm))
        lib.phost_get_leaves(h, P(ip), P(npl), P(act), None, None)
        lib.phost_get_tasks(h, P(tt), P(ts))
    finally:
        lib.phost_free(h)
    return {"perm": perm, "leaf_ip": ip, "leaf_np": npl, "active": act, "task_t": tt, "task_s": ts}


def tree_workload(name: str, pos: np.ndarray, box: float, mesh: int, maxleaf: int, open_angle: float,
                  eps: float, softening: int, periodic_ghosts: bool = True, nthreads: int = 0,
                  meta: dict | None = None) -> Workload:
    """Local particles in [0, box) -> reference integer grid, ghost copies
    within cutoffRadius, tree + walk lists.  The returned positions are the
    grid positions (posi * INT2POS, FP64) in tree order, ghosts included;
    meta['tag'] marks local particles (1) and ghosts (0), meta['local_index']
    maps each row back to the input particle."""
    cell = box / mesh
    rs = 1.25 * cell                                   # src/setparam.c:29
    cutoff_radius = 4.5 * rs                           # src/setparam.c:30
    int2pos = int_grid(box, cutoff_radius)
    posi, box_i = to_grid(pos, box, int2pos)
    if periodic_ghosts:
        width_i = int(np.ceil(np.float32(cutoff_radius) / int2pos))
        gpos, gsrc = ghost_images(posi, box_i, width_i)
    else:
        gpos, gsrc = np.zeros((0, 3), np.int32), np.zeros(0, np.int64)
    allp = np.concatenate([posi, gpos])
    tag = np.concatenate([np.ones(posi.shape[0], np.int32), np.zeros(gpos.shape[0], np.int32)])
    local_index = np.concatenate([np.arange(posi.shape[0]), gsrc])
    bdl, bdr = combine_bounds(allp, int2pos, box)
    lists = build_lists(allp, tag, int2pos, maxleaf, np.float32(open_angle), np.float32(cutoff_radius),
                        bdl, bdr, nthreads)
    perm = lists["perm"]
    m = dict(meta or {})
    m.update({"int2pos": int2pos, "box_i": box_i, "posi": allp[perm], "tag": tag[perm],
              "local_index": local_index[perm], "nlocal": posi.shape[0], "maxleaf": maxleaf,
              "open_angle": open_angle, "cutoff_radius": cutoff_radius, "mesh": mesh,
              "bdl": bdl, "bdr": bdr, "active_leaf": lists["active"]})
    w = Workload(name, allp[perm].astype(np.float64) * int2pos, lists["leaf_ip"], lists["leaf_np"],
                 lists["task_t"], lists["task_s"], 0.0, rs, eps, softening=softening, meta=m)
    return w

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
    lib.phost_build_tree.restype = vp
    lib.phost_build_tree.argtypes = [ctypes.c_longlong, vp, vp, ctypes.c_double, ctypes.c_int, vp, vp, ctypes.c_int]
    lib.phost_walk.restype = ctypes.c_int
    lib.phost_walk.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.phost_walk_stream.restype = ctypes.c_int
    lib.phost_walk_stream.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      vp, vp, vp, ctypes.c_int]
    lib.phost_count.restype = ctypes.c_longlong
    lib.phost_count.argtypes = [vp, ctypes.c_int]
    lib.phost_get_perm.argtypes = [vp, vp]
    lib.phost_get_leaves.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.phost_get_nodes.argtypes = [vp, vp, vp, vp]
    lib.phost_get_tasks.argtypes = [vp, vp, vp]
    lib.phost_get_m2l.argtypes = [vp, vp, vp]
    lib.phost_free.argtypes = [vp]
    return lib


def build_lists(posi: np.ndarray, tag: np.ndarray, int2pos: float, maxleaf: int, open_angle: float,
                cutoff_radius: float, bdl: np.ndarray, bdr: np.ndarray, nthreads: int = 0,
                walk: bool = True, with_tree: bool = False, with_m2l: bool = False) -> dict:
    """kd-tree + dual walk (libphotons_host.so).  Returns perm (tree order ->
    input index), leaf_ip, leaf_np, active, task_t, task_s (0-based leaf ids)
    and the host seconds of the two steps (tree_s, walk_s).  walk=False skips
    the walk (empty task arrays); with_tree adds the tree as the device walk
    reads it: node_son (nnode x 2), node_center, node_width, leaf_center,
    leaf_width (x 3, FP32).  with_m2l adds the well-separated cell pairs of the
    walk, m2l_t / m2l_s (cells: leaf id, or nleaf + node id)."""
    import time
    lib = _host_lib()
    P = lambda a: a.ctypes.data_as(__import__("ctypes").c_void_p)  # noqa: E731
    posi = np.ascontiguousarray(posi, np.int32)
    tag = np.ascontiguousarray(tag, np.int32)
    bdl = np.ascontiguousarray(bdl, np.float32)
    bdr = np.ascontiguousarray(bdr, np.float32)
    t0 = time.perf_counter()
    h = lib.phost_build_tree(posi.shape[0], P(posi), P(tag), int2pos, maxleaf, P(bdl), P(bdr), nthreads)
    if not h:
        raise ValueError("phost_build rejected its input")
    t1 = time.perf_counter()
    try:
        if walk and lib.phost_walk(h, open_angle, cutoff_radius, nthreads) != 0:
            raise ValueError("phost_walk failed")
        t2 = time.perf_counter()
        n = lib.phost_count(h, 0)
        nl = lib.phost_count(h, 1)
        nn = lib.phost_count(h, 2)
        nt = lib.phost_count(h, 3)
        perm = np.empty(n, np.int32)
        ip = np.empty(nl, np.int32)
        npl = np.empty(nl, np.int32)
        act = np.empty(nl, np.int32)
        tt = np.empty(nt, np.int32)
        ts = np.empty(nt, np.int32)
        out = {"perm": perm, "leaf_ip": ip, "leaf_np": npl, "active": act, "task_t": tt, "task_s": ts}
        lib.phost_get_perm(h, P(perm))
        if with_tree:
            out.update(node_son=np.empty((nn, 2), np.int32), node_center=np.empty((nn, 3), np.float32),
                       node_width=np.empty((nn, 3), np.float32), leaf_center=np.empty((nl, 3), np.float32),
                       leaf_width=np.empty((nl, 3), np.float32))
            lib.phost_get_leaves(h, P(ip), P(npl), P(act), P(out["leaf_center"]), P(out["leaf_width"]))
            lib.phost_get_nodes(h, P(out["node_son"]), P(out["node_center"]), P(out["node_width"]))
        else:
            lib.phost_get_leaves(h, P(ip), P(npl), P(act), None, None)
        lib.phost_get_tasks(h, P(tt), P(ts))
        if with_m2l:
            nm = lib.phost_count(h, 4)
            out.update(m2l_t=np.empty(nm, np.int32), m2l_s=np.empty(nm, np.int32))
            lib.phost_get_m2l(h, P(out["m2l_t"]), P(out["m2l_s"]))
    finally:
        lib.phost_free(h)
    out.update(tree_s=t1 - t0, walk_s=t2 - t1 if walk else 0.0)
    return out


def tree_workload(name: str, pos: np.ndarray, box: float, mesh: int, maxleaf: int, open_angle: float,
                  eps: float, softening: int, periodic_ghosts: bool = True, nthreads: int = 0,
                  meta: dict | None = None, progress=None) -> Workload:
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
    if progress:
        progress(f"{allp.shape[0]} particles ({gpos.shape[0]} ghost copies) on the grid; building tree + lists")
    lists = build_lists(allp, tag, int2pos, maxleaf, np.float32(open_angle), np.float32(cutoff_radius),
                        bdl, bdr, nthreads, with_tree=True, with_m2l=True)
    perm = lists["perm"]
    m = dict(meta or {})
    m.update({"int2pos": int2pos, "box_i": box_i, "posi": allp[perm], "tag": tag[perm], "perm": perm,
              "local_index": local_index[perm], "nlocal": posi.shape[0], "maxleaf": maxleaf,
              "open_angle": open_angle, "cutoff_radius": cutoff_radius, "mesh": mesh,
              "bdl": bdl, "bdr": bdr, "active_leaf": lists["active"],
              "tree": {k: lists[k] for k in ("node_son", "node_center", "node_width", "leaf_center", "leaf_width")},
              "m2l": (lists["m2l_t"], lists["m2l_s"]),
              "list_seconds": {"tree": lists["tree_s"], "walk": lists["walk_s"], "threads": nthreads}})
    w = Workload(name, allp[perm].astype(np.float64) * int2pos, lists["leaf_ip"], lists["leaf_np"],
                 lists["task_t"], lists["task_s"], 0.0, rs, eps, softening=softening, meta=m)
    return w


# --------------------------------------------------------------------------
# Particle generators

def zeldovich_grid(n: int, box: float, rms_frac: float, g: np.random.Generator,
                   dtype=np.float32) -> np.ndarray:
    """n^3 grid in [0, box) displaced by a Zel'dovich field psi = i k / k^2 delta
    of a Gaussian random field with P(k) ~ k^-2, scaled to an rms |psi| of
    rms_frac * (box / n).  Returns (n^3, 3) float64, wrapped into [0, box)."""
    d = box / n
    w = g.standard_normal((n, n, n), dtype=np.float32 if dtype == np.float32 else np.float64)
    wk = np.fft.rfftn(w)
    del w
    kx = np.fft.fftfreq(n).astype(dtype)[:, None, None]
    ky = np.fft.fftfreq(n).astype(dtype)[None, :, None]
    kz = np.fft.rfftfreq(n).astype(dtype)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = (k2 ** -1.5).astype(dtype)          # |delta| ~ k^-1, psi ~ k / k^2 delta
    amp[0, 0, 0] = 0.0
    psi = np.empty((n ** 3, 3), np.float64)
    for j, kj in enumerate((kx, ky, kz)):
        comp = np.fft.irfftn(1j * kj * amp * wk, s=(n, n, n), axes=(0, 1, 2))
        psi[:, j] = comp.reshape(-1)
        del comp
    del wk
    psi *= rms_frac * d / np.sqrt(np.mean(np.sum(psi * psi, axis=1)))
    gr = (np.arange(n, dtype=np.float64) + 0.5) * d
    q = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3)
    return np.mod(q + psi, box)


def _split_counts(total: int, weights: np.ndarray) -> np.ndarray:
    """Integer counts proportional to weights summing to total (largest remainder)."""
    x = total * weights / weights.sum()
    c = np.floor(x).astype(np.int64)
    rem = total - int(c.sum())
    if rem > 0:
        c[np.argsort(-(x - c), kind="stable")[:rem]] += 1
    return c


def _directions(g: np.random.Generator, m: int) -> np.ndarray:
    v = g.standard_normal((m, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def plummer_spheres(n_total: int, n_spheres: int, box: float, spacing: float, g: np.random.Generator,
                    mass_decades: float = 3.0, a_range=(0.5, 5.0), trunc: float = 10.0) -> np.ndarray:
    """n_total particles in n_spheres Plummer spheres with uniform centres,
    masses log-uniform over `mass_decades` (particle counts proportional to
    mass), scale radius a uniform in a_range * spacing, truncated at trunc a
    (inverse CDF of M(<r)/M = r^3 / (r^2 + a^2)^{3/2})."""
    mass = 10.0 ** (mass_decades * g.random(n_spheres))
    counts = _split_counts(n_total, mass)
    centres = g.random((n_spheres, 3)) * box
    a = (a_range[0] + (a_range[1] - a_range[0]) * g.random(n_spheres)) * spacing
    owner = np.repeat(np.arange(n_spheres), counts)
    xmax = trunc ** 3 / (trunc ** 2 + 1.0) ** 1.5
    x = g.random(n_total) * xmax
    x = np.clip(x, 1e-300, None)
    r = a[owner] / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    pos = centres[owner] + r[:, None] * _directions(g, n_total)
    return np.mod(pos, box)


def nfw_halos(n_total: int, n_halos: int, box: float, n_mean_total: int, g: np.random.Generator,
              mass_decades: float = 3.0, c_range=(5.0, 15.0)) -> np.ndarray:
    """n_total particles in n_halos NFW halos: masses log-uniform over
    `mass_decades` (counts proportional to mass), concentration uniform in
    c_range, truncated at r200 of 200 times the mean number density
    n_mean_total / box^3, radii by inverting mu(x) = ln(1 + x) - x / (1 + x)."""
    mass = 10.0 ** (mass_decades * g.random(n_halos))
    counts = _split_counts(n_total, mass)
    centres = g.random((n_halos, 3)) * box
    conc = c_range[0] + (c_range[1] - c_range[0]) * g.random(n_halos)
    rho = n_mean_total / box ** 3
    r200 = (3.0 * counts / (4.0 * np.pi * 200.0 * rho)) ** (1.0 / 3.0)
    owner = np.repeat(np.arange(n_halos), counts)
    mu = lambda x: np.log1p(x) - x / (1.0 + x)  # noqa: E731
    target = g.random(n_total) * mu(conc[owner])
    lo = np.zeros(n_total)
    hi = conc[owner].copy()
    for _ in range(60):                           # bisection to ~1e-18 relative
        mid = 0.5 * (lo + hi)
        below = mu(mid) < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    r = 0.5 * (lo + hi) * (r200 / conc)[owner]
    pos = centres[owner] + r[:, None] * _directions(g, n_total)
    return np.mod(pos, box)


def clustered_mix(n1: int, box: float, seed: int, n_spheres: int = 2000) -> np.ndarray:
    """Half of a Zel'dovich-displaced n1^3 grid (seeded random half) plus the
    other half in Plummer spheres (configs at 256^3 and 512^3)."""
    g = rng_for(seed)
    n = n1 ** 3
    grid = zeldovich_grid(n1, box, 0.3, g)
    keep = g.permutation(n)[: n // 2]
    keep.sort()
    sph = plummer_spheres(n - keep.size, n_spheres, box, box / n1, g)
    return np.concatenate([grid[keep], sph])


# --------------------------------------------------------------------------
# The benchmark configurations (sizes scale down through n1 for tests)

def config_plain(n1: int = 64, box: float = 100.0, seed: int = 20261202, softening: int = 1,
                 maxleaf: int = 32, open_angle: float = 0.5, nthreads: int = 0) -> Workload:
    """Plain P2P: n1^3 Zel'dovich grid (rms 0.3 spacing), mesh n1^3, leaves
    <= 32, opening angle 0.5; Plummer eps = spacing / 40 (REF: 0.03 cell,
    the reference's SoftenScale, src/setparam.c:16-23)."""
    g = rng_for(seed)
    pos = zeldovich_grid(n1, box, 0.3, g)
    d = box / n1
    eps = d / 40.0 if softening == 1 else 0.03 * d
    return tree_workload(f"plain{n1}", pos, box, n1, maxleaf, open_angle, eps, softening, nthreads=nthreads,
                         meta={"seed": seed, "n1": n1})


def config_clustered(n1: int = 256, box: float = 400.0, seed: int = 20261203, softening: int = 1,
                     maxleaf: int = 64, open_angle: float = 0.5, n_spheres: int = 2000,
                     nthreads: int = 0, progress=None) -> Workload:
    """Plain P2P at scale: half Zel'dovich grid, half in Plummer spheres;
    leaves <= 64, mesh n1^3, Plummer eps = spacing / 40."""
    pos = clustered_mix(n1, box, seed, n_spheres)
    if progress:
        progress(f"{pos.shape[0]} particles generated")
    d = box / n1
    eps = d / 40.0 if softening == 1 else 0.03 * d
    return tree_workload(f"clustered{n1}", pos, box, n1, maxleaf, open_angle, eps, softening, nthreads=nthreads,
                         meta={"seed": seed, "n1": n1, "n_spheres": n_spheres}, progress=progress)


def config_precision(n1: int = 128, box: float = 100.0, seed: int = 20261204, n_halos: int = 500,
                     halo_frac: float = 0.6, offset_boxes: float = 1000.0, maxleaf: int = 32,
                     open_angle: float = 0.5, nthreads: int = 0) -> Workload:
    """Precision stress: 60% of n1^3 particles in NFW halos, the rest on a
    displaced grid; Plummer eps = spacing / 50.  Lists are built on the box
    coordinates; meta['offset'] is the translation (offset_boxes * box per
    axis) that the precision comparison applies to absolute coordinates."""
    g = rng_for(seed)
    n = n1 ** 3
    nh = int(round(halo_frac * n))
    grid = zeldovich_grid(n1, box, 0.3, g)
    keep = np.sort(g.permutation(n)[: n - nh])
    halos = nfw_halos(nh, n_halos, box, n, g)
    pos = np.concatenate([grid[keep], halos])
    d = box / n1
    w = tree_workload(f"precision{n1}", pos, box, n1, maxleaf, open_angle, d / 50.0, 1, nthreads=nthreads,
                      meta={"seed": seed, "n1": n1})
    w.meta["offset"] = offset_boxes * box
    return w


def active_subset(w: Workload, fraction: float, seed: int = 20261207, level: int = 1,
                  n_grid: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Hierarchical-time-step active mask of a tree workload: the first
    `fraction` of the local particles in a seeded order that puts members of
    Plummer spheres (input index >= n_grid, clustered_mix order) first get
    act = level, all others (and every ghost copy) act = 0.  Returns (act,
    task mask): the tasks whose target leaf holds an active particle, the
    targets the reference's active walk emits (src/fmm.c:1308-1320, :1046).
    Without n_grid the order is a plain seeded permutation."""
    m = w.meta
    tag = m["tag"]
    local = m["local_index"]
    nlocal = int(m["nlocal"])
    g = rng_for(seed)
    order = g.permutation(nlocal)
    if n_grid is not None:
        member = order >= n_grid
        order = np.concatenate([order[member], order[~member]])
    nact = int(round(fraction * nlocal))
    chosen = np.zeros(nlocal, bool)
    chosen[order[:nact]] = True
    act = np.where((tag == 1) & chosen[local], level, 0).astype(np.uint8)
    # targets: leaves holding an active particle
    leaf_of = np.repeat(np.arange(w.nleaf), w.leaf_np)
    has = np.zeros(w.nleaf, bool)
    has[leaf_of[act == level]] = True
    return act, has[w.task_t]

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

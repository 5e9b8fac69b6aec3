"""Writes tests/golden/lists_*.npz: interaction lists produced by the
reference's own list producer (build_localtree + walk_task_p2p from
src/fmm.c, compiled into oracle/_ref/libreflist.so by oracle/ref.mk) for small
seeded particle sets.  Each fixture holds the inputs (integer positions,
tags, tree parameters, root box) and the reference's outputs (particle
permutation, leaf ranges, leaf pairs).  Run in the build container, where
the reference tree exists:  python tests/golden/make_golden_lists.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "photons-on-mt3000_amd"), os.path.join(ROOT, "oracle")]

import pyoracle  # noqa: E402
from photons_p2p import workloads as W  # noqa: E402

CASES = [
    # name, particles per side, mesh, maxleaf, opening angle, seed
    ("lists_n12_leaf8_theta04", 12, 12, 8, 0.4, 101),
    ("lists_n16_leaf32_theta05", 16, 16, 32, 0.5, 102),
]


def inputs(n1, mesh, seed, box=100.0):
    g = np.random.default_rng(seed)
    d = box / n1
    gr = (np.arange(n1) + 0.5) * d
    pos = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3)
    pos = np.mod(pos + g.normal(0.0, 0.3 * d, pos.shape), box)
    rs = 1.25 * box / mesh
    cr = 4.5 * rs
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(pos, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(len(posi), np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    return allp, tag, int2pos, np.float32(cr), bdl, bdr


def main():
    for name, n1, mesh, maxleaf, theta, seed in CASES:
        posi, tag, int2pos, cr, bdl, bdr = inputs(n1, mesh, seed)
        out = pyoracle.reflist_build(posi, tag, int2pos, maxleaf, np.float32(theta), cr, bdl, bdr)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), posi=posi, tag=tag, int2pos=int2pos,
                            maxleaf=maxleaf, open_angle=np.float32(theta), cutoff_radius=cr, bdl=bdl, bdr=bdr,
                            **out)
        print(name, posi.shape[0], out["leaf_ip"].size, out["task_t"].size)


if __name__ == "__main__":
    main()

"""The whole force of the split against an FP64 Ewald sum in a periodic box:
P2P and the multipole pass (the short-range erfc part) plus the particle mesh
(the long-range erf part), all on the device from grid positions, equal
Newton's periodic force.  Ewald's split with alpha = 1 / (2 rs) is the same
split, so each half is also compared with its own sum: real space for the
short range (truncated at the reference's cutoff), reciprocal space for the
mesh.  This ties the normalisations of the pieces together (the walk finds no
well-separated pair at this size, so the multipole pass adds nothing here;
it is checked against its oracle in test_gpu_fmm.py).  Measured: short range
median 3e-5, max 8e-4; mesh median 4.4e-3, p90 9.7e-3, max 4.0e-2 of |a|."""
import math

import numpy as np
import pytest
from scipy.special import erfc

import photons_p2p as pp
from photons_p2p import workloads as W

pytestmark = pytest.mark.gpu


def _ewald(x, m, box, alpha, eps, kcut=30.0):
    """Real-space (nearest 27 images, Plummer-softened distance as the P2P
    kernel's) and reciprocal-space accelerations, FP64."""
    n = x.shape[0]
    imgs = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], float) * box
    acc_r = np.zeros((n, 3))
    for i in range(n):
        d = (x[None, :, :] + imgs[:, None, :]) - x[i]
        r = np.sqrt((d * d).sum(2) + eps * eps)
        f = (erfc(alpha * r) + 2 * alpha * r / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2)) / r ** 3
        f[(d * d).sum(2) == 0] = 0.0
        acc_r[i] = (m[None, :, None] * d * f[:, :, None]).sum((0, 1))
    nmax = int(math.ceil(math.sqrt(kcut) * 2 * alpha / (2 * math.pi / box)))
    g = np.arange(-nmax, nmax + 1)
    K = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    K = K[np.any(K != 0, axis=1)] * (2 * math.pi / box)
    k2 = (K ** 2).sum(1)
    keep = k2 / (4 * alpha ** 2) < kcut
    K, k2 = K[keep], k2[keep]
    w = np.exp(-k2 / (4 * alpha ** 2)) / k2
    S = (m[None, :] * np.exp(-1j * (K @ x.T))).sum(1)
    im = (np.exp(1j * (x @ K.T)) * S[None, :]).imag
    acc_k = -(4 * math.pi / box ** 3) * (im * w[None, :]) @ K
    return acc_r, acc_k


def _stats(a, ref, norm):
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(norm, axis=1)
    return float(np.median(e)), float(np.percentile(e, 90)), float(e.max())


def test_whole_split_force_matches_ewald(gpu_ctx):
    g = np.random.default_rng(2026)
    box, N, n, mass = 100.0, 32, 600, 0.5
    rs = 1.25 * box / N
    cr = 4.5 * rs
    x = g.uniform(0, box, (n, 3))
    x[:200] = (40.0 + g.normal(size=(200, 3)) * 4.0) % box
    int2pos = W.int_grid(box, cr)
    posi, box_i = W.to_grid(x, box, int2pos)
    gp, _ = W.ghost_images(posi, box_i, int(np.ceil(np.float32(cr) / int2pos)))
    allp = np.concatenate([posi, gp])
    tag = np.concatenate([np.ones(n, np.int32), np.zeros(len(gp), np.int32)])
    bdl, bdr = W.combine_bounds(allp, int2pos, box)
    boxf = float(box_i * int2pos)

    gpu_ctx.set_tuning()
    eps = 0.02 * box / N                                 # the reference's softening scale
    gpu_ctx.set_law(pp.ForceLaw(eps=eps, rs=rs, mass=mass, softening=pp.SOFT_PLUMMER))
    gpu_ctx.build_tree(allp, tag, int2pos, 16, bdl, bdr)
    perm = gpu_ctx.get_tree()["perm"]
    gpu_ctx.walk_tree(np.float32(0.5), np.float32(cr))
    gpu_ctx.zero_acc()
    gpu_ctx.compute(pp.EXEC_QUEUE)
    gpu_ctx.fmm_short()
    short = gpu_ctx.acc()
    gpu_ctx.pm_add_tree(boxf, N, rs, mass)
    total = gpu_ctx.acc()

    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)                    # input index -> tree position
    short, total = short[inv[:n]], total[inv[:n]]
    xs = posi.astype(np.float64) * int2pos
    acc_r, acc_k = _ewald(xs, np.full(n, mass), boxf, 1.0 / (2.0 * rs), eps)
    ref = acc_r + acc_k
    s_short = _stats(short, acc_r, ref)
    s_long = _stats(total - short, acc_k, ref)
    s_tot = _stats(total, ref, ref)
    print(f"case: n={n} mesh {N}^3; errors relative to |a| (median, p90, max): short {s_short}, "
          f"long {s_long}, total {s_tot}", flush=True)
    assert s_short[0] < 2e-4 and s_short[2] < 5e-3, s_short
    assert s_long[0] < 1e-2 and s_long[1] < 2.5e-2 and s_long[2] < 0.1, s_long
    assert s_tot[0] < 1e-2 and s_tot[1] < 2.5e-2 and s_tot[2] < 0.1, s_tot

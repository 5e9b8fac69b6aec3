"""PM (particle-mesh) long-range force, numpy restatement -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and bench.py's checker; the product library never uses it.

Follows the reference's partmesh_kick for INTXYZ particles on one domain
(src/partmesh.c:1476-2954) and its convolution (src/conv.f90:168-352):

* positions: posf = (float)(posi * INT2POS) (src/partmesh.c:1569-1571)
* cell-centred CIC: i = (int)(posf * norm), w = (posf - (i + 0.5) delta) norm,
  neighbour i + 1 when w > 0, else i - 1 with w = -w (src/partmesh.c:1573-1612);
  deposit MASSPART * w_x * w_y * w_z in FP32 products (:1622-1672)
* the mass per cell becomes a density: pmesh *= renormal, renormal =
  (NSIDEMESH / BOXSIZE)^3 in FP32 (src/partmesh.c:1797-1807)
* Green's function on the r2c spectrum (src/conv.f90:262-318):
  G = pref exp(-k2 ismth2) / k2 / (sinc_x sinc_y sinc_z)^4, G(0) = 0,
  ismth2 = (2 pi splitRadius / box)^2, pref = box^2 / (pi N^3), pi = 3.1415926
  as the reference's REAL constant; unnormalised inverse transform
* acceleration: 4-point central differences of the potential,
  f1 invx (p[+1] - p[-1]) - f2 invx (p[+2] - p[-2]) with f1 = 4/3, f2 = 1/6,
  invx = 0.5 N / box, at the 8 CIC cells, weighted by the same CIC weights
  (src/partmesh.c:2596-2922)

The reference keeps the mesh in FP32 (cpuType) and hands the FFT to
2DECOMP&FFT/FFTW (not in the reference tree; its precision is a build option
of that library).  This restatement keeps the CIC weights in the reference's
FP32 arithmetic and runs the mesh, the transforms and the gradient in FP64.
One domain holds the whole periodic mesh: cell indices wrap modulo N where
the reference reads the ghost layers (NPAD = 3) of its neighbours.
"""
from __future__ import annotations

import numpy as np

PI_REF = float(np.float32(3.1415926))     # src/conv.f90:183, a REAL(cputype) parameter


def positions(posi: np.ndarray, int2pos: float) -> np.ndarray:
    """(float)(posi * INT2POS): FP64 product rounded once to FP32."""
    return (np.asarray(posi, np.int32).astype(np.float64) * np.float64(int2pos)).astype(np.float32)


def cic(posf: np.ndarray, box: float, nside: int):
    """Per axis: cell i, neighbour ii, neighbour weight w and own weight wn
    (FP32), in the reference's operation order (src/partmesh.c:1573-1612)."""
    norm = np.float32(np.float32(nside) / np.float32(box))        # NSIDEMESH / BOXSIZE, FP32
    delta = np.float32(1.0 / np.float64(norm))                     # 1.0 / norm in FP64, stored FP32
    i = (posf * norm).astype(np.int32)                             # FP32 product, truncation
    w = ((posf.astype(np.float64) - (i.astype(np.float64) + 0.5) * np.float64(delta))
         * np.float64(norm)).astype(np.float32)
    up = w > 0
    ii = np.where(up, i + 1, i - 1).astype(np.int32)
    w = np.where(up, w, -w).astype(np.float32)
    wn = (1.0 - w.astype(np.float64)).astype(np.float32)
    return i, ii, w, wn


def _corners(cx, cy, cz):
    """The 8 (cell, weight) corners in the reference's order."""
    (i, ii, wi, win), (j, jj, wj, wjn), (k, kk, wk, wkn) = cx, cy, cz
    for (c3, w3) in ((k, wkn), (kk, wk)):
        for (c2, w2) in ((j, wjn), (jj, wj)):
            for (c1, w1) in ((i, win), (ii, wi)):
                yield c1, c2, c3, w1, w2, w3


def deposit(posi: np.ndarray, tag: np.ndarray | None, int2pos: float, box: float, nside: int,
            mass: float) -> np.ndarray:
    """Mass per cell, C order [x][y][z], FP64 sums of the FP32 CIC products."""
    posi = np.asarray(posi, np.int32).reshape(-1, 3)
    if tag is not None:
        posi = posi[np.asarray(tag) == 1]
    posf = positions(posi, int2pos)
    ax = [cic(posf[:, d], box, nside) for d in range(3)]
    m = np.float32(mass)
    N = nside
    rho = np.zeros(N ** 3, np.float64)
    for c1, c2, c3, w1, w2, w3 in _corners(*ax):
        wt = ((m * w1) * w2) * w3                                  # FP32, left to right
        idx = ((c1 % N).astype(np.int64) * N + (c2 % N)) * N + (c3 % N)
        np.add.at(rho, idx, wt.astype(np.float64))
    return rho.reshape(N, N, N)


def renormal(nside: int, box: float) -> float:
    """(NSIDEMESH / BOXSIZE)^3 as the reference's FP32 products (src/partmesh.c:1797-1798)."""
    r = np.float32(np.float32(nside) / np.float32(box))
    return float(np.float32(np.float32(r * r) * r))


def green(nside: int, split_radius: float, box: float) -> np.ndarray:
    """G on the r2c half spectrum [N][N][N/2 + 1] (src/conv.f90:262-318)."""
    N = nside
    smooth = float(np.float32(split_radius))
    bx = float(np.float32(box))
    ismth2 = (2 * PI_REF * smooth / bx) ** 2
    pref = bx * bx / (PI_REF * N * N * N)

    def wav(n):
        return np.where(n > N // 2, n - N, n)

    def sinc(n):
        x = PI_REF * n / N
        with np.errstate(invalid="ignore", divide="ignore"):
            f = np.sin(x) / x
        return np.where(n == 0, 1.0, f)

    l = wav(np.arange(N))[:, None, None].astype(np.float64)
    m = wav(np.arange(N))[None, :, None].astype(np.float64)
    n = np.arange(N // 2 + 1)[None, None, :].astype(np.float64)
    k2 = l * l + m * m + n * n
    ff = 1.0 / (sinc(l) * sinc(m) * sinc(n))
    with np.errstate(invalid="ignore", divide="ignore"):
        ff2 = ff * ff
        g = np.exp(-k2 * ismth2) / k2 * (ff2 * ff2) * pref
    g[0, 0, 0] = 0.0
    return g


def potential(rho: np.ndarray, split_radius: float, box: float) -> np.ndarray:
    """Unnormalised inverse transform of G * FFT(renormal * rho), rho the mass
    per cell (numpy's irfftn divides by N^3)."""
    N = rho.shape[0]
    spec = np.fft.rfftn(rho * renormal(N, box)) * green(N, split_radius, box)
    return np.fft.irfftn(spec, s=rho.shape, axes=(0, 1, 2)) * float(N) ** 3


def accelerations(pot: np.ndarray, posi: np.ndarray, int2pos: float, box: float) -> np.ndarray:
    """(n, 3) FP64 accelerations at the particles (src/partmesh.c:2596-2922)."""
    N = pot.shape[0]
    posi = np.asarray(posi, np.int32).reshape(-1, 3)
    posf = positions(posi, int2pos)
    ax = [cic(posf[:, d], box, N) for d in range(3)]
    invx = float(np.float32(0.5 * N / np.float64(np.float32(box))))
    f1, f2 = float(np.float32(4.0 / 3.0)), float(np.float32(1.0 / 6.0))
    P = pot
    acc = np.zeros((posi.shape[0], 3))
    for d in range(3):
        for c1, c2, c3, w1, w2, w3 in _corners(*ax):
            c = [c1, c2, c3]

            def at(o):
                cc = list(c)
                cc[d] = cc[d] + o
                return P[cc[0] % N, cc[1] % N, cc[2] % N]
            dp = f1 * invx * (at(1) - at(-1)) - f2 * invx * (at(2) - at(-2))
            wt = w1.astype(np.float64) * w2.astype(np.float64) * w3.astype(np.float64)
            acc[:, d] += wt * dp
    return acc


def pm_forces(posi, tag, int2pos, box, nside, split_radius, mass):
    """Accelerations of the particles with tag == 1 (all when tag is None);
    the others get zero.  Returns (acc (n, 3) FP64, rho, pot)."""
    posi = np.asarray(posi, np.int32).reshape(-1, 3)
    rho = deposit(posi, tag, int2pos, box, nside, mass)
    pot = potential(rho, split_radius, box)
    acc = accelerations(pot, posi, int2pos, box)
    if tag is not None:
        acc[np.asarray(tag) != 1] = 0.0
    return acc, rho, pot

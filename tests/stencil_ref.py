"""The pair list DESGRID_NEIGH_BUILD keeps, restated in Python from the
reference text so that it can pin the oracle's (and the library's) search
in the regimes where the reference's stencil is incomplete.

* DES cell of a particle (desgrid_mod.f:391-415, 905-911):
  i = min(iend2, max(istart2, floor((x - dg_xstart) * dg_dxinv) + istart1)),
  istart1 = 2, istart2 = 1, iend2 = isize + 2; dg_xstart = XE(1) = X_MIN,
  dg_dxinv = isize / (XE(IMAX+1) - XE(1)) with isize = NINT(XLENGTH /
  (XLENGTH / DESGRIDSEARCH_IMAX)) (desgrid_mod.f:572-574,833-847); the linear
  index DG_FUNIJK = J + (I-1) NJ + (K-1) NJ NI (:184-188, 776-786), J fastest.
* Stencil (desgrid_mod.f:1107-1152): 27 cells, or per axis 2 cells when
  1 > 2 MAX_RADIUS FACTOR_RLM dg_dxinv ("coarse"): the own cell and the
  neighbour on the side of the particle's half of its cell.
* Kept (:1160-1182): j != i, not (j < i and j normal), j exists, and
  |x_i - x_j|^2 <= (FACTOR_RLM (r_i + r_j))^2; rows in stencil order (K, J, I
  outer to inner), each cell's particles in ascending index (the serial
  DESGRID_PIC fill order).

Between 2.4 r_max and 4.8 r_max (coarse) and below 2.4 r_max (27 cells
narrower than the cut-off) this list misses pairs within the cut-off; the
brute-force set (NSQUARE's criterion without a grid) shows which.
"""
from __future__ import annotations

import math

import numpy as np


def faces(lo, d, n):
    """XE(0..n+1) built as SET_GEOMETRY_DES does for a uniform grid: XE(0) =
    lo - d, then cumulative sums of d."""
    f = [lo - d]
    for _ in range(n + 1):
        f.append(f[-1] + d)
    return f


def des_grid(p):
    """(start, inv, size) per axis of the DES grid."""
    out = []
    for lo, length, cells, dg in ((p.x_min, p.xlength, p.imax, p.desgridsearch_imax),
                                  (p.y_min, p.ylength, p.jmax, p.desgridsearch_jmax),
                                  (p.z_min, p.zlength, p.kmax, p.desgridsearch_kmax)):
        d = length / cells
        f = faces(lo, d, cells)
        if dg <= 0:
            width = 3.0 * (2.0 * p.max_radius)
            dg = max(1, int(length / min(width, d)))
        size = int(round((f[cells + 1] - f[1]) / (length / dg)))
        inv = size / (f[cells + 1] - f[1])
        out.append((f[1], inv, size))
    return out


def des_cells(p, pos):
    """(I, J, K) 1-based DES cells of every particle (pos (3, n)) and the
    1-based DG_PIJK."""
    g = des_grid(p)
    ijk = []
    for a in range(3):
        start, inv, size = g[a]
        c = np.array([math.floor((x - start) * inv) + 2 for x in pos[a]])
        ijk.append(np.clip(c, 1, size + 2))
    ni, nj = g[0][2] + 2, g[1][2] + 2
    dg = ijk[1] + (ijk[0] - 1) * nj + (ijk[2] - 1) * nj * ni
    return ijk, dg


def reference_pairs(p, pos, radius):
    """NEIGHBOR_INDEX (1-based exclusive ends) and NEIGHBORS (1-based) of the
    serial DESGRID_NEIGH_BUILD on these particles (all normal, none rogue)."""
    n = pos.shape[1]
    g = des_grid(p)
    (ci, cj, ck), _ = des_cells(p, pos)
    members = {}
    for q in range(n):
        members.setdefault((int(ci[q]), int(cj[q]), int(ck[q])), []).append(q)
    frlm = p.factor_rlm
    coarse = [1.0 > 2.0 * p.max_radius * frlm * g[a][1] for a in range(3)]
    index = np.zeros(n, dtype=np.int64)
    nbrs = []
    for q in range(n):
        lo, hi = [1, 1, 1], [1, 1, 1]
        for a in range(3):
            if coarse[a]:
                t = (pos[a, q] - g[a][0]) * g[a][1]
                if t - math.floor(t) >= 0.5:
                    lo[a] = 0
                else:
                    hi[a] = 0
        for lk in range(int(ck[q]) - lo[2], int(ck[q]) + hi[2] + 1):
            for lj in range(int(cj[q]) - lo[1], int(cj[q]) + hi[1] + 1):
                for li in range(int(ci[q]) - lo[0], int(ci[q]) + hi[0] + 1):
                    for j in members.get((li, lj, lk), ()):
                        if j <= q:
                            continue
                        sr = frlm * (radius[q] + radius[j])
                        d = pos[:, q] - pos[:, j]
                        if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > sr * sr:
                            continue
                        nbrs.append(j + 1)
        index[q] = len(nbrs) + 1
    return index, np.array(nbrs, dtype=np.int64)


def brute_pairs(p, pos, radius):
    """Every pair (i < j, 1-based) within the cut-off, no grid."""
    n = pos.shape[1]
    out = set()
    for q in range(n):
        d = pos[:, q + 1:] - pos[:, q:q + 1]
        ds = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        sr = p.factor_rlm * (radius[q] + radius[q + 1:])
        for j in np.nonzero(ds <= sr * sr)[0]:
            out.add((q + 1, int(q + 2 + j)))
    return out


def as_set(index, nbrs):
    out, start = set(), 1
    for i, end in enumerate(index, start=1):
        for c in range(start, end):
            out.add((i, int(nbrs[c - 1])))
        start = end
    return out


# beds in the two regimes where the reference's stencil misses pairs
REGIMES = {"coarse-2x2x2": (5, 17, 5),     # DES cell 2.0 mm: 2.4 r_max < w < 4.8 r_max
           "narrow-27": (10, 35, 10)}      # DES cell 1.0 mm: w < 2.4 r_max


def regime_bed(name, n=2000, seed=3):
    from mfixdem import workloads
    p, parts, _ = workloads.settling_bed(n=n, seed=seed, box=(0.01, 0.035, 0.01))
    p, parts = workloads.jostled(p, parts, seed=seed + 1)
    p.desgridsearch_imax, p.desgridsearch_jmax, p.desgridsearch_kmax = REGIMES[name]
    return p, parts

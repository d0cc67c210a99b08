"""Pins the oracle's neighbour searches against each other and brute force.

NSQUARE (nsquare.f:15-142) and the DES-grid search (desgrid_mod.f:1002-1299)
are two different algorithms for the same pair set whenever the grid
stencil is complete (DES cell >= 2 x the cut-off): on random positions, where
no pair sits exactly at the cut-off, the strict (NSQUARE) and non-strict
(grid) tests agree, so both must give exactly the set a vectorised brute
force gives, with rows in ascending partner id for NSQUARE.
"""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem


def system(search, seed=3, n=2000, jostle=True):
    p, parts, _ = workloads.settling_bed(n=n, seed=seed, box=(0.01, 0.035, 0.01))
    if jostle:
        p, parts = workloads.jostled(p, parts, seed=seed + 1)
    p.des_neighbor_search = search
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    return p, parts, o


def brute_pairs(p, parts):
    x = parts.pos.T
    r = parts.radius
    d = x[:, None, :] - x[None, :, :]
    ds = np.einsum("ijk,ijk->ij", d, d)
    cut = (p.factor_rlm * (r[:, None] + r[None, :])) ** 2
    i, j = np.nonzero(np.triu(ds < cut, 1))
    return set(zip((i + 1).tolist(), (j + 1).tolist()))


def csr_pairs(idx, nb):
    out, start = [], 1
    for q, end in enumerate(idx):
        row = nb[start - 1:end - 1] if end >= start else []
        out += [(q + 1, int(v)) for v in row]
        start = max(end, start) if end else start
    return out


@pytest.mark.parametrize("seed", [3, 7])
def test_nsquare_equals_brute_force_and_grid(seed):
    p, parts, og = system(abi.NSEARCH_GRID, seed=seed)
    _, _, on = system(abi.NSEARCH_NSQUARE, seed=seed)
    want = brute_pairs(p, parts)
    assert len(want) > 500
    gi, gn, _ = og.get_pairs()
    ni, nn, _ = on.get_pairs()
    gset, nlist = set(csr_pairs(gi, gn)), csr_pairs(ni, nn)
    assert set(nlist) == want
    assert gset == want
    # NSQUARE rows: ascending partner id
    for q in range(parts.n):
        row = [b for a, b in nlist if a == q + 1]
        assert row == sorted(row)


def test_nsquare_trajectory_matches_grid():
    """Same pair set, same row order once re-sorted: the forces differ only
    in the order in which a particle's own row is summed (stencil order vs
    id order), so 40 steps agree to round-off."""
    pg, parts, og = system(abi.NSEARCH_GRID)
    _, _, on = system(abi.NSEARCH_NSQUARE)
    og.run(1, 40)
    on.run(1, 40)
    a, b = og.get_particles(), on.get_particles()
    assert np.array_equal(a[3], b[3])
    assert np.max(np.abs(a[0] - b[0])) <= 1e-13
    assert np.max(np.abs(a[1] - b[1])) <= 1e-9

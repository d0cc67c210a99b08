"""Pins the oracle's DES-grid search (DESGRID_PIC, DESGRID_NEIGH_BUILD;
desgrid_mod.f:877-1299) where the reference's stencil is incomplete, against
the Python restatement in stencil_ref.py:

* DG_PIJK equals the closed form floor((x - dg_xstart) dg_dxinv) + 2,
  clamped, linearised J fastest; dg_xstart and dg_dxinv equal the closed
  form bit for bit;
* the neighbour list (set and row order) equals the pairs the reference
  keeps, which in these beds is a strict subset of the pairs within the
  cut-off (so the test would notice a search that found them all).
"""
import numpy as np
import pytest

import _paths  # noqa: F401
import stencil_ref as sr
from oracle_py import OracleSystem


@pytest.mark.parametrize("regime", list(sr.REGIMES))
def test_des_cells_closed_form(regime):
    p, parts = sr.regime_bed(regime)
    o = OracleSystem(p)
    d = o.derived()
    g = sr.des_grid(p)
    assert (d.dg_xstart, d.dg_ystart, d.dg_zstart) == (g[0][0], g[1][0], g[2][0])
    assert (d.dg_dxinv, d.dg_dyinv, d.dg_dzinv) == (g[0][1], g[1][1], g[2][1])
    assert (d.dg_imax, d.dg_jmax, d.dg_kmax) == (g[0][2], g[1][2], g[2][2])
    o.set_particles(parts)
    o.init()
    dg, _ = o.get_cells()
    _, want = sr.des_cells(p, parts.pos)
    assert np.array_equal(dg, want)
    o.close()


@pytest.mark.parametrize("regime", list(sr.REGIMES))
def test_incomplete_stencil_pairs(regime):
    p, parts = sr.regime_bed(regime)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    idx, nb, _ = o.get_pairs()
    o.close()
    want_idx, want_nb = sr.reference_pairs(p, parts.pos, parts.radius)
    assert np.array_equal(idx, want_idx), "NEIGHBOR_INDEX"
    assert np.array_equal(nb, want_nb), "NEIGHBORS (set or row order)"
    kept = sr.as_set(want_idx, want_nb)
    brute = sr.brute_pairs(p, parts.pos, parts.radius)
    assert kept < brute, "the bed must exercise the incomplete stencil"
    assert len(brute - kept) >= 10

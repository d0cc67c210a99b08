"""GPU parity on non-uniform fluid grids (DX(I), DY(J), DZ(K) through
mfixdem_params.dx/dy/dz): the HIP library against the oracle, whose
stretched-grid routines are pinned in test_oracle_stretched.py.

Bars as on the uniform grid: fluid and DES cells bit for bit; P_FORCE from
CALC_PG_GRAD bit for bit; deposition at set-up within 1e-14 (fixed point
against serial sums); after coupled steps positions within 1e-12 m (pow
ulps, deposition rounding); the gas-side drag within 1e-13 at set-up with
the reference-order deposition."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem
from oracle_py import OracleSystem
from test_oracle_stretched import stretch

pytestmark = pytest.mark.gpu


def fluidised(n=4000, seed=2):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=0.02, ly=0.04, lz=0.02,
                                            cells=(10, 20, 10), dg=(0, 0, 0))
    p, parts = workloads.jostled(p, parts, seed=seed + 1, v0=0.2)
    p.coupled = 1
    return stretch(p), parts, gas


def start(p, parts, gas, p_g=None, debug=0):
    g, o = DemSystem(p), OracleSystem(p)
    if debug:
        g.set_debug(debug)
        o.set_debug(debug)
    for s in (g, o):
        s.set_particles(parts)
        s.set_gas(gas)
        if p_g is not None:
            s.set_gas_pressure(p_g)
        s.init()
    return g, o


@pytest.mark.parametrize("scheme,interp_on,width", [
    (abi.INTERP_GARG_2012, 1, None), (abi.INTERP_LINEAR_HAT, 1, None),
    (abi.INTERP_SQUARE_DPVM, 1, 1.2e-3)], ids=["garg", "linear_hat", "square_dpvm"])
def test_coupled_run_on_a_stretched_grid(scheme, interp_on, width):
    p, parts, gas = fluidised()
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    rng = np.random.default_rng(11)
    nc = workloads.n_cells(p)
    p_g = 1.0e5 - 2500.0 * rng.uniform(0.0, 0.04, nc) + rng.normal(0.0, 5.0, nc)
    g, o = start(p, parts, gas, p_g)
    assert np.array_equal(g.get_p_force(), o.get_p_force())
    gd, gpijk = g.get_cells()
    od, opijk = o.get_cells()
    assert np.array_equal(gd, od) and np.array_equal(gpijk, opijk)
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-14
    assert np.max(np.abs(gr - orr)) <= 1e-14 * np.max(np.abs(orr))
    g.run(1, 40)
    o.run(1, 40)
    gp, gv, _, gs = g.get_particles()
    op, ov, _, os_ = o.get_particles()
    assert np.array_equal(gs, os_)
    assert np.max(np.abs(gp - op)) <= 1e-12
    assert np.max(np.abs(gv - ov)) <= 1e-9 * max(1.0, np.max(np.abs(ov)))
    assert np.array_equal(g.get_cells()[1], o.get_cells()[1])
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-12
    g.close()
    o.close()


def test_gas_side_drag_on_a_stretched_grid():
    p, parts, gas = fluidised(n=3000)
    g, o = start(p, parts, gas, debug=abi.DEBUG_REFERENCE_DEPOSITION)
    assert np.array_equal(g.get_gas()[0], o.get_gas()[0])
    fg, ag, bg = g.drag_gs_gas()
    fo, ao, bo = o.drag_gs_gas()
    assert fo.max() > 0.0 and ao.max() > 0.0
    assert np.max(np.abs(fg - fo)) <= 1e-13 * fo.max()
    assert np.max(np.abs(ag - ao)) <= 1e-13 * ao.max()
    assert np.max(np.abs(bg - bo)) <= 1e-13 * np.max(np.abs(bo))
    g.close()
    o.close()


def test_bad_cell_sizes_are_refused():
    p, parts, gas = fluidised(n=500)
    dx = np.full(p.imax, p.xlength / p.imax)
    dx[3] *= 1.5                                   # no longer sums to XLENGTH
    keep = p.xlength
    p.set_cell_sizes(dx=dx)
    p.xlength = keep
    with pytest.raises(abi.MfixDemError):
        DemSystem(p)

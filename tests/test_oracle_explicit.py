"""Explicit coupling (DES_EXPLICITLY_COUPLED) in the oracle, pinned by a
closed form on one sphere.

The fluid-step sequence is CALC_DES_2FLUID (PARTICLES_IN_CELL, weights,
mean fields, DRAG_GS_GAS1, which keeps F_GP), the gas solve (here: a new
frozen field), then DES_TIME_INIT's DRAG_GS_DES1 + CALC_PG_GRAD, which form
DRAG_FC = F_GP_old * (u_new - v) + P_FORCE * PVOL; every DEM step then adds
DRAG_FC to FC (calc_drag_des.f:34-43) and does not rebin.  A sphere far
from every wall and every other sphere feels FC = DRAG_FC exactly, so FC of
its first step must equal the closed form built from the textbook Wen-Yu
law (F_GP with the OLD gas speed and the cell voidage) and the NEW gas
speed.  Using the new field for F_GP, or forgetting the pressure term, moves
FC by far more than the 1e-12 tolerance (the pow of Re**0.687 and
eps**-2.65 is libm's, as in the reference)."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem


def _one_sphere(scheme):
    width = 3.0e-3 if scheme in (abi.INTERP_SQUARE_DPVM, abi.INTERP_GAUSS_DPVM) else None
    p, parts, gas, p_g, _ = workloads.dem06(scheme=scheme, width=width, u_gas=0.4)
    p.des_explicitly_coupled = 1
    p.dtsolid = 1.0e-6
    return p, parts, gas, p_g


def wen_yu_fgp(epg, rho_g, mu, vrel, d, pvol):
    """DRAG_WEN_YU (drag_gs.f:703-758) times PVOL (drag_gp_des.f:249-253)."""
    ropg = rho_g * epg
    re = d * vrel * ropg / mu
    cd = 24.0 / (re + 1e-15) * (1.0 + 0.15 * re ** 0.687) if re <= 1000.0 else 0.44
    beta = 0.75 * cd * vrel * ropg * epg ** -2.65 / d if re > 0.0 else 0.0
    return beta * pvol


@pytest.mark.parametrize("scheme", [abi.INTERP_NONE, abi.INTERP_SQUARE_DPVM],
                         ids=["none", "square_dpvm"])
def test_drag_fc_is_old_fgp_times_new_slip_plus_pressure(scheme):
    p, parts, gas, p_g = _one_sphere(scheme)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.set_gas_pressure(p_g)
    o.init()
    # CALC_DES_2FLUID with the old field (v_g = 0.4), F_GP kept
    o.particles_in_cell()
    o.calc_interp_weights()
    o.comp_mean_fields()
    o.drag_gs_gas()
    ep_g = o.get_gas()[0]
    ugc_old = o.ugc()
    # the gas solve moves the field: v_g = 0.7
    gas2 = workloads.Gas(gas.u_g, np.full_like(gas.v_g, 0.7), gas.w_g, gas.p_force)
    o.set_gas(gas2)
    o.set_gas_pressure(p_g)
    ugc_new = o.ugc()
    pfc = o.get_p_force()
    o.des_time_init()
    o.set_debug(abi.DEBUG_STORE_FORCES)
    o.step(2)
    fc = o.get_forces()[0][:, 0]
    # closed form for a sphere at rest: fields at the sphere from its filter
    # weights (or its cell), Wen-Yu F_GP with the old slip, the new slip
    r = parts.radius[0]
    pvol = (4.0 / 3.0) * np.pi * (r * r * r)
    cell = o.get_cells()[1][3, 0] - 1
    cells, w = o.filter(0)
    if scheme == abi.INTERP_NONE:
        cells, w = np.array([cell]), np.array([1.0])
    lepg = float(np.sum(ep_g[cells] * w))
    v_old = (ugc_old[cells] * w[:, None]).sum(0)
    v_new = (ugc_new[cells] * w[:, None]).sum(0)
    lpf = (pfc[cells] * w[:, None]).sum(0)
    fgp_old = wen_yu_fgp(lepg, p.ro_g0, p.mu_g0, float(np.linalg.norm(v_old)), 2.0 * r, pvol)
    expect = fgp_old * v_new + lpf * pvol
    assert fgp_old > 0.0 and abs(lpf[1]) > 0.0
    assert np.max(np.abs(fc - expect)) <= 1e-12 * np.abs(expect).max()
    # the new field's F_GP, or no pressure term, would be far off
    fgp_new = wen_yu_fgp(lepg, p.ro_g0, p.mu_g0, float(np.linalg.norm(v_new)), 2.0 * r, pvol)
    assert np.max(np.abs(fgp_new * v_new + lpf * pvol - expect)) > 1e-6 * np.abs(expect).max()
    assert np.max(np.abs(fgp_old * v_new - expect)) > 1e-6 * np.abs(expect).max()
    o.close()


def test_explicit_steps_do_not_rebin():
    """PARTICLES_IN_CELL runs in CALC_DES_2FLUID only: a sphere that falls
    through a fluid-cell face during the substeps keeps its old PIJK."""
    p, parts, gas, p_g = _one_sphere(abi.INTERP_NONE)
    # start just above a J face so that free fall crosses it
    dy = p.ylength / p.jmax
    y0 = 18 * dy + 2.0e-6
    parts = workloads.make_particles(np.array([[0.005], [y0], [0.005]]), parts.radius,
                                     parts.ro_sol, vel=np.array([[0.0], [-1.0], [0.0]]))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.set_gas_pressure(p_g)
    o.init()
    j0 = o.get_cells()[1][1, 0]
    o.run(1, 10)   # 10 us at 1 m/s: 10 um down, across the face
    assert o.get_particles()[0][1, 0] < 18 * dy
    assert o.get_cells()[1][1, 0] == j0
    o.particles_in_cell()
    assert o.get_cells()[1][1, 0] == j0 - 1
    o.close()

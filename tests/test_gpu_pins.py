"""GPU side of the closed-form pins: the HIP library through the C ABI
against the same closed forms that pin the oracle (linear gas field for the
filter-scheme drag, incomplete DES-grid stencils) and against the oracle.

Bar: neighbour lists and DES cells bit-exact; trajectories of the stencil
beds bit-exact (no libm pow on that path); drag forces within 1e-12 of the
largest force of the closed form and 1e-13 of the oracle's (the drag law
calls pow, whose last bit may differ between ocml and glibc).
"""
import numpy as np
import pytest

import _paths  # noqa: F401
import stencil_ref as sr
from mfixdem import abi, workloads
from mfixdem.system import DemSystem
from oracle_py import OracleSystem
from linear_gas import LINEAR_CASES, LINEAR_IDS, linear_case, linear_expectation

pytestmark = pytest.mark.gpu


def _run_forces(cls, p, parts, gas):
    s = cls(p)
    s.set_debug(abi.DEBUG_STORE_FORCES)
    s.set_particles(parts)
    s.set_gas(gas)
    s.init()
    s.step(1)
    fc, _ = s.get_forces()
    s.close()
    return fc


@pytest.mark.parametrize("drag", [abi.DRAG_GIDASPOW, abi.DRAG_WEN_YU], ids=["gidaspow", "wenyu"])
@pytest.mark.parametrize("scheme,interp_on,width", LINEAR_CASES, ids=LINEAR_IDS)
def test_des1_drag_linear_field_on_gpu(scheme, interp_on, width, drag):
    p, parts, gas = linear_case(scheme, interp_on, width, drag)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    ep, _ = o.get_gas()
    for match in (False, True):
        vel, want = linear_expectation(p, parts, ep, o.filter, match_gas_velocity=match)
        q = parts if not match else workloads.make_particles(parts.pos, parts.radius,
                                                             parts.ro_sol, vel=vel)
        fg = _run_forces(DemSystem, p, q, gas)
        fo = _run_forces(OracleSystem, p, q, gas)
        scale = np.max(np.abs(want))
        assert np.max(np.abs(fg - want)) <= 1e-12 * scale, (match, "closed form")
        assert np.max(np.abs(fg - fo)) <= 1e-13 * scale, (match, "oracle")
    o.close()


@pytest.mark.parametrize("regime", list(sr.REGIMES))
def test_incomplete_stencil_on_gpu(regime):
    """DG_PIJK against the closed form; NEIGHBOR_INDEX / NEIGHBORS against the
    reference rule (a strict subset of the pairs within the cut-off) and the
    oracle; then 60 steps (two rebuilds) bit-exact with the oracle."""
    p, parts = sr.regime_bed(regime)
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        s.init()
    dg, _ = g.get_cells()
    _, want_dg = sr.des_cells(p, parts.pos)
    assert np.array_equal(dg, want_dg)
    want_idx, want_nb = sr.reference_pairs(p, parts.pos, parts.radius)
    gi, gn, _ = g.get_pairs()
    assert np.array_equal(gi, want_idx) and np.array_equal(gn, want_nb)
    assert sr.as_set(want_idx, want_nb) < sr.brute_pairs(p, parts.pos, parts.radius)
    for s in (g, o):
        s.run(1, 60)
    for a, b in zip(g.get_particles()[:3], o.get_particles()[:3]):
        assert np.array_equal(a, b)
    gi, gn, gp = g.get_pairs()
    oi, on, op = o.get_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gn, on) and np.array_equal(gp, op)
    g.close()
    o.close()


def _fluidised(n=4000, seed=2):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=0.02, ly=0.04, lz=0.02,
                                            cells=(10, 20, 10), dg=(0, 0, 0))
    p, parts = workloads.jostled(p, parts, seed=seed + 1, v0=0.2)
    return p, parts, gas


def test_state_calls_are_ordered_after_queued_steps():
    """Calls on a handle are ordered on its stream: mfixdem_set_gas and
    mfixdem_set_particles issued right after mfixdem_run (which returns with
    graph-replayed steps still queued) give the same result, bit for bit, as
    the same calls after an explicit mfixdem_synchronize; the CPU oracle
    agrees to the drag law's tolerance."""
    p, parts, gas = _fluidised()
    gas2 = workloads.Gas(0.5 * gas.u_g, 1.3 * gas.v_g, -gas.w_g, 0.9 * gas.p_force)
    out = []
    for wait in (False, True):
        g = DemSystem(p)
        g.set_particles(parts)
        g.set_gas(gas)
        g.init()
        g.run(1, 40)
        if wait:
            g.synchronize()
        g.set_gas(gas2)
        g.run(41, 40)
        out.append(g.get_particles()[:3])
        g.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    o.run(1, 40)
    o.set_gas(gas2)
    o.run(41, 40)
    op = o.get_particles()
    o.close()
    assert np.max(np.abs(out[0][0] - op[0])) <= 1e-12 * np.max(np.abs(op[0]))
    assert np.max(np.abs(out[0][1] - op[1])) <= 1e-10 * np.max(np.abs(op[1]))
    # a new particle set right after a run replaces the state completely
    g = DemSystem(p)
    g.set_particles(parts)
    g.set_gas(gas)
    g.init()
    g.run(1, 40)
    g.set_particles(parts)
    g.init()
    g.run(1, 20)
    fresh = DemSystem(p)
    fresh.set_particles(parts)
    fresh.set_gas(gas)
    fresh.init()
    fresh.run(1, 20)
    for a, b in zip(g.get_particles()[:3], fresh.get_particles()[:3]):
        assert np.array_equal(a, b)
    g.close()
    fresh.close()

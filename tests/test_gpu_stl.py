"""GPU parity of STL walls and moving walls: the HIP library (C ABI) against
the CPU oracle on the same seeded inputs, bit for bit in FP64.

* A square funnel of eight STL triangles (two per face, so each face has a
  coplanar diagonal whose edge contacts are ignored) inside the default box
  walls, with jostled spheres settling into it: positions, velocities,
  spins, the pair list and its tangential history equal the oracle's after
  600 steps, for both integrators.  The same run without the funnel moves
  hundreds of spheres elsewhere, so the facets were hit.
* DEM01 with its floor given as two STL triangles and no box walls matches
  the reference's DEM01 output (POST_POS, 8 decimals) on the GPU too.
* Moving walls: a floor sliding along x and z and a side wall turning about
  a vertical axis (BC_STL wall velocity and omega), changed part-way through
  a graph-replayed run; every state equals the oracle's.
* Mixed precision: the FP32 wall arithmetic against FP64 on the funnel.
"""
import json
import os

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem
from oracle_py import OracleSystem
from test_oracle_stl import sphere_in_cone

pytestmark = pytest.mark.gpu

BOX = (0.01, 0.035, 0.01)


def funnel_case(integrator=abi.EULER, precision=abi.FP64):
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=BOX)
    p.integrator = integrator
    p.precision = precision
    p, parts = workloads.jostled(p, parts, seed=4)
    v, n, planes = workloads.funnel_facets(BOX, 0.004, 0.012, 0.0015)
    keep = workloads.inside_funnel(parts.pos, parts.radius, planes, margin=0.9)
    sub = workloads.make_particles(parts.pos[:, keep], parts.radius[keep], parts.ro_sol[keep],
                                   parts.vel[:, keep], parts.omega[:, keep])
    p.max_pip = sub.n
    return p, sub, v, n


def start(cls, p, parts, v=None, n=None, default_walls=True, debug=0):
    s = cls(p)
    if debug:
        s.set_debug(debug)
    s.set_particles(parts)
    if v is not None:
        s.set_facets(v, n, default_walls)
    s.init()
    return s


def same_state(g, o):
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    gi, gn, gp = g.get_pairs()
    oi, on, op = o.get_pairs()
    assert np.array_equal(gi, oi) and np.array_equal(gn, on)
    assert np.array_equal(gp, op)


@pytest.mark.parametrize("integrator", [abi.EULER, abi.ADAMS_BASHFORTH], ids=["euler", "ab2"])
def test_funnel_stl_bit_exact(integrator):
    p, parts, v, n = funnel_case(integrator)
    g = start(DemSystem, p, parts, v, n)
    o = start(OracleSystem, p, parts, v, n)
    g.run(1, 300)
    o.run(1, 300)
    same_state(g, o)
    for nn in range(301, 601):   # eager steps as well as graph replay
        if nn < 320:
            g.step(nn)
        else:
            g.run(nn, 1)
        o.step(nn)
    same_state(g, o)
    assert o.status().n_contacts > 0
    plain = start(OracleSystem, p, parts)
    plain.run(1, 600)
    moved = np.abs(plain.get_particles()[0] - o.get_particles()[0]).max(axis=0) > 1e-9
    assert moved.sum() > 100
    for s in (g, o, plain):
        s.close()


def test_dem01_on_an_stl_floor_on_gpu():
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dem01_post.json")))
    run = next(r for r in gold["runs"] if r["scheme"] == "EULER")
    p, parts, _ = workloads.dem01(kn=run["kn"], en=run["en"], integrator=abi.EULER)
    a = np.array([-1.0, 0.0, -0.7])
    b, c, d = np.array([2.0, 0.0, -0.7]), np.array([2.0, 0.0, 2.3]), np.array([-1.0, 0.0, 2.3])
    v = np.array([np.stack(t, axis=1) for t in ((a, c, b), (a, d, c))])
    n = np.array([[0.0, 1.0, 0.0]] * 2)
    g = start(DemSystem, p, parts, v, n, default_walls=False)
    o = start(OracleSystem, p, parts, v, n, default_walls=False)
    dt = g.derived().dtsolid
    rows = np.array(run["rows"])
    steps = np.rint(rows[:, 0] / dt).astype(int)
    ys = {0: 0.5}
    for nn in range(1, steps.max() + 1):
        g.step(nn)
        o.step(nn)
        if nn in set(steps.tolist()):
            ys[nn] = g.get_particles()[0][1, 0]
    assert np.max(np.abs(np.array([ys[s] for s in steps]) - rows[:, 1])) <= 1.0e-8
    same_state(g, o)
    g.close()
    o.close()


def test_moving_walls_bit_exact():
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=BOX)
    p, parts = workloads.jostled(p, parts, seed=4)
    g = start(DemSystem, p, parts)
    o = start(OracleSystem, p, parts)
    centre = (0.5 * BOX[0], 0.5 * BOX[1], 0.5 * BOX[2])
    for s in (g, o):
        s.set_wall_motion(2, 1, (0.2, 0.0, -0.1))                    # floor (S)
        s.set_wall_motion(1, 1, (0.0, 0.0, 0.0), (0.0, 40.0, 0.0), centre)   # E wall turning
    g.run(1, 200)
    o.run(1, 200)
    same_state(g, o)
    for s in (g, o):
        s.set_wall_motion(2, 1, (-0.3, 0.0, 0.0))
        s.set_wall_motion(0, 1, (0.0, 0.05, 0.0), (0.0, 0.0, 0.0), centre)   # W wall sliding
    g.run(201, 200)
    o.run(201, 200)
    same_state(g, o)
    still = start(OracleSystem, p, parts)
    still.run(1, 400)
    assert not np.array_equal(still.get_particles()[1], o.get_particles()[1])
    for s in (g, o, still):
        s.close()


def test_funnel_mixed_precision_wall_forces():
    """FP32 wall and pair arithmetic on the funnel: max force error against
    FP64 on the same state <= 1e-4 of max |F|."""
    p64, parts, v, n = funnel_case()
    pmx, _, _, _ = funnel_case(precision=abi.MIXED)
    a = start(DemSystem, p64, parts, v, n, debug=abi.DEBUG_STORE_FORCES)
    b = start(DemSystem, pmx, parts, v, n, debug=abi.DEBUG_STORE_FORCES)
    a.step(1)
    b.step(1)
    fa, _ = a.get_forces()
    fb, _ = b.get_forces()
    assert np.max(np.abs(fa - fb)) / np.max(np.abs(fa)) <= 1e-4
    a.close()
    b.close()


def test_set_facets_argument_errors():
    p, parts, v, n = funnel_case()
    g = DemSystem(p)
    g.set_particles(parts)
    with pytest.raises(abi.MfixDemError):
        g.set_wall_motion(0, 100, (0.0, 0.0, 0.0))    # beyond the 6 box facets
    g.set_facets(v, n, True)
    g.set_wall_motion(0, 14, (0.0, 0.0, 0.0))        # 8 funnel + 6 box facets
    g.init()
    with pytest.raises(abi.MfixDemError):
        g.set_facets(v, n, True)                      # after init
    g.close()


def test_wall_contact_capacity_matches_oracle():
    """Eight facets touching one sphere run bit-exact; nine are
    MFIXDEM_ERR_CAPACITY (reported by the next rebuild or synchronize)."""
    p, parts, v, n = sphere_in_cone(8)
    g = start(DemSystem, p, parts, v, n)
    o = start(OracleSystem, p, parts, v, n)
    g.run(1, 60)
    o.run(1, 60)
    g.synchronize()
    same_state(g, o)
    g.close()
    o.close()
    p, parts, v, n = sphere_in_cone(9)
    g = start(DemSystem, p, parts, v, n)
    with pytest.raises(abi.MfixDemError) as e:
        g.step(1)
        g.synchronize()
    assert e.value.code == abi.ERRORS_BY_NAME["CAPACITY"]
    g.close()


def test_both_wall_paths_agree():
    """Walls of at most eight facets run with slot k = facet k and the
    facets in the kernel arguments; more facets use the facet table and the
    reference's slot search.  The box walls alone, and the box walls after
    three extra facets far outside the domain (nine facets: the general
    path, same facets per cell), give the same run bit for bit, equal to
    the oracle's."""
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=BOX)
    p, parts = workloads.jostled(p, parts, seed=4)
    # triangles in z = 5 m planes, at y = 5, 6, 7 m
    far = np.array([[[5.0, 6.0, 5.0], [5.0, 5.0, 6.0], [5.0, 5.0, 5.0]]] * 3)
    far[:, 1, :] += np.arange(3)[:, None]
    n = np.array([[0.0, 0.0, 1.0]] * 3)
    small = start(DemSystem, p, parts)
    general = start(DemSystem, p, parts, far, n, default_walls=True)
    o = start(OracleSystem, p, parts, far, n, default_walls=True)
    for s in (small, general, o):
        s.set_wall_motion(2 if s is small else 5, 1, (0.1, 0.0, 0.0))   # the floor slides
        s.run(1, 300)
    same_state(general, o)
    same_state(small, o)
    for s in (small, general, o):
        s.close()

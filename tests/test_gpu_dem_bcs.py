"""GPU parity of the DEM mass inlets and outlets (mfixdem_set_dem_bcs): the
HIP library against the oracle, whose inlet and outlet routines are pinned
in test_oracle_dem_bcs.py, on the same seeded stream.

Bar: bit for bit -- the state, position, velocity and rotation of every
slot at each checkpoint and the neighbour pairs with their row order at
the end -- through seeding, the ENTERING -> NORMAL turn, contacts between
normal and entering or exiting particles, EXITING at both outlet kinds
(with UPDATE_EXIT_DES_VEL), deletion with its forced neighbour search, and
new particles taking the lowest free slots.  Running out of slots fails at
the same step on both."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem
from oracle_py import OracleSystem
from test_gpu_parity import assert_same_pairs

pytestmark = pytest.mark.gpu

D, RHO = 1.0e-3, 2500.0


def params(box, cells, nphases=1, nsearch=25, grav=(-20.0, 0.0, 0.0), max_pip=1500,
           diameters=(D,)):
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = nphases
    p.max_pip = max_pip
    p.xlength, p.ylength, p.zlength = box
    p.imax, p.jmax, p.kmax = cells
    p.kn = p.kn_w = 1.0e3
    p.mew = p.mew_w = 0.3
    for m in range(nphases):
        p.des_en_wall_input[m] = 0.9
        p.d_p0[m] = diameters[m]
        p.ro_s0[m] = RHO
    for k in range(nphases * (nphases + 1) // 2):
        p.des_en_input[k] = 0.9
    p.max_radius = 0.5 * max(diameters[:nphases])
    p.grav[:] = list(grav)
    p.dtsolid = 1.0e-5
    p.neighbor_search_n = nsearch
    return p


def start(p, bcs, parts, holes=None, seed=11):
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        if holes is not None:
            s.set_particle_state(holes)
        s.set_dem_bcs(bcs, seed=seed)
        s.init()
    return g, o


def assert_same(g, o, where):
    for a, b, name in zip(g.get_particles(), o.get_particles(),
                          ("position", "velocity", "omega", "state")):
        assert np.array_equal(a, b), f"{name} differs {where}"


def run_compare(g, o, steps, every):
    nn = 1
    while nn <= steps:
        k = min(every, steps - nn + 1)
        g.run(nn, k)
        o.run(nn, k)
        nn += k
        assert_same(g, o, f"after step {nn - 1}")


def column():
    """Inlet in the bottom face (y = 0) of a 20 x 16 x 20 mm box; the
    particles cross it in about 3200 steps and drift to the x = 0 wall."""
    box = (0.02, 0.016, 0.02)
    p = params(box, (4, 8, 4))
    inlet = abi.dem_bc(abi.BC_MASS_INFLOW, abi.PLANE_N, (0.005, 0.0, 0.005), (0.015, 0.0, 0.015),
                       vel=[(0.0, 0.5, 0.0)], ep_s=[0.3])
    return p, box, inlet


@pytest.mark.parametrize("kind", [abi.BC_MASS_OUTFLOW, abi.BC_OUTFLOW],
                         ids=["mass_outflow", "outflow"])
def test_column_inlet_to_outlet_bit_exact(kind):
    p, box, inlet = column()
    outlet = abi.dem_bc(kind, abi.PLANE_S, (0.0, box[1], 0.0), box, vel=[(0.0, 0.8, 0.0)])
    empty = workloads.make_particles(np.zeros((3, 0)), 0.5 * D, RHO)
    g, o = start(p, [inlet, outlet], empty)
    run_compare(g, o, 4500, 500)
    st = o.get_particles()[3]
    assert (st == 3).any() and (st == 2).any() and (st == 1).any()
    assert o.status().rebuilds > 4500 // p.neighbor_search_n + 50   # deletions forced them
    assert_same_pairs(g, o)
    g.close()
    o.close()


def test_column_hertz_bit_exact():
    """The column with Hertz contacts (E = 1e7, nu = 0.3, normal and
    tangential restitution 0.9, walls alike): the contact law, its derived
    collision time and the wall contact of the particles at the bottom all
    follow the inlet's seeding bit for bit."""
    p, box, inlet = column()
    p.coll_model = abi.HERTZ
    p.e_young[0], p.v_poisson[0] = 1.0e7, 0.3
    p.ew_young, p.vw_poisson = 1.0e7, 0.3
    p.des_et_input[0] = 0.9
    p.des_et_wall_input[0] = 0.9
    outlet = abi.dem_bc(abi.BC_MASS_OUTFLOW, abi.PLANE_S, (0.0, box[1], 0.0), box,
                        vel=[(0.0, 0.8, 0.0)])
    empty = workloads.make_particles(np.zeros((3, 0)), 0.5 * D, RHO)
    g, o = start(p, [inlet, outlet], empty)
    run_compare(g, o, 3000, 500)
    st = o.get_particles()[3]
    assert (st == 1).any() and (st == 2).any()
    assert_same_pairs(g, o)
    g.close()
    o.close()


def test_side_inlet_two_phases_into_a_bed_bit_exact():
    """Two phases (by mass flow) through an inlet in the x = 0 face into a
    bed of normal particles with holes among them, out through an outflow
    in the x = 20 mm face; gravity along -y."""
    box = (0.02, 0.012, 0.012)
    p = params(box, (10, 6, 6), nphases=2, nsearch=10, grav=(0.0, -5.0, 0.0),
               diameters=(1.0e-3, 0.8e-3))
    rng = np.random.default_rng(5)
    n0 = 400
    pos = np.stack([rng.uniform(0.004, 0.016, n0), rng.uniform(0.001, 0.011, n0),
                    rng.uniform(0.001, 0.011, n0)])
    keep = np.ones(n0, dtype=bool)
    for q in range(n0):                       # no initial overlaps
        d = np.linalg.norm(pos[:, :q] - pos[:, q:q + 1], axis=0)
        keep[q] = not np.any(d[keep[:q]] < 1.05e-3)
    pos = pos[:, keep]
    n = pos.shape[1]
    phase = rng.integers(1, 3, n).astype(np.int32)
    radius = np.where(phase == 1, 0.5e-3, 0.4e-3)
    parts = workloads.make_particles(pos, radius, RHO, vel=rng.normal(0.0, 0.05, (3, n)),
                                     phase=phase)
    holes = np.ones(n, dtype=np.int8)
    holes[rng.choice(n, 40, replace=False)] = 0
    inlet = abi.dem_bc(abi.BC_MASS_INFLOW, abi.PLANE_E, (0.0, 0.002, 0.002), (0.0, 0.010, 0.010),
                       vel=[(0.6, 0.0, 0.0), (0.6, 0.0, 0.0)], ep_s=[0.1, 0.1],
                       massflow=[2.0e-3, 1.0e-3])
    outlet = abi.dem_bc(abi.BC_OUTFLOW, abi.PLANE_W, (box[0], 0.0, 0.0), box)
    g, o = start(p, [inlet, outlet], parts, holes)
    run_compare(g, o, 5000, 500)
    assert o.status().rebuilds > 5000 // p.neighbor_search_n + 1   # deletions forced some
    assert_same_pairs(g, o)
    g.close()
    o.close()


def test_running_out_of_slots_fails_at_the_same_step():
    p, box, inlet = column()
    p.max_pip = 60
    empty = workloads.make_particles(np.zeros((3, 0)), 0.5 * D, RHO)
    g, o = start(p, [inlet], empty)
    failed = {}
    for name, s in (("gpu", g), ("oracle", o)):
        for nn in range(1, 2000):
            try:
                s.step(nn)
            except abi.MfixDemError as e:
                failed[name] = (nn, e.code)
                break
    assert "gpu" in failed and failed["gpu"] == failed["oracle"]
    assert failed["gpu"][1] == abi.ERRORS_BY_NAME["CAPACITY"]
    g.close()
    o.close()


def test_refused_combinations():
    p, box, inlet = column()
    p.integrator = abi.ADAMS_BASHFORTH
    g = DemSystem(p)
    with pytest.raises(abi.MfixDemError):
        g.set_dem_bcs([inlet])
    g.close()


def test_hopper_inlet_funnel_outlet_bit_exact():
    """A hopper: an inlet in the top face feeds spheres that fall into an
    STL funnel (eight triangles inside the box walls), pile on its faces,
    drain through its throat and leave through a MASS_OUTFLOW outlet in
    the bottom face; only normal spheres meet the walls."""
    box = (0.01, 0.035, 0.01)
    p = params(box, (2, 7, 2), nsearch=20, grav=(0.0, -9.81, 0.0), max_pip=1200)
    v, n, _ = workloads.funnel_facets(box, 0.004, 0.012, 0.0015)
    inlet = abi.dem_bc(abi.BC_MASS_INFLOW, abi.PLANE_S, (0.002, box[1], 0.002),
                       (0.008, box[1], 0.008), vel=[(0.0, -0.3, 0.0)], ep_s=[0.2])
    outlet = abi.dem_bc(abi.BC_MASS_OUTFLOW, abi.PLANE_N, (0.0, 0.0, 0.0), (box[0], 0.0, box[2]),
                        vel=[(0.0, -0.5, 0.0)])
    empty = workloads.make_particles(np.zeros((3, 0)), 0.5 * D, RHO)
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(empty)
        s.set_facets(v, n, True)
        s.set_dem_bcs([inlet, outlet], seed=5)
        s.init()
    run_compare(g, o, 12000, 1500)
    assert o.status().rebuilds > 12000 // p.neighbor_search_n + 20   # deletions at the outlet
    assert_same_pairs(g, o)
    g.close()
    o.close()

"""DEM mass inlets and outlets in the oracle (MASS_INFLOW_DEM,
MASS_OUTFLOW_DEM, SET_BC_DEM_MI / LAYOUT_MI_DEM, SET_BC_DEM_MO), pinned by
closed forms and invariants of the reference's rules:

* the inlet feeds BC_EP_s x BC_V_s x area of solids volume per second (or
  BC_MASSFLOW_s): after a long run the count of seeded spheres equals the
  closed-form rate to within two spheres (the first seeding waits one
  period, the last period may be partial), for one phase and for two
  (each phase by its own rate);
* every seeded sphere starts outside the inlet plane by the largest
  diameter, inside its window of the inlet rectangle, at the inlet
  velocity, ENTERING; it turns NORMAL only once it is fully inside and
  has been binned (a neighbour search since it entered);
* spheres that reach an outlet turn EXITING with the outlet's velocity
  (MASS_OUTFLOW) or only their normal velocity (another outflow) and are
  deleted when their centre crosses the plane: seeded = existing + deleted;
* new spheres take the lowest NONEXISTENT slots (slot reuse).
"""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem

BOX = (0.02, 0.04, 0.02)
D, RHO, V_IN, EPS = 1.0e-3, 2500.0, 0.5, 0.3
INLET_LO, INLET_HI = (0.005, 0.0, 0.005), (0.015, 0.0, 0.015)


def params(max_pip=3000, nphases=1, nsearch=25):
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = nphases
    p.max_pip = max_pip
    p.xlength, p.ylength, p.zlength = BOX
    p.imax, p.jmax, p.kmax = 4, 8, 4
    p.kn = p.kn_w = 1.0e3
    p.mew = p.mew_w = 0.3
    for m in range(nphases):
        p.des_en_wall_input[m] = 0.9
        p.d_p0[m] = D
        p.ro_s0[m] = RHO
    for k in range(nphases * (nphases + 1) // 2):
        p.des_en_input[k] = 0.9
    p.max_radius = 0.5 * D
    p.grav[:] = [0.0, 0.0, 0.0]
    p.dtsolid = 1.0e-5
    p.neighbor_search_n = nsearch
    return p


def start(p, bcs, seed=7, parts=None):
    o = OracleSystem(p)
    o.set_particles(parts if parts is not None else
                    workloads.make_particles(np.zeros((3, 0)), 0.5 * D, RHO))
    o.set_dem_bcs(bcs, seed=seed)
    o.init()
    return o


def inlet(vel=V_IN, ep=(EPS,), massflow=None, t1=1.0e30):
    return abi.dem_bc(abi.BC_MASS_INFLOW, abi.PLANE_N, INLET_LO, INLET_HI,
                      vel=[(0.0, vel, 0.0)] * len(ep), ep_s=list(ep), massflow=massflow, t1=t1)


def outlet(kind=abi.BC_MASS_OUTFLOW, vel=V_IN):
    return abi.dem_bc(kind, abi.PLANE_S, (0.0, BOX[1], 0.0), BOX, vel=[(0.0, vel, 0.0)])


def seeded_rate(ep=EPS, vel=V_IN):
    area = (INLET_HI[0] - INLET_LO[0]) * (INLET_HI[2] - INLET_LO[2])
    return ep * vel * area / (np.pi / 6.0 * D ** 3)      # spheres per second


def test_inflow_rate_closed_form():
    p = params()
    o = start(p, [inlet()])
    steps = 4000
    o.run(1, steps)
    st = o.get_particles()[3]
    want = seeded_rate() * steps * p.dtsolid
    assert abs(int((st != 0).sum()) - want) <= 2.0
    o.close()


def test_inflow_massflow_and_two_phases():
    """BC_MASSFLOW_s sets the rate (kg/s over the sphere mass); two phases
    share the stream in proportion to their rates (a seed's phase is drawn
    against DEM_BC_POLY_LAYOUT).  The phases are told apart by a small
    tangential inlet velocity of the second one."""
    p = params(nphases=2)
    mdot = (2.0e-3, 6.0e-3)
    b = inlet(ep=(0.1, 0.2), massflow=list(mdot))
    b.u_s[1] = 1.0e-3
    o = start(p, [b])
    steps = 6000
    n_phase = [0, 0]
    st = o.get_particles()[3].copy()
    for nn in range(1, steps + 1):
        o.step(nn)
        _, vel, _, st2 = o.get_particles()
        for q in np.flatnonzero((st == 0) & (st2 != 0)):
            n_phase[1 if vel[0, q] != 0.0 else 0] += 1
        st = st2.copy()
    mass = RHO * np.pi / 6.0 * D ** 3
    t = steps * p.dtsolid
    assert abs(sum(n_phase) - sum(mdot) / mass * t) <= 2.0
    assert int((st != 0).sum()) == sum(n_phase)
    assert abs(n_phase[1] / sum(n_phase) - mdot[1] / sum(mdot)) < 0.06
    o.close()


def test_seeds_start_outside_in_their_window_and_enter():
    p = params(nsearch=10)
    o = start(p, [inlet()])
    r = 0.5 * D
    seen_normal = 0
    for nn in range(1, 601):
        before = o.get_particles()[3].copy()
        o.step(nn)
        pos, vel, _, st = o.get_particles()
        new = (before == 0) & (st != 0)
        for q in np.flatnonzero(new):
            assert st[q] == 2
            assert pos[1, q] == -D                    # OFFSET: the plane less MAX_DIA
            assert INLET_LO[0] + r - 1e-12 <= pos[0, q] <= INLET_HI[0] - r + 1e-12
            assert INLET_LO[2] + r - 1e-12 <= pos[2, q] <= INLET_HI[2] - r + 1e-12
            assert np.array_equal(vel[:, q], [0.0, V_IN, 0.0])
        # NORMAL only fully inside; ENTERING ones keep the inlet velocity
        # and turn NORMAL at the first check after a neighbour search
        normal = st == 1
        assert np.all(pos[1, normal] > r)
        entering = st == 2
        assert np.all(vel[1, entering] == V_IN)
        assert np.all(pos[1, entering] <= r + 10 * V_IN * p.dtsolid)
        seen_normal += int(normal.sum())
    assert seen_normal > 0
    o.close()


@pytest.mark.parametrize("kind", [abi.BC_MASS_OUTFLOW, abi.BC_OUTFLOW],
                         ids=["mass_outflow", "outflow"])
def test_outflow_balance_and_exit_velocities(kind):
    p = params()
    o = start(p, [inlet(), outlet(kind, vel=0.8)])
    seeded = 0
    st = o.get_particles()[3].copy()
    steps = 11000
    for nn in range(1, steps + 1):
        o.step(nn)
        st2 = o.get_particles()[3]
        seeded += int(((st == 0) & (st2 != 0)).sum())
        st = st2.copy()
    pos, vel, _, st = o.get_particles()
    assert abs(seeded - seeded_rate() * steps * p.dtsolid) <= 2.0
    deleted = seeded - int((st != 0).sum())
    assert deleted > 100
    assert np.all(pos[1, st != 0] <= BOX[1])
    exiting = st == 3
    assert exiting.any()
    assert np.all(pos[1, exiting] > BOX[1] - 0.5 * D)
    if kind == abi.BC_MASS_OUTFLOW:
        assert np.all(vel[1, exiting] == 0.8) and np.all(vel[0, exiting] == 0.0)
    else:
        assert np.all(vel[0, exiting] == 0.0) and np.all(vel[2, exiting] == 0.0)
        assert np.all(vel[1, exiting] > 0.0)
    o.close()


def test_new_spheres_take_the_lowest_free_slots():
    p = params(max_pip=3000)
    o = start(p, [inlet(), outlet()])
    o.run(1, 9000)                 # deletions at the top have begun
    checked = 0
    for nn in range(9001, 9400):
        before = o.get_particles()[3].copy()
        o.step(nn)
        st = o.get_particles()[3]
        new = np.flatnonzero((before == 0) & (st != 0))
        if new.size:
            free = np.flatnonzero(before == 0)
            # lowest free slots, unless a deletion in the same step freed a
            # lower one after the seeding (MASS_OUTFLOW_DEM runs second)
            assert np.array_equal(new, free[:new.size])
            checked += 1
    assert checked > 50
    o.close()

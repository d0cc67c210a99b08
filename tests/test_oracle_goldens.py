"""Pins the oracle against more of the reference's own DEM regression data
(AUTOTEST files, extracted to tests/golden/ by make_dem_fixtures.py):

* DEM03 (tests/dem/DEM03): two stacked spheres of two phases compressed
  between two walls -- particle-particle LSD contact with unequal masses
  (DES_ETAN(1,2), m_eff), two walls at once, Adams-Bashforth.  The MFIX
  columns are printed with 8 decimals.
* DEM04 (tests/dem/DEM04): a sphere slipping on a rough wall -- the wall's
  tangential spring, tangential history, Coulomb cap and torque.  Compared:
  the end-of-slip time and the translational and angular velocities at that
  moment, as the reference's user routine computes them
  (tests/dem/DEM04/write_usr1.f), printed with 8 decimals.  The end-of-slip
  step agrees exactly (time to 1e-8).  The angular velocity in the AUTOTEST
  file is higher by 1.5e-5 relative in every run: the file was written by a
  reference version whose wall torque used the lever arm R, while the
  current source uses |c - x| = R - overlap (calc_collision_wall_mod.f:606).
  With an undamped wall contact the force-weighted mean overlap is
  1.5 m g / KN_W = 7.7e-9 m, i.e. 1.54e-5 R, which is the offset seen; an
  oracle variant with lever arm R matches the file to 1e-7.  So omega r/u0
  is held to 1.5e-5 and v/u0 (which feels the torque only through the last
  sticking steps) to 2e-7.
"""
import json
import os

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem

HERE = os.path.join(os.path.dirname(__file__), "golden")
DEM03 = json.load(open(os.path.join(HERE, "dem03_post.json")))["runs"]
DEM04 = json.load(open(os.path.join(HERE, "dem04_post.json")))["runs"]
GRAV_Y = -9.80665  # reference default gravity (constant_mod.f)


def _system(p, parts):
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    return o


@pytest.mark.parametrize("k", range(len(DEM03)), ids=[f"en{r['en']:g}" for r in DEM03])
def test_dem03_two_body_stack(k):
    run = DEM03[k]
    p, parts, _ = workloads.dem03(en=run["en"])
    o = _system(p, parts)
    dt = o.dtsolid()
    rows = np.array(run["rows"])
    steps = np.rint(rows[:, 0] / dt).astype(int)
    # output times are whole steps of DTSOLID = TCOLL/50
    assert np.allclose(steps * dt, rows[:, 0], atol=6e-9)
    want = set(steps.tolist())
    y = {0: (0.00045, 0.00135)}
    for nn in range(1, steps.max() + 1):
        o.step(nn)
        if nn in want:
            pos = o.get_particles()[0]
            y[nn] = (pos[1, 0], pos[1, 1])
    got = np.array([y[s] for s in steps])
    assert np.max(np.abs(got[:, 0] - rows[:, 1])) <= 1.0e-8
    assert np.max(np.abs(got[:, 1] - rows[:, 2])) <= 1.0e-8
    o.close()


def _compare(v1, v2, tol=1.0e-4, small=1.0e-15):
    """Relative comparison of the reference's COMPARE (toleranc_mod.f:138)."""
    if abs(v1) <= small:
        return abs(v2) <= small
    return abs(v2 / v1 - 1.0) <= tol


@pytest.mark.parametrize("k", range(len(DEM04)), ids=[f"mew{r['mew']:g}" for r in DEM04])
def test_dem04_slip_to_roll(k):
    run = DEM04[k]
    p, parts, _ = workloads.dem04(mew=run["mew"])
    o = _system(p, parts)
    dt = o.dtsolid()
    u0, r = 0.1, 5.0e-4
    t = 0.0
    for nn in range(1, 200000):
        o.step(nn)
        t += dt
        _, vel, om, _ = o.get_particles()
        slip = vel[0, 0] + om[2, 0] * r
        if _compare(abs(slip), 1.0e-6) or slip < 0.0:
            break
    else:
        pytest.fail("slip never ended")
    time_nd = abs(run["mew"] * GRAV_Y / u0) * t
    assert abs(time_nd - run["time_nd"]) <= 1.0e-8
    assert abs(abs(vel[0, 0] / u0) - run["tvel_nd"]) <= 2.0e-7
    d_avel = abs(om[2, 0] * r / u0) - run["avel_nd"]
    # lever arm R - overlap (current source) vs R (file): omega lower, by
    # 1.5 * overlap_mean / R relative
    assert -1.5e-5 <= d_avel <= 0.0
    o.close()


DEM05 = json.load(open(os.path.join(HERE, "dem05_post.json")))


@pytest.mark.parametrize("model", ["LSD", "HERTZ"])
def test_dem05_oblique_wall_impacts(model):
    """DEM05 (tests/dem/DEM05): 31 spheres hit the south wall at impact
    angles 0-65 deg with friction 0.092 -- sticking/sliding transitions of
    the wall's tangential history, LSD and Hertz.  Compared (part-wall
    column, printed with 4 decimals): rebound angle, tangential restitution
    and spin after TSTOP = 2.5 ms, computed as tests/dem/DEM05/usr3_des.f
    does.  The other 62 particles (the particle-particle half of the test,
    whose targets the reference re-pins every step from user code) never
    touch these 31."""
    run = next(r for r in DEM05["runs"] if r["model"] == model)
    p, parts, _ = workloads.dem05(DEM05["particles"], model=model)
    o = _system(p, parts)
    dt = o.dtsolid()
    # pure-DEM step count: CEILING(real((TSTOP-TIME)/DTSOLID)), des_time_march.f:130
    nsteps = int(np.ceil(np.float32(0.0025 / dt)))
    o.run(1, nsteps)
    _, vel, om, _ = o.get_particles()
    v0 = parts.vel
    vt0 = np.sqrt(v0[0, :31] ** 2 + v0[2, :31] ** 2)
    ang0 = np.abs(np.arctan(vt0 / v0[1, :31])) * 180.0 / np.pi
    vt = np.sqrt(vel[0, :31] ** 2 + vel[2, :31] ** 2)
    alpha = np.arctan(vt / vel[1, :31]) * 180.0 / np.pi
    spin = np.sqrt(om[0, :31] ** 2 + om[2, :31] ** 2)
    ga, gc, gw = (np.array(run[k]) for k in ("alpha", "coeff", "omega"))
    assert np.allclose(ang0, ga[:, 0], atol=5.1e-5)
    tol = 5.1e-5  # half a unit of the 4th printed decimal
    assert np.max(np.abs(alpha - ga[:, 2])) <= tol
    assert np.max(np.abs(spin - gw[:, 2])) <= tol
    assert np.max(np.abs(vt[1:] / vt0[1:] - gc[:, 2])) <= tol
    o.close()


@pytest.mark.parametrize("model", ["LSD", "HERTZ"])
def test_dem05_oblique_particle_impacts(model):
    """DEM05 particle-particle half: spheres 32-62 (rho 4000) hit spheres
    63-93 (rho 5000), which the reference's user hook puts back every step
    at y = 0.0475 under their partner's x, z with zero velocity and spin
    (tests/dem/DEM05/usr2_des.f).  The same hook is applied to the oracle
    between steps.  Pins particle-particle LSD and Hertz contact with
    tangential history, the Coulomb cap and the torque split between two
    phases.  Compared (part-part column, 4 printed decimals): rebound angle,
    tangential restitution and spin."""
    run = next(r for r in DEM05["runs"] if r["model"] == model)
    p, parts, _ = workloads.dem05(DEM05["particles"], model=model)
    o = _system(p, parts)
    dt = o.dtsolid()
    nsteps = int(np.ceil(np.float32(0.0025 / dt)))
    zero = np.zeros((3, 31))
    tgt = np.zeros((3, 31))
    tgt[1, :] = 0.0475
    for nn in range(1, nsteps + 1):
        o.step(nn)
        pos = o.get_particles()[0]
        tgt[0, :] = pos[0, 31:62]
        tgt[2, :] = pos[2, 31:62]
        o.user_set_particles(62, pos=tgt, vel=zero, omega=zero)
    _, vel, om, _ = o.get_particles()
    sl = slice(31, 62)
    v0 = parts.vel
    vt0 = np.sqrt(v0[0, sl] ** 2 + v0[2, sl] ** 2)
    vt = np.sqrt(vel[0, sl] ** 2 + vel[2, sl] ** 2)
    alpha = np.arctan(vt / vel[1, sl]) * 180.0 / np.pi
    spin = np.sqrt(om[0, sl] ** 2 + om[2, sl] ** 2)
    ga, gc, gw = (np.array(run[k]) for k in ("alpha", "coeff", "omega"))
    tol = 5.1e-5  # half a unit of the 4th printed decimal
    assert np.max(np.abs(alpha - ga[:, 1])) <= tol
    assert np.max(np.abs(spin - gw[:, 1])) <= tol
    assert np.max(np.abs(vt[1:] / vt0[1:] - gc[:, 1])) <= tol
    o.close()

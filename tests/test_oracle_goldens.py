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

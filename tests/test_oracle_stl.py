"""STL walls and moving walls in the oracle, pinned against the reference's
DEM01 output and against rigid-body closed forms.

* DEM01 with its south wall given as two user STL triangles (default box
  walls off): the bouncing sphere must reproduce the reference's DEM01
  output (POST_POS / POST_VEL, 8 decimals) exactly as the box wall does.
* A sphere released on an STL incline at 20 degrees with enough friction
  to roll (mu = 0.3 > 2/7 tan 20 deg): its speed along the slope must grow
  as (5/7) g sin(theta) t, the rolling-without-slipping closed form.
* A sphere at rest on a floor moving at U (a BC wall velocity): friction
  spins it up until its contact point moves with the floor, which leaves
  it at (2/7) U.
A wrong sign or a dropped term of the wall velocity, or a facet order that
loses the STL floor, fails these by far more than their tolerances."""
import json
import os

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dem01_post.json")))


def quad(p0, p1, p2, p3, n):
    """Two triangles (p0, p1, p2) and (p0, p2, p3) with normal n, in
    VERTEX(3,3,nf) memory order: vertex[f, coordinate, vertex]."""
    tris = [(p0, p1, p2), (p0, p2, p3)]
    v = np.array([[[t[k][c] for k in range(3)] for c in range(3)] for t in tris])
    return v, np.array([n, n], dtype=np.float64)


def test_dem01_on_an_stl_floor_matches_reference_output():
    run = next(r for r in GOLD["runs"] if r["scheme"] == "EULER")
    p, parts, _ = workloads.dem01(kn=run["kn"], en=run["en"], integrator=abi.EULER)
    # the shared diagonal (z = x + 0.3) passes away from the sphere at
    # x = z = 0.5: contacts on an edge between coplanar facets are ignored
    v, n = quad((-1.0, 0.0, -0.7), (2.0, 0.0, -0.7), (2.0, 0.0, 2.3), (-1.0, 0.0, 2.3),
                (0.0, 1.0, 0.0))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_facets(v, n, default_walls=False)
    o.init()
    dt = o.dtsolid()
    rows = np.array(run["rows"])
    steps = np.rint(rows[:, 0] / dt).astype(int)
    ys, vs = {0: 0.5}, {0: 0.0}
    for nn in range(1, steps.max() + 1):
        o.step(nn)
        if nn in set(steps.tolist()):
            pos, vel, _, _ = o.get_particles()
            ys[nn], vs[nn] = pos[1, 0], vel[1, 0]
    y = np.array([ys[s] for s in steps])
    vy = np.array([vs[s] for s in steps])
    assert np.max(np.abs(y - rows[:, 1])) <= 1.0e-8
    assert np.max(np.abs(vy - rows[:, 2])) <= 1.0e-8
    o.close()


def _floor_box(p, with_walls=True):
    """A 0.2 m cube, one sphere of d = 10 mm, LSD."""
    p.xlength = p.ylength = p.zlength = 0.2
    p.imax = p.jmax = p.kmax = 4
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = 4
    return p


def one_sphere(pos, vel=(0.0, 0.0, 0.0), mew=0.3, grav=(0.0, -9.81, 0.0)):
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = 1
    p.max_pip = 1
    p = _floor_box(p)
    p.kn = p.kn_w = 1.0e4
    p.mew = p.mew_w = mew
    p.des_en_input[0] = 0.9
    p.des_en_wall_input[0] = 0.9
    p.d_p0[0] = 0.01
    p.ro_s0[0] = 2500.0
    p.max_radius = 0.005
    p.grav[:] = grav
    parts = workloads.make_particles(np.array(pos, dtype=np.float64).reshape(3, 1), 0.005,
                                     2500.0, vel=np.array(vel, dtype=np.float64).reshape(3, 1))
    return p, parts


def test_sphere_rolls_down_an_stl_incline():
    th = np.deg2rad(20.0)
    t_hat = np.array([np.cos(th), -np.sin(th), 0.0])      # down the slope
    n_hat = np.array([np.sin(th), np.cos(th), 0.0])        # up from the surface
    c0 = np.array([0.05, 0.15, 0.1])                       # a point on the plane
    w = 0.3
    corners = [c0 - w * t_hat - np.array([0, 0, w]), c0 + w * t_hat - np.array([0, 0, w]),
               c0 + w * t_hat + np.array([0, 0, w]), c0 - w * t_hat + np.array([0, 0, w])]
    v, n = quad(*corners, n_hat)
    r = 0.005
    start = c0 + 0.999 * r * n_hat                         # resting on the plane
    p, parts = one_sphere(start)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_facets(v, n, default_walls=False)
    o.init()
    dt = o.dtsolid()
    nsteps = int(round(0.05 / dt))
    o.run(1, nsteps)
    pos, vel, om, _ = o.get_particles()
    t = nsteps * dt
    v_slope = float(vel[:, 0] @ t_hat)
    expect = 5.0 / 7.0 * 9.81 * np.sin(th) * t
    assert abs(v_slope - expect) <= 0.01 * expect
    # no slip: the contact point is at rest, omega r = v
    assert abs(abs(om[2, 0]) * r - v_slope) <= 0.01 * expect
    o.close()


@pytest.mark.parametrize("u", [0.5, -0.3])
def test_sphere_on_a_moving_floor_reaches_two_sevenths(u):
    # the floor (y = 0 plane of the box, facet 2 = south of W E S N B T)
    # moves along x at u; no user facets
    p, parts = one_sphere((0.1, 0.005 * 0.999, 0.1))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    o.set_wall_motion(2, 1, (u, 0.0, 0.0))
    dt = o.dtsolid()
    o.run(1, int(round(0.2 / dt)))
    _, vel, om, _ = o.get_particles()
    assert abs(vel[0, 0] - 2.0 / 7.0 * u) <= 0.02 * abs(u)
    assert abs(vel[2, 0]) < 1e-9
    o.close()


def cone(apex, half_angle, height, sides):
    """A downward cone of `sides` triangles (apex at the bottom, axis +y),
    normals pointing into the cone."""
    apex = np.asarray(apex, dtype=np.float64)
    rad = height * np.tan(half_angle)
    verts, norms = [], []
    for k in range(sides):
        a0, a1 = 2 * np.pi * k / sides, 2 * np.pi * (k + 1) / sides
        p0 = apex + np.array([rad * np.cos(a0), height, rad * np.sin(a0)])
        p1 = apex + np.array([rad * np.cos(a1), height, rad * np.sin(a1)])
        n = np.cross(p1 - apex, p0 - apex)
        n /= np.linalg.norm(n)
        if n[1] < 0.0:
            n = -n
        tri = (apex, p0, p1) if np.cross(p0 - apex, p1 - apex) @ n > 0 else (apex, p1, p0)
        verts.append(np.stack(tri, axis=1))
        norms.append(n)
    return np.array(verts), np.array(norms)


def sphere_in_cone(sides):
    """A sphere settled at the bottom of a cone touches every one of its
    `sides` facets at once."""
    apex = (0.1, 0.02, 0.1)
    v, n = cone(apex, np.deg2rad(40.0), 0.05, sides)
    r = 0.005
    h = 0.999 * r / np.sin(np.deg2rad(40.0))            # just overlapping every face
    p, parts = one_sphere((0.1, 0.02 + h, 0.1))
    return p, parts, v, n


def test_more_than_eight_wall_contacts_is_a_capacity_error():
    """COLLISION_ARRAY_MAX is 8 (get_collision, calc_collision_wall_mod.f:
    663-700); the reference grows the table, this build reports
    MFIXDEM_ERR_CAPACITY.  Eight facets at once are fine, nine are not."""
    p, parts, v, n = sphere_in_cone(8)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_facets(v, n, default_walls=True)
    o.init()
    o.run(1, 20)
    o.close()
    p, parts, v, n = sphere_in_cone(9)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_facets(v, n, default_walls=True)
    o.init()
    with pytest.raises(abi.MfixDemError) as e:
        o.step(1)
    assert e.value.code == abi.ERRORS_BY_NAME["CAPACITY"]
    o.close()

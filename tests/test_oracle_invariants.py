"""Invariants that pin the oracle's contact, history-carry, GARG_2012
deposition and GARG_2012 drag interpolation independently of its code:

* Newton's third law and the torque lever arms (calc_force_dem.f:356-382,
  dist_cl + dist_ci = |d|): with no walls, gas or gravity, the Euler
  integrator (cfnewvalues.f:700-738) conserves total linear momentum and
  total angular momentum sum m x x v + I omega exactly up to rounding
  (x_{n+1} x v_{n+1} = x_n x v_{n+1}, and the pair torques cancel the
  moment of the pair forces).
* The tangential history carried over a rebuild (neighbour.f:77-117) is
  the history itself: isolated colliding pairs give bit-identical
  trajectories whether the pair list is rebuilt every step or never.
* COMP_MEAN_FIELDS0 (comp_mean_fields0.f:186-210, 339-432): a particle's
  mass goes to the 8 corners of its fluid cell with trilinear weights and
  every interior node spreads its mass evenly over its 8 cells, so
  ROP_S(c) = m / (8 VOL) * sum of the particle's weights on c's corners;
  the total solids mass is conserved.
* DRAG_GS_DES0 interpolation (drag_gs_des0.f:148-179, 507-566): node
  velocities are averages of 4 staggered face values and the particle
  velocity is trilinear in the nodes, so a linear gas field is reproduced
  exactly at the particle.
"""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem


def cloud(model, n=300, seed=11):
    """Polydisperse spheres jostled (overlapping, random translation and
    spin) in the middle of a large box: no wall is reached, no gravity.
    Unequal radii make the two lever arms of a contact differ."""
    p, parts, _ = workloads.settling_bed(n=n, seed=seed, box=(0.05, 0.05, 0.05))
    r = 0.5 * workloads.lognormal_diameters(np.random.default_rng(seed), n)
    r = r[np.argsort(-r, kind="stable")]
    pos = workloads.rsa_positions(seed, r, [0.0205] * 3, [0.0295] * 3)
    p.max_radius = float(r.max())
    parts = workloads.make_particles(pos, r, parts.ro_sol)
    p, parts = workloads.jostled(p, parts, seed=seed, v0=0.2, w0=100.0)
    p.grav[:] = [0.0, 0.0, 0.0]
    if model == "HERTZ":
        p.coll_model = abi.HERTZ
        p.des_et_input[0] = 0.9
        p.des_et_wall_input[0] = 0.9
        p.e_young[0] = 1.0e7
        p.v_poisson[0] = 0.3
        p.ew_young = 1.0e7
        p.vw_poisson = 0.3
    return p, parts


def momenta(pos, vel, om, radius, rho):
    m = rho * (4.0 / 3.0) * np.pi * radius ** 3
    inertia = m * radius ** 2 / 2.5
    lin = (m * vel).sum(axis=1)
    ang = (m * np.cross(pos.T, vel.T).T).sum(axis=1) + (inertia * om).sum(axis=1)
    scale_l = (m * np.abs(vel)).sum()
    scale_a = (m * np.linalg.norm(pos, axis=0) * np.linalg.norm(vel, axis=0)).sum() + \
        (inertia * np.abs(om)).sum()
    return lin, ang, scale_l, scale_a


@pytest.mark.parametrize("model", ["LSD", "HERTZ"])
def test_momentum_and_angular_momentum_conserved(model):
    p, parts = cloud(model)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    l0, a0, sl, sa = momenta(parts.pos, parts.vel, parts.omega, parts.radius, parts.ro_sol)
    o.run(1, 300)
    pos, vel, om, st = o.get_particles()
    assert st.min() == 1
    l1, a1, _, _ = momenta(pos, vel, om, parts.radius, parts.ro_sol)
    m = parts.ro_sol * (4.0 / 3.0) * np.pi * parts.radius ** 3
    ke0 = 0.5 * (m * parts.vel ** 2).sum()
    ke1 = 0.5 * (m * vel ** 2).sum()
    assert ke1 < 0.97 * ke0                    # dissipative contacts took part
    assert np.max(np.abs(l1 - l0)) <= 1e-14 * sl
    assert np.max(np.abs(a1 - a0)) <= 1e-13 * sa
    o.close()


def colliding_pairs(npairs=8, seed=5):
    """Isolated pairs of d = 1 mm spheres approaching each other obliquely
    (normal and tangential relative velocity, spin), far apart from each
    other and from the walls; soft springs make each contact last about
    80 steps."""
    rng = np.random.default_rng(seed)
    p, _, _ = workloads.settling_bed(n=2 * npairs, seed=seed, box=(0.05, 0.05, 0.05))
    p.grav[:] = [0.0, 0.0, 0.0]
    p.kn = p.kn_w = 10.0
    p.mew = p.mew_w = 0.3
    pos = np.zeros((3, 2 * npairs))
    vel = np.zeros((3, 2 * npairs))
    for q in range(npairs):
        c = np.array([0.01 + 0.01 * (q % 4), 0.015 + 0.02 * (q // 4), 0.025])
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        pos[:, 2 * q] = c - 0.51e-3 * n
        pos[:, 2 * q + 1] = c + 0.51e-3 * n
        vel[:, 2 * q] = 0.02 * n + 0.03 * t
        vel[:, 2 * q + 1] = -0.02 * n - 0.01 * t
    om = rng.uniform(-20.0, 20.0, (3, 2 * npairs))
    return p, workloads.make_particles(pos, 0.5e-3, 2500.0, vel, om)


def test_history_carry_is_exact_across_rebuilds():
    runs = {}
    for nsn in (1, 7, 1000):
        p, parts = colliding_pairs()
        p.neighbor_search_n = nsn
        o = OracleSystem(p)
        o.set_particles(parts)
        o.init()
        o.run(1, 400)
        runs[nsn] = o.get_particles()[:3]
        o.close()
    base = runs[1000]
    # friction acted: the spins changed by the tangential impulses
    assert np.max(np.abs(base[2] - colliding_pairs()[1].omega)) > 1.0
    for nsn in (1, 7):
        for a, b in zip(runs[nsn], base):
            assert np.array_equal(a, b)


def garg_single(pos, cells=(6, 8, 6), box=(0.012, 0.016, 0.012)):
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.max_pip = 1
    p.coupled = 1
    p.interp_scheme = abi.INTERP_GARG_2012
    p.des_interp_on = p.des_interp_mean_fields = 1
    p.xlength, p.ylength, p.zlength = box
    p.imax, p.jmax, p.kmax = cells
    p.kn = p.kn_w = 1.0e3
    p.mew = p.mew_w = 0.0
    p.des_en_input[0] = 0.9
    p.des_en_wall_input[0] = 0.9
    p.dtsolid = 1.0e-5
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 2500.0
    p.max_radius = 0.5e-3
    parts = workloads.make_particles(np.array(pos, dtype=float).reshape(3, 1), 0.5e-3, 2500.0)
    return p, parts


def cell_index(p, i, j, k):
    """0-based linear index of fluid cell (I, J, K) (1-based, ghosts
    included; J fastest, then I, then K)."""
    ni, nj = p.imax + 2, p.jmax + 2
    return (j - 1) + (i - 1) * nj + (k - 1) * ni * nj


def test_garg_deposition_closed_form():
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    nc = workloads.n_cells(p)
    gas = workloads.Gas(np.zeros(nc), np.zeros(nc), np.zeros(nc), np.zeros((nc, 3)))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    ep, rop = o.get_gas()
    d = np.array([p.xlength / p.imax, p.ylength / p.jmax, p.zlength / p.kmax])
    vol = d.prod()
    m = 2500.0 * (4.0 / 3.0) * np.pi * 0.5e-3 ** 3
    # cell (I,J,K) spans [(I-2) d, (I-1) d]; its corners are nodes I-1..I
    cell = (np.floor(x / d) + 2).astype(int)
    zeta = x / d - (cell - 2)
    want = np.zeros(nc)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = ((zeta[0] if a else 1 - zeta[0]) * (zeta[1] if b else 1 - zeta[1]) *
                     (zeta[2] if c else 1 - zeta[2]))
                node = cell - 1 + np.array([a, b, c])
                for da in (0, 1):          # the 8 cells around node (i,j,k)
                    for db in (0, 1):
                        for dc in (0, 1):
                            want[cell_index(p, node[0] + da, node[1] + db, node[2] + dc)] += \
                                w * m / (8.0 * vol)
    assert np.max(np.abs(rop[0] - want)) <= 1e-12 * want.max()
    assert np.max(np.abs(ep - (1.0 - want / 2500.0))) <= 1e-15
    o.close()


def test_garg_deposition_conserves_mass():
    p, parts, gas = workloads.fluidised_bed(n=3000, seed=4, lx=0.02, ly=0.04, lz=0.02,
                                            cells=(10, 20, 10), dg=(0, 0, 0))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    o.run(1, 20)
    _, rop = o.get_gas()
    # fluid cells only: the reference's node loop also adds density x VOL
    # to ghost cells and divides by VOL only in fluid cells
    # (comp_mean_fields0.f:339-432), so ghost entries are not densities
    fluid = rop[0].reshape(12, 12, 22)[1:-1, 1:-1, 1:-1]
    vol = (0.02 / 10) * (0.04 / 20) * (0.02 / 10)
    mass = (2500.0 * (4.0 / 3.0) * np.pi * parts.radius ** 3).sum()
    assert abs(fluid.sum() * vol - mass) <= 1e-13 * mass
    o.close()


def test_garg_drag_reproduces_linear_gas_field():
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    d = np.array([p.xlength / p.imax, p.ylength / p.jmax, p.zlength / p.kmax])
    k, i, j = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), np.arange(1, nj + 1),
                          indexing="ij")
    face = [(i - 1) * d[0], (j - 1) * d[1], (k - 1) * d[2]]          # east/north/top faces
    ctr = [(i - 1.5) * d[0], (j - 1.5) * d[1], (k - 1.5) * d[2]]
    a0 = np.array([0.05, 0.6, -0.03])
    A = np.array([[3.0, -2.0, 1.5], [4.0, 7.0, -5.0], [-1.0, 2.5, 6.0]])

    def field(c, xs):
        return (a0[c] + A[c, 0] * xs[0] + A[c, 1] * xs[1] + A[c, 2] * xs[2]).ravel()
    u = field(0, [face[0], ctr[1], ctr[2]])
    v = field(1, [ctr[0], face[1], ctr[2]])
    w = field(2, [ctr[0], ctr[1], face[2]])
    nc = u.size
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(workloads.Gas(np.ascontiguousarray(u), np.ascontiguousarray(v),
                            np.ascontiguousarray(w), np.zeros((nc, 3))))
    o.init()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    eps = ep[pijk[3, 0] - 1]
    o.step(1)
    fc, _ = o.get_forces()
    ug = a0 + A @ x
    vrel = np.linalg.norm(ug)
    fgp = o.lib.orc_drag_fgp(p.drag_type, eps, p.mu_g0, p.ro_g0, vrel, 1.0e-3,
                             np.pi / 6.0 * 1.0e-9)
    assert np.allclose(fc[:, 0], fgp * ug, rtol=1e-13, atol=0)
    o.close()

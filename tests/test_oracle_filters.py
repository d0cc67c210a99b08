"""Pins the oracle's filter-scheme path (the DES1 side of the reference:
CALC_INTERP_WEIGHTS, DRAG_GS_DES1, COMP_MEAN_FIELDS1) by what mathematics
fixes, independently of the restated code:

* SQUARE_DPVM / GAUSS_DPVM weights (calc_interp_weights.f:47-156) are the
  mass of the Epanechnikov kernel K(s) = 3/(4h) (1 - s^2/h^2), h = width/2,
  lying beyond each cell face; checked against numerical integration of K
  and the 27-entry order (K outer, then I, then J fastest).
* LINEAR_HAT weights (calc_interp_weights.f:351-528) are the linear hat
  about the cell centre; checked against the closed form and the order
  (K outer, J, I fastest).
* Weights sum to one (DPVM always, LINEAR_HAT in the interior); deposition
  conserves solids volume; the NONE scheme deposits PVOL into the own cell.
* DES1 drag with a uniform gas field equals the textbook Gidaspow / Wen-Yu
  force on the interpolated voidage (drag_gs.f:536-758 describe the laws).
* DES1 drag with a gas field linear in space (staggered U_G/V_G/W_G, cell
  P_FORCE) equals the law at the closed-form interpolated velocity plus the
  closed-form interpolated pressure force (linear_gas.py): pins the
  cell-centre face averages of CALC_CELL_CENTER_GAS_VEL and the cells the
  27 weights are applied to, for every scheme, with and without
  DES_INTERP_ON.
"""
import warnings

import numpy as np
import pytest
from scipy import integrate

import _paths  # noqa: F401
from mfixdem import abi, workloads
from linear_gas import (LINEAR_CASES, LINEAR_IDS, linear_case, linear_expectation,
                        textbook_fgp)
from oracle_py import OracleSystem

LX, LY, LZ, CELLS = 0.02, 0.04, 0.02, (10, 20, 10)
DX = LX / CELLS[0]


def bed(scheme, width=None, interp_on=1, n=3000, seed=4):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=LX, ly=LY, lz=LZ, cells=CELLS,
                                            dg=(0, 0, 0))
    p.interp_scheme = scheme
    p.des_interp_on = interp_on if scheme != abi.INTERP_NONE else 0
    p.des_interp_width = abi.UNDEFINED if width is None else width
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    return p, parts, gas, o


def interior(o, parts):
    """Particles whose 27 surrounding cells are all fluid cells."""
    _, pijk = o.get_cells()
    i, j, k = pijk[0], pijk[1], pijk[2]
    ok = (i >= 3) & (i <= CELLS[0]) & (j >= 3) & (j <= CELLS[1]) & (k >= 3) & (k <= CELLS[2])
    return np.nonzero(ok)[0], pijk


def faces(n, d):
    return -d + d * np.arange(n + 3)  # xe(0..n+2), 0-based index = face index


def epan_beyond(dist, h):
    """Mass of the Epanechnikov kernel of half-width h beyond a face at
    distance dist from its centre."""
    if dist >= h:
        return 0.0
    k = lambda s: 0.75 / h * (1.0 - (s / h) ** 2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        return integrate.quad(k, dist, h, epsabs=1e-15, epsrel=1e-14)[0]


@pytest.mark.parametrize("scheme", [abi.INTERP_SQUARE_DPVM, abi.INTERP_GAUSS_DPVM],
                         ids=["square", "gauss"])
def test_dpvm_weights_are_kernel_mass(scheme):
    width = 1.6e-3
    h = 0.5 * width
    p, parts, _, o = bed(scheme, width=width)
    idx, pijk = interior(o, parts)
    xe = [faces(CELLS[d], [LX, LY, LZ][d] / CELLS[d]) for d in range(3)]
    for q in idx[:40]:
        ii = [pijk[0][q], pijk[1][q], pijk[2][q]]
        w = []
        for d in range(3):
            x = parts.pos[d, q]
            lo = epan_beyond(x - xe[d][ii[d] - 1], h)
            hi = epan_beyond(xe[d][ii[d]] - x, h)
            w.append([lo, 1.0 - lo - hi, hi])
        want = np.array([w[0][ic] * w[1][jc] * w[2][kc]
                         for kc in range(3) for ic in range(3) for jc in range(3)])
        cells, got = o.filter(q)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
        wc = np.array([(ii[1] - 1 + jc - 1) + (ii[0] - 1 + ic - 1) * (CELLS[1] + 2) +
                       (ii[2] - 1 + kc - 1) * (CELLS[1] + 2) * (CELLS[0] + 2)
                       for kc in range(3) for ic in range(3) for jc in range(3)])
        assert np.array_equal(cells, wc)
    o.close()


def test_linear_hat_weights_closed_form():
    p, parts, _, o = bed(abi.INTERP_LINEAR_HAT)
    idx, pijk = interior(o, parts)
    xe = [faces(CELLS[d], [LX, LY, LZ][d] / CELLS[d]) for d in range(3)]
    for q in idx[:40]:
        ii = [pijk[0][q], pijk[1][q], pijk[2][q]]
        w = []
        for d in range(3):
            dd = [LX, LY, LZ][d] / CELLS[d]
            xc = xe[d][ii[d] - 1] + 0.5 * dd
            t = (parts.pos[d, q] - xc) / dd
            w.append([-t, 1.0 + t, 0.0] if t < 0 else [0.0, 1.0 - t, t])
        want = np.array([w[0][ic] * w[1][jc] * w[2][kc]
                         for kc in range(3) for jc in range(3) for ic in range(3)])
        cells, got = o.filter(q)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
        assert abs(got.sum() - 1.0) <= 1e-14
    o.close()


@pytest.mark.parametrize("scheme", [abi.INTERP_SQUARE_DPVM, abi.INTERP_LINEAR_HAT])
def test_weight_sums(scheme):
    width = 1.6e-3 if scheme == abi.INTERP_SQUARE_DPVM else None
    p, parts, _, o = bed(scheme, width=width)
    idx, _ = interior(o, parts)
    sums = np.array([o.filter(q)[1].sum() for q in range(parts.n)])
    if scheme == abi.INTERP_SQUARE_DPVM:   # weights at walls go to the own cell
        assert np.max(np.abs(sums - 1.0)) <= 1e-14
    else:                                  # at walls the hat loses its outer part
        assert np.max(np.abs(sums[idx] - 1.0)) <= 1e-14
        assert np.all(sums <= 1.0 + 1e-14)
    o.close()


@pytest.mark.parametrize("scheme", [abi.INTERP_SQUARE_DPVM, abi.INTERP_NONE])
def test_deposition_conserves_solids_volume(scheme):
    width = 1.6e-3 if scheme == abi.INTERP_SQUARE_DPVM else None
    p, parts, _, o = bed(scheme, width=width)
    ep, rop = o.get_gas()
    vol = DX ** 3
    got = rop.sum() * vol / 2500.0
    want = np.sum((4.0 / 3.0) * np.pi * parts.radius ** 3)
    assert abs(got - want) <= 1e-13 * want
    o.close()


def test_none_scheme_deposits_into_own_cell():
    p, parts, _, o = bed(abi.INTERP_NONE)
    _, pijk = o.get_cells()
    ep, rop = o.get_gas()
    vols = (4.0 / 3.0) * np.pi * parts.radius ** 3
    want = np.zeros(rop.shape[1])
    np.add.at(want, pijk[3] - 1, vols)
    want = 2500.0 * want / DX ** 3
    assert np.allclose(rop[0], want, rtol=1e-13, atol=0)
    assert np.allclose(ep[want > 0], 1.0 - want[want > 0] / 2500.0, rtol=0, atol=1e-15)
    o.close()




@pytest.mark.parametrize("scheme,drag", [(abi.INTERP_SQUARE_DPVM, abi.DRAG_GIDASPOW),
                                         (abi.INTERP_LINEAR_HAT, abi.DRAG_WEN_YU),
                                         (abi.INTERP_NONE, abi.DRAG_GIDASPOW)],
                         ids=["dpvm-gidaspow", "lhat-wenyu", "none-gidaspow"])
def test_des1_drag_uniform_field(scheme, drag):
    """One sphere in a uniform gas: FC = F_gp (U - v) + P_FORCE PVOL with
    F_gp from the textbook law at the interpolated voidage."""
    p = abi.default_params()
    p.max_pip = 1
    p.coupled = 1
    p.drag_type = drag
    p.interp_scheme = scheme
    p.des_interp_on = 1 if scheme != abi.INTERP_NONE else 0
    p.des_interp_width = 1.6e-3 if scheme == abi.INTERP_SQUARE_DPVM else abi.UNDEFINED
    p.xlength, p.ylength, p.zlength = LX, LY, LZ
    p.imax, p.jmax, p.kmax = CELLS
    p.dtsolid = 1.0e-5
    p.kn = p.kn_w = 1.0e3
    p.des_en_input[0] = p.des_en_wall_input[0] = 0.9
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 2500.0
    p.max_radius = 5.0e-4
    parts = workloads.make_particles(np.array([[0.0101], [0.0203], [0.0097]]), 5.0e-4, 2500.0,
                                     vel=np.array([[0.05], [-0.3], [0.02]]))
    nc = (CELLS[0] + 2) * (CELLS[1] + 2) * (CELLS[2] + 2)
    U = np.array([0.1, 1.2, -0.2])
    gas = workloads.Gas(np.full(nc, U[0]), np.full(nc, U[1]), np.full(nc, U[2]),
                        np.tile([0.0, 11.8, 0.3], nc).astype(np.float64))
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    ep, _ = o.get_gas()
    if scheme == abi.INTERP_NONE:
        eps = ep[o.get_cells()[1][3][0] - 1]
    else:
        cells, w = o.filter(0)
        eps = float(np.sum(ep[cells] * w))
    o.step(1)
    fc, _ = o.get_forces()
    r = 5.0e-4
    pvol = (4.0 / 3.0) * np.pi * r ** 3
    v = parts.vel[:, 0]
    fgp = textbook_fgp(drag, eps, p.mu_g0, p.ro_g0, np.linalg.norm(U - v), 2 * r, pvol)
    want = fgp * (U - v) + np.array([0.0, 11.8, 0.3]) * pvol
    assert np.allclose(fc[:, 0], want, rtol=1e-12, atol=0)
    o.close()


@pytest.mark.parametrize("drag", [abi.DRAG_GIDASPOW, abi.DRAG_WEN_YU], ids=["gidaspow", "wenyu"])
def test_drag_law_textbook(drag):
    """DES_DRAG_GP's F_gp (oracle orc_drag_fgp) over dense and dilute
    voidage, low and high Reynolds number."""
    from oracle_py import load_oracle
    lib = load_oracle()
    dp, mu, rho = 1.0e-3, 1.8e-5, 1.2
    pvol = np.pi / 6.0 * dp ** 3
    for eps in (0.4, 0.6, 0.8, 0.80001, 0.9, 0.999):
        for vrel in (0.0, 1e-3, 0.5, 3.0, 40.0):
            got = lib.orc_drag_fgp(drag, eps, mu, rho, vrel, dp, pvol)
            want = textbook_fgp(drag, eps, mu, rho, vrel, dp, pvol)
            assert abs(got - want) <= 1e-13 * max(abs(want), 1e-300), (eps, vrel)


def cell_centres(p):
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    dx, dy, dz = p.xlength / p.imax, p.ylength / p.jmax, p.zlength / p.kmax
    kk, ii, jj = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), np.arange(1, nj + 1),
                             indexing="ij")
    # J fastest, then I, then K: the reference's cell numbering
    x = (ii - 1.5) * dx
    y = (jj - 1.5) * dy
    z = (kk - 1.5) * dz
    fluid = (ii >= 2) & (ii <= p.imax + 1) & (jj >= 2) & (jj <= p.jmax + 1) & \
        (kk >= 2) & (kk <= p.kmax + 1)
    return x.ravel(), y.ravel(), z.ravel(), fluid.ravel(), (dx, dy, dz)


def test_pressure_gradient_linear_and_quadratic_fields():
    """CALC_PG_GRAD: the 3-point derivative (with linear extrapolation next
    to walls) is exact for a linear pressure field everywhere and for a
    quadratic one away from walls; P_FORCE = -grad p, 0 outside fluid."""
    p, parts, gas, o = bed(abi.INTERP_GARG_2012, n=500)
    x, y, z, fluid, (dx, dy, dz) = cell_centres(p)
    a = np.array([30.0, -2500.0, 7.0])
    o.set_gas_pressure(1.0e5 + a[0] * x + a[1] * y + a[2] * z)
    pf = o.get_p_force()
    assert np.allclose(pf[fluid], -a, rtol=1e-9, atol=1e-9)
    assert not pf[~fluid].any()
    o.set_gas_pressure(3.0e4 * y * y)
    pf = o.get_p_force()
    _, pijk = o.get_cells()
    ny = p.jmax
    jj = (np.arange(len(y)) % (ny + 2)) + 1
    inner = fluid & (jj > 2) & (jj < ny + 1)
    assert np.allclose(pf[inner, 1], -6.0e4 * y[inner], rtol=1e-9, atol=1e-6)
    o.close()


# --------------------------------------------------------------------------
# DES1 drag in a gas field linear in space (closed forms in linear_gas.py)


@pytest.mark.parametrize("drag", [abi.DRAG_GIDASPOW, abi.DRAG_WEN_YU], ids=["gidaspow", "wenyu"])
@pytest.mark.parametrize("scheme,interp_on,width", LINEAR_CASES, ids=LINEAR_IDS)
def test_des1_drag_linear_field_closed_form(scheme, interp_on, width, drag):
    """CALC_CELL_CENTER_GAS_VEL's face averages and the 27-point
    interpolation of the cell-centre velocity and P_FORCE
    (drag_gs_des1.f:96-116,313-370) in a linear gas field: FC of each
    sphere equals the drag law at the closed-form interpolated velocity plus
    the closed-form interpolated pressure force, first with the spheres'
    own velocities, then with spheres moving at the interpolated gas
    velocity (drag ~ 0, the pressure force alone)."""
    p, parts, gas = linear_case(scheme, interp_on, width, drag)
    for match in (False, True):
        o = OracleSystem(p)
        o.set_particles(parts)
        o.set_gas(gas)
        o.init()
        ep, _ = o.get_gas()
        vel, want = linear_expectation(p, parts, ep, o.filter, match_gas_velocity=match)
        if match:
            o.close()
            parts = workloads.make_particles(parts.pos, parts.radius, parts.ro_sol, vel=vel)
            o = OracleSystem(p)
            o.set_particles(parts)
            o.set_gas(gas)
            o.init()
        o.step(1)
        fc, _ = o.get_forces()
        o.close()
        scale = np.max(np.abs(want))
        assert np.max(np.abs(fc - want)) <= 1e-12 * scale, (match, np.max(np.abs(fc - want)) / scale)

"""Non-uniform fluid grids in the oracle: cell sizes DX(I), DY(J), DZ(K)
(shift_dxyz.f:36-44), faces XE(I) = XE(I-1) + DX(I) (set_geometry_des.f:
72-87), VOL = DX*DY*DZ (set_geometry1.f:125), pinned by brute force and by
closed forms that a uniform-grid shortcut would miss:

* PIJK of every sphere, before and after it moves, equals a brute-force
  search over the faces built here from the cell sizes;
* the GARG_2012 deposition of one sphere equals the trilinear closed form
  with its cell's own sizes, and a bed's ROP_S x VOL sums to its mass with
  VOL = DX(I) DY(J) DZ(K);
* CALC_PG_GRAD is exact for a linear pressure everywhere and a quadratic
  one away from walls (the 3-point formula with the three cells' sizes,
  calc_grad_des.f:85-133);
* the filter-scheme drag reproduces a linear gas field (the closed forms of
  linear_gas.py with the cell centres of the stretched grid; LINEAR_HAT's
  oDX_E = 1/(0.5 (DX(I+1) + DX(I)))), and DRAG_GS_GAS0's node sums
  weighted by DES_VOL_NODE (each cell's dx dy dz, cfassign.f:154-157) give
  back sum F_GP;
* the default DES grid follows the smallest cell (check_geometry.f:137-148).
Using the uniform size anywhere moves these far beyond their tolerances."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from linear_gas import (LINEAR_CASES, LINEAR_IDS, cell_sizes, grid, linear_case,
                        linear_expectation, stretched_sizes)
from oracle_py import OracleSystem
from test_oracle_invariants import cell_index, garg_single

STRETCH = (1.15, 1.0 / 1.1, 1.07)


def stretch(p):
    p.set_cell_sizes(stretched_sizes(p.imax, p.xlength, STRETCH[0]),
                     stretched_sizes(p.jmax, p.ylength, STRETCH[1]),
                     stretched_sizes(p.kmax, p.zlength, STRETCH[2]))
    return p


def bed(n=3000, seed=4):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=0.02, ly=0.04, lz=0.02,
                                            cells=(10, 20, 10), dg=(0, 0, 0))
    return stretch(p), parts, gas


def cell_volumes(p):
    dc = cell_sizes(p)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    k, i, j = np.meshgrid(np.arange(nk), np.arange(ni), np.arange(nj), indexing="ij")
    return (dc[0][i] * dc[1][j] * dc[2][k]).ravel()


def test_grid_is_stretched():
    p, _, _ = bed()
    dc = cell_sizes(p)
    assert dc[0][1:-1].max() / dc[0][1:-1].min() > 3.0
    assert dc[1][1:-1].max() / dc[1][1:-1].min() > 5.0


def test_cells_of_spheres_by_brute_force():
    p, parts, gas = bed()
    f, _ = grid(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    for _ in range(2):
        pos = o.get_particles()[0]
        _, pijk = o.get_cells()
        for a in range(3):
            want = np.searchsorted(f[a], pos[a], side="right")
            assert np.array_equal(pijk[a], want)
        o.run(1, 40)
    o.close()


def test_garg_deposition_single_sphere_closed_form():
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    stretch(p)
    nc = workloads.n_cells(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(workloads.Gas(np.zeros(nc), np.zeros(nc), np.zeros(nc), np.zeros((nc, 3))))
    o.init()
    ep, rop = o.get_gas()
    f, _ = grid(p)
    vol = cell_volumes(p)
    m = 2500.0 * (4.0 / 3.0) * np.pi * 0.5e-3 ** 3
    cell = np.array([np.searchsorted(f[a], x[a], side="right") for a in range(3)])
    zeta = np.array([(x[a] - f[a][cell[a] - 1]) / (f[a][cell[a]] - f[a][cell[a] - 1])
                     for a in range(3)])
    # node (I, J, K) = the upper corner of cell (I, J, K); its volume is the
    # fluid part of its 8 cells (all fluid here), VOL_SURR
    node_w = {}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = ((zeta[0] if a else 1 - zeta[0]) * (zeta[1] if b else 1 - zeta[1]) *
                     (zeta[2] if c else 1 - zeta[2]))
                node_w[tuple(cell - 1 + np.array([a, b, c]))] = w
    want = np.zeros(nc)
    for node, w in node_w.items():
        cells = [(node[0] + da, node[1] + db, node[2] + dd)
                 for da in (0, 1) for db in (0, 1) for dd in (0, 1)]
        vsurr = sum(vol[cell_index(p, *q)] for q in cells)
        for q in cells:
            want[cell_index(p, *q)] += w * m / vsurr
    assert np.max(np.abs(rop[0] - want)) <= 1e-12 * want.max()
    assert np.max(np.abs(ep - (1.0 - want / 2500.0))) <= 1e-15
    o.close()


def test_garg_deposition_conserves_mass():
    p, parts, gas = bed()
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    o.run(1, 20)
    _, rop = o.get_gas()
    _, _, _, state = o.get_particles()
    vol = cell_volumes(p).reshape(p.kmax + 2, p.imax + 2, p.jmax + 2)
    fluid = rop[0].reshape(p.kmax + 2, p.imax + 2, p.jmax + 2)[1:-1, 1:-1, 1:-1]
    mass = (2500.0 * (4.0 / 3.0) * np.pi * parts.radius[state != 0] ** 3).sum()
    got = (fluid * vol[1:-1, 1:-1, 1:-1]).sum()
    assert abs(got - mass) <= 1e-13 * mass
    o.close()


def test_pressure_gradient_linear_and_quadratic():
    p, parts, gas = bed(n=300)
    f, dc = grid(p)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    k, i, j = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), np.arange(1, nj + 1),
                          indexing="ij")
    x = (f[0][i - 1] + 0.5 * dc[0][i - 1]).ravel()
    y = (f[1][j - 1] + 0.5 * dc[1][j - 1]).ravel()
    z = (f[2][k - 1] + 0.5 * dc[2][k - 1]).ravel()
    fluid = ((i >= 2) & (i <= p.imax + 1) & (j >= 2) & (j <= p.jmax + 1) & (k >= 2) &
             (k <= p.kmax + 1)).ravel()
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    a = np.array([30.0, -2500.0, 7.0])
    o.set_gas_pressure(a[0] * x + a[1] * y + a[2] * z)
    pf = o.get_p_force()
    assert np.max(np.abs(pf[fluid] + a)) <= 1e-12 * np.abs(a).max()
    assert not pf[~fluid].any()
    o.set_gas_pressure(3.0e4 * y * y + 1.0e4 * x * x)
    pf = o.get_p_force()
    jj, ii = j.ravel(), i.ravel()
    inner_y = fluid & (jj > 2) & (jj < p.jmax + 1)
    inner_x = fluid & (ii > 2) & (ii < p.imax + 1)
    assert np.allclose(pf[inner_y, 1], -6.0e4 * y[inner_y], rtol=1e-9, atol=1e-6)
    assert np.allclose(pf[inner_x, 0], -2.0e4 * x[inner_x], rtol=1e-9, atol=1e-6)
    o.close()


@pytest.mark.parametrize("scheme,interp_on,width", LINEAR_CASES, ids=LINEAR_IDS)
def test_des1_drag_linear_field_on_a_stretched_grid(scheme, interp_on, width):
    p, parts, gas = linear_case(scheme, interp_on, width, stretch=1.06)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    ep, _ = o.get_gas()
    _, want = linear_expectation(p, parts, ep, o.filter)
    o.step(1)
    fc, _ = o.get_forces()
    o.close()
    scale = np.max(np.abs(want))
    assert np.max(np.abs(fc - want)) <= 1e-12 * scale


def test_gas_side_node_sums_weighted_by_node_volume():
    p, parts, _ = bed()
    nc = workloads.n_cells(p)
    u_gas = np.array([0.3, 1.1, -0.2])
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(workloads.Gas(np.full(nc, u_gas[0]), np.full(nc, u_gas[1]), np.full(nc, u_gas[2]),
                            np.zeros((nc, 3))))
    o.init()
    o.run(1, 20)
    _, vel, _, state = o.get_particles()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    _, am, bm = o.drag_gs_gas()
    vol = cell_volumes(p)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    vn = np.zeros(nc)
    for k in range(1, p.kmax + 2):
        for i in range(1, p.imax + 2):
            for j in range(1, p.jmax + 2):
                for q in [(i + a, j + b, k + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
                    if 2 <= q[0] <= p.imax + 1 and 2 <= q[1] <= p.jmax + 1 and \
                            2 <= q[2] <= p.kmax + 1:
                        vn[cell_index(p, i, j, k)] += vol[cell_index(p, *q)] / 8.0
    live = np.flatnonzero(state != 0)
    fgp = np.array([o.lib.orc_drag_fgp(p.drag_type, ep[pijk[3, q] - 1], p.mu_g0, p.ro_g0,
                                       np.linalg.norm(u_gas - vel[:, q]), 2.0 * parts.radius[q],
                                       np.pi / 6.0 * (2.0 * parts.radius[q]) ** 3)
                    for q in live])
    assert abs((am * vn).sum() - fgp.sum()) <= 1e-12 * fgp.sum()
    for dd in range(3):
        want = (fgp * vel[dd, live]).sum()
        assert abs((bm[dd] * vn).sum() - want) <= 1e-12 * (fgp * np.abs(vel[dd, live])).sum()
    o.close()


def test_default_des_grid_follows_the_smallest_cell():
    p, parts, _ = workloads.settling_bed(n=200, seed=2, box=(0.02, 0.035, 0.02))
    p.imax, p.jmax, p.kmax = 5, 7, 5
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = 0
    stretch(p)
    o = OracleSystem(p)
    d = o.derived()
    width = 3.0 * 2.0 * p.max_radius
    dc = cell_sizes(p)
    for a, (ln, got) in enumerate(((p.xlength, d.dg_imax), (p.ylength, d.dg_jmax),
                                   (p.zlength, d.dg_kmax))):
        assert got == max(int(ln / min(width, dc[a][1:-1].min())), 1)
    o.close()

"""Pins of the oracle's DRAG_GS_GAS0 (drag_gs_des0.f:235-497), the gas side
of the GARG_2012 drag: a single particle against the closed form of its
node weights and F_GDS, and many particles against the conservation of
the drag coefficient and momentum carried onto the nodes.

In a uniform gas field the interpolated gas velocity is that field at
every particle, so each particle's F_GP follows from orc_drag_fgp (pinned
by the textbook drag laws) and the node sums must give back
sum F_GP and sum F_GP v_p once weighted by the node volumes."""
import numpy as np

import _paths  # noqa: F401
from mfixdem import workloads
from oracle_py import OracleSystem
from test_oracle_invariants import cell_index, garg_single

U_GAS = np.array([0.3, 1.1, -0.2])


def uniform_gas(nc):
    return workloads.Gas(np.full(nc, U_GAS[0]), np.full(nc, U_GAS[1]), np.full(nc, U_GAS[2]),
                         np.zeros((nc, 3)))


def node_volumes(p):
    """DES_VOL_NODE of a box of fluid cells: the fluid cells among the 8
    around node (i, j, k), times VOL / 8, for i in [1, imax+1] (etc.)."""
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    vol = (p.xlength / p.imax) * (p.ylength / p.jmax) * (p.zlength / p.kmax)
    fl = np.zeros((nk + 2, ni + 2, nj + 2))
    fl[2:p.kmax + 2, 2:p.imax + 2, 2:p.jmax + 2] = 1.0     # [K, I, J], 1-based + 1 pad
    vn = np.zeros(ni * nj * nk)
    for k in range(1, p.kmax + 2):
        for i in range(1, p.imax + 2):
            for j in range(1, p.jmax + 2):
                cnt = fl[k:k + 2, i:i + 2, j:j + 2].sum()
                vn[cell_index(p, i, j, k)] = cnt * vol / 8.0
    return vn


def test_single_particle_closed_form():
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    vp = np.array([0.05, -0.4, 0.12])
    parts.vel[:, 0] = vp
    nc = workloads.n_cells(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(uniform_gas(nc))
    o.init()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    f_gds, am, bm = o.drag_gs_gas()
    d = np.array([p.xlength / p.imax, p.ylength / p.jmax, p.zlength / p.kmax])
    vol = d.prod()
    fgp = o.lib.orc_drag_fgp(p.drag_type, ep[pijk[3, 0] - 1], p.mu_g0, p.ro_g0,
                             np.linalg.norm(U_GAS - vp), 1.0e-3, np.pi / 6.0 * 1.0e-9)
    cell = (np.floor(x / d) + 2).astype(int)
    zeta = x / d - (cell - 2)
    want_am = np.zeros(nc)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = ((zeta[0] if a else 1 - zeta[0]) * (zeta[1] if b else 1 - zeta[1]) *
                     (zeta[2] if c else 1 - zeta[2]))
                node = cell - 1 + np.array([a, b, c])
                want_am[cell_index(p, *node)] = fgp * w / vol
    assert np.max(np.abs(am - want_am)) <= 1e-13 * want_am.max()
    for dd in range(3):
        assert np.max(np.abs(bm[dd] - want_am * vp[dd])) <= 1e-13 * want_am.max()
    # F_GDS: the own cell sees all 8 nodes, its east neighbour the 4 east ones
    own = cell_index(p, *cell)
    east = cell_index(p, cell[0] + 1, cell[1], cell[2])
    assert abs(f_gds[own] - fgp / (8.0 * vol)) <= 1e-13 * f_gds[own]
    assert abs(f_gds[east] - fgp * zeta[0] / (8.0 * vol)) <= 1e-13 * f_gds[own]
    ghost = cell_index(p, 1, cell[1], cell[2])
    assert f_gds[ghost] == 0.0
    o.close()


def test_node_sums_conserve_drag_and_momentum():
    p, parts, _ = workloads.fluidised_bed(n=3000, seed=4, lx=0.02, ly=0.04, lz=0.02,
                                          cells=(10, 20, 10), dg=(0, 0, 0))
    nc = workloads.n_cells(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(uniform_gas(nc))
    o.init()
    o.run(1, 30)
    pos, vel, _, state = o.get_particles()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    f_gds, am, bm = o.drag_gs_gas()
    vn = node_volumes(p)
    live = state != 0
    fgp = np.array([o.lib.orc_drag_fgp(p.drag_type, ep[pijk[3, q] - 1], p.mu_g0, p.ro_g0,
                                       np.linalg.norm(U_GAS - vel[:, q]),
                                       2.0 * parts.radius[q],
                                       np.pi / 6.0 * (2.0 * parts.radius[q]) ** 3)
                    for q in np.flatnonzero(live)])
    assert abs((am * vn).sum() - fgp.sum()) <= 1e-12 * fgp.sum()
    for dd in range(3):
        want = (fgp * vel[dd, live]).sum()
        scale = (fgp * np.abs(vel[dd, live])).sum()
        assert abs((bm[dd] * vn).sum() - want) <= 1e-12 * scale
    # every node the particles reach is a node of the deposit range
    assert np.all(am[vn == 0.0] == 0.0)
    assert np.all(f_gds >= 0.0)
    o.close()


def filter_params(p, scheme, interp_on, width):
    from mfixdem import abi
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_mean_fields = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    return p


def test_filter_gas_drag_single_particle_without_interpolation():
    """DRAG_GS_GAS1 with DES_INTERP_ON off: everything lands in the
    particle's own cell, F_GP/VOL and F_GP v_p/VOL; no DRAG_AM."""
    from mfixdem import abi
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    filter_params(p, abi.INTERP_NONE, 0, None)
    vp = np.array([0.05, -0.4, 0.12])
    parts.vel[:, 0] = vp
    nc = workloads.n_cells(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(uniform_gas(nc))
    o.init()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    own = pijk[3, 0] - 1
    f_gds, am, bm = o.drag_gs_gas()
    vol = (p.xlength / p.imax) * (p.ylength / p.jmax) * (p.zlength / p.kmax)
    fgp = o.lib.orc_drag_fgp(p.drag_type, ep[own], p.mu_g0, p.ro_g0,
                             np.linalg.norm(U_GAS - vp), 1.0e-3, np.pi / 6.0 * 1.0e-9)
    assert f_gds[own] == fgp * (1.0 / vol)
    assert np.count_nonzero(f_gds) == 1 and not am.any()
    for dd in range(3):
        assert bm[dd, own] == fgp * vp[dd] * (1.0 / vol)
    o.close()


import pytest  # noqa: E402


@pytest.mark.parametrize("scheme,interp_on,width", [
    (1, 0, None), (2, 1, 1.6e-3), (3, 1, 1.6e-3), (4, 1, None)],
    ids=["none", "square_dpvm", "gauss_dpvm", "linear_hat"])
def test_filter_gas_drag_conserves_drag_and_momentum(scheme, interp_on, width):
    """DRAG_GS_GAS1: each particle's filter weights sum to 1, so the cell sums
    weighted by VOL give back sum F_GP and sum F_GP v_p, with F_GP from the
    particle's own filter interpolation of EP_G (uniform gas velocity)."""
    p, parts, _ = workloads.fluidised_bed(n=1500, seed=5, lx=0.02, ly=0.04, lz=0.02,
                                          cells=(10, 20, 10), dg=(0, 0, 0))
    filter_params(p, scheme, interp_on, width)
    nc = workloads.n_cells(p)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(uniform_gas(nc))
    o.init()
    o.run(1, 20)
    _, vel, _, state = o.get_particles()
    ep, _ = o.get_gas()
    _, pijk = o.get_cells()
    f_gds, am, bm = o.drag_gs_gas()
    vol = (p.xlength / p.imax) * (p.ylength / p.jmax) * (p.zlength / p.kmax)
    fgp, fv = 0.0, np.zeros(3)
    scale = np.zeros(3)
    for q in np.flatnonzero(state == 1):
        own = pijk[3, q] - 1
        if interp_on:
            cells, w = o.filter(q)
            lepg = 0.0
            for c, wc in zip(cells, w):
                lepg = lepg + ep[c] * wc
            if lepg < 1.0e-6:
                lepg = ep[own]
        else:
            lepg = ep[own]
        f = o.lib.orc_drag_fgp(p.drag_type, lepg, p.mu_g0, p.ro_g0,
                               np.linalg.norm(U_GAS - vel[:, q]), 2.0 * parts.radius[q],
                               np.pi / 6.0 * (2.0 * parts.radius[q]) ** 3)
        fgp += f
        fv += f * vel[:, q]
        scale += f * np.abs(vel[:, q])
    assert not am.any()
    assert abs(f_gds.sum() * vol - fgp) <= 1e-12 * fgp
    for dd in range(3):
        assert abs(bm[dd].sum() * vol - fv[dd]) <= 1e-12 * scale[dd]
    o.close()

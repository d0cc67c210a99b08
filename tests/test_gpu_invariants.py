"""GPU side of the invariant checks: the HIP library meets the same
closed forms and conservation laws that pin the oracle (momentum and
angular momentum, exact history carry over rebuilds, GARG_2012 deposition
and linear-field drag interpolation), and matches the oracle bit for bit
where no pow enters."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem
from oracle_py import OracleSystem
from test_oracle_invariants import (cell_index, cloud, colliding_pairs, garg_single,
                                    momenta)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", ["LSD", "HERTZ"])
def test_cloud_bit_exact_and_momenta_conserved(model):
    p, parts = cloud(model)
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        s.init()
        s.run(1, 300)
    pg, vg, wg, _ = g.get_particles()
    po, vo, wo, _ = o.get_particles()
    g.close()
    o.close()
    assert np.array_equal(pg, po) and np.array_equal(vg, vo) and np.array_equal(wg, wo)
    l0, a0, sl, sa = momenta(parts.pos, parts.vel, parts.omega, parts.radius, parts.ro_sol)
    l1, a1, _, _ = momenta(pg, vg, wg, parts.radius, parts.ro_sol)
    assert np.max(np.abs(l1 - l0)) <= 1e-14 * sl
    assert np.max(np.abs(a1 - a0)) <= 1e-13 * sa


def test_history_carry_is_exact_across_rebuilds_on_gpu():
    runs = {}
    for nsn in (1, 7, 1000):
        p, parts = colliding_pairs()
        p.neighbor_search_n = nsn
        g = DemSystem(p)
        g.set_particles(parts)
        g.init()
        g.run(1, 400)
        runs[nsn] = g.get_particles()[:3]
        g.close()
    for nsn in (1, 7):
        for a, b in zip(runs[nsn], runs[1000]):
            assert np.array_equal(a, b)


def test_garg_single_particle_on_gpu_matches_closed_forms():
    """One particle: deposition (fixed-point sums, 1e-12 of the largest
    cell value) and the drag of a linear gas field (pow: 1e-12)."""
    x = np.array([0.0051, 0.00737, 0.0069])
    p, parts = garg_single(x)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    d = np.array([p.xlength / p.imax, p.ylength / p.jmax, p.zlength / p.kmax])
    k, i, j = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), np.arange(1, nj + 1),
                          indexing="ij")
    face = [(i - 1) * d[0], (j - 1) * d[1], (k - 1) * d[2]]
    ctr = [(i - 1.5) * d[0], (j - 1.5) * d[1], (k - 1.5) * d[2]]
    a0 = np.array([0.05, 0.6, -0.03])
    A = np.array([[3.0, -2.0, 1.5], [4.0, 7.0, -5.0], [-1.0, 2.5, 6.0]])

    def field(c, xs):
        return np.ascontiguousarray(
            (a0[c] + A[c, 0] * xs[0] + A[c, 1] * xs[1] + A[c, 2] * xs[2]).ravel())
    gas = workloads.Gas(field(0, [face[0], ctr[1], ctr[2]]), field(1, [ctr[0], face[1], ctr[2]]),
                        field(2, [ctr[0], ctr[1], face[2]]), np.zeros((ni * nj * nk, 3)))
    g, o = DemSystem(p), OracleSystem(p)
    g.set_debug(abi.DEBUG_STORE_FORCES)
    for s in (g, o):
        s.set_particles(parts)
        s.set_gas(gas)
        s.init()
    epg, ropg = g.get_gas()
    epo, ropo = o.get_gas()
    # deposition closed form (the oracle meets it to 1e-12)
    vol = d.prod()
    m = 2500.0 * (4.0 / 3.0) * np.pi * 0.5e-3 ** 3
    cell = (np.floor(x / d) + 2).astype(int)
    zeta = x / d - (cell - 2)
    want = np.zeros(ni * nj * nk)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = ((zeta[0] if a else 1 - zeta[0]) * (zeta[1] if b else 1 - zeta[1]) *
                     (zeta[2] if c else 1 - zeta[2]))
                node = cell - 1 + np.array([a, b, c])
                for da in (0, 1):
                    for db in (0, 1):
                        for dc in (0, 1):
                            want[cell_index(p, node[0] + da, node[1] + db, node[2] + dc)] += \
                                w * m / (8.0 * vol)
    fluid = np.zeros((nk, ni, nj), dtype=bool)
    fluid[1:-1, 1:-1, 1:-1] = True
    fluid = fluid.ravel()
    assert np.max(np.abs(ropg[0][fluid] - want[fluid])) <= 1e-12 * want.max()
    assert np.max(np.abs(epg[fluid] - epo[fluid])) <= 1e-12
    # drag of a linear field at the particle
    _, pijk = g.get_cells()
    eps = epg[pijk[3, 0] - 1]
    g.step(1)
    o.step(1)
    fg, _ = g.get_forces()
    fo, _ = o.get_forces()
    ug = a0 + A @ x
    fgp = o.lib.orc_drag_fgp(p.drag_type, eps, p.mu_g0, p.ro_g0, float(np.linalg.norm(ug)),
                             1.0e-3, np.pi / 6.0 * 1.0e-9)
    assert np.allclose(fg[:, 0], fgp * ug, rtol=1e-12, atol=0)
    assert np.allclose(fg[:, 0], fo[:, 0], rtol=1e-12, atol=0)
    g.close()
    o.close()


def test_gas_side_drag_conserved_at_full_size():
    """DRAG_GS_GAS0 at the benchmark's size (1M spheres, 64x128x64 cells):
    each particle's node weights sum to 1, so the node sums weighted by
    DES_VOL_NODE give back sum F_GP and sum F_GP v_p (to 1e-11), with F_GP
    from the drag law (orc_drag_fgp, pinned by the textbook laws) at each
    particle's cell voidage in a uniform gas field."""
    from oracle_py import load_oracle
    p, parts, _ = workloads.fluidised_bed(n=1_000_000)
    nc = workloads.n_cells(p)
    u = np.array([0.3, 1.1, -0.2])
    gas = workloads.Gas(np.full(nc, u[0]), np.full(nc, u[1]), np.full(nc, u[2]),
                        np.zeros((nc, 3)))
    g = DemSystem(p)
    g.set_particles(parts)
    g.set_gas(gas)
    g.init()
    g.run(1, 20)
    f_gds, am, bm = g.drag_gs_gas()
    _, vel, _, state = g.get_particles()
    ep, _ = g.get_gas()
    _, pijk = g.get_cells()
    g.close()
    live = np.flatnonzero(state == 1)
    eps = ep[pijk[3, live] - 1]
    vrel = np.linalg.norm(u[:, None] - vel[:, live], axis=0)
    d = 2.0 * parts.radius[live]
    fgp_law = load_oracle().orc_drag_fgp
    fgp = np.array([fgp_law(p.drag_type, e, p.mu_g0, p.ro_g0, w, dd, np.pi / 6.0 * dd ** 3)
                    for e, w, dd in zip(eps.tolist(), vrel.tolist(), d.tolist())])
    # DES_VOL_NODE of a box: VOL/8 per fluid cell of the node's 2x2x2 block
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    vol = (p.xlength / p.imax) * (p.ylength / p.jmax) * (p.zlength / p.kmax)
    fl = np.zeros((nk, ni, nj))
    fl[1:p.kmax + 1, 1:p.imax + 1, 1:p.jmax + 1] = 1.0
    cnt = sum(fl[a:nk - 1 + a, b:ni - 1 + b, c:nj - 1 + c]
              for a in (0, 1) for b in (0, 1) for c in (0, 1))
    vn = np.zeros((nk, ni, nj))
    vn[:nk - 1, :ni - 1, :nj - 1] = cnt * vol / 8.0
    vn = vn.ravel()
    assert np.all(am[vn == 0.0] == 0.0) and np.all(f_gds >= 0.0)
    assert abs((am * vn).sum() - fgp.sum()) <= 1e-11 * fgp.sum()
    for k in range(3):
        want = (fgp * vel[k, live]).sum()
        assert abs((bm[k] * vn).sum() - want) <= 1e-11 * (fgp * np.abs(vel[k, live])).sum()

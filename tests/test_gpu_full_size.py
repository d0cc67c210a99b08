"""The workloads at their full sizes on the GPU against the oracle.

* Settling bed, 20,000 monodisperse LSD spheres in a 0.02 x 0.08 x 0.02 m
  box, 2,000 steps (80 neighbour rebuilds): positions, velocities, omega,
  NEIGHBOR_INDEX, NEIGHBORS and PFT_NEIGHBOR bit for bit at four steps.
* Fluidised bed, 1M polydisperse Hertz spheres with the frozen gas field,
  grown by 1/0.93 after placement and given random velocities so that
  about half of the listed pairs are in contact from the first step:
  DG_PIJK, PIJK and the pair lists bit for bit after set-up; FP64 FC and
  TOW of step 1 within 1e-12 of the largest value (the drag law's pow is
  libm's on the CPU and the device's on the GPU); the mixed-precision
  FC within 1e-4 of the largest value; ten FP64 steps within 1e-12 m.
* The 256-member parameter-sweep ensemble (2k-200k spheres each): a sample
  of members bit for bit against the same systems run alone."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.ensemble import Ensemble
from mfixdem.system import DemSystem
from oracle_py import OracleSystem

pytestmark = pytest.mark.gpu


def _pairs_equal(g, o):
    gi, gn, gp = g.get_pairs()
    oi, on, op = o.get_pairs()
    assert np.array_equal(gi, oi), "NEIGHBOR_INDEX differs"
    assert np.array_equal(gn, on), "NEIGHBORS differs"
    return gp, op


def test_settling_bed_full_size_2000_steps_bit_exact():
    p, parts, _ = workloads.settling_bed()   # 20k spheres, 0.02 x 0.08 x 0.02 m
    assert parts.n == 20_000
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        s.init()
    nn = 1
    for stop in (250, 1000, 1500, 2000):
        for s in (g, o):
            s.run(nn, stop - nn + 1)
        nn = stop + 1
        for a, b, what in zip(g.get_particles()[:3], o.get_particles()[:3],
                              ("pos", "vel", "omega")):
            assert np.array_equal(a, b), f"{what} differs after step {stop}"
        gp, op = _pairs_equal(g, o)
        assert np.array_equal(gp, op), f"PFT differs after step {stop}"
    # spheres collide during the fall (0.02 s): tangential histories are live
    assert np.count_nonzero(op) > 100
    g.close()
    o.close()


@pytest.fixture(scope="module")
def dense_1m():
    p, parts, gas = workloads.fluidised_bed(n=1_000_000, dg=(0, 0, 0))
    p, parts = workloads.jostled(p, parts, seed=3, shrink=0.93, v0=0.2)
    return p, parts, gas


def _run_one_step(p, parts, gas, oracle=False):
    s = OracleSystem(p) if oracle else DemSystem(p)
    s.set_particles(parts)
    s.set_gas(gas)
    s.set_debug(abi.DEBUG_STORE_FORCES)
    s.init()
    s.step(1)
    return s


def test_fluidised_bed_1m_fp64_and_mixed_against_oracle(dense_1m):
    p, parts, gas = dense_1m
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.set_debug(abi.DEBUG_STORE_FORCES)
    o.init()
    g = DemSystem(p)
    g.set_particles(parts)
    g.set_gas(gas)
    g.set_debug(abi.DEBUG_STORE_FORCES)
    g.init()
    gdg, gpijk = g.get_cells()
    odg, opijk = o.get_cells()
    assert np.array_equal(gdg, odg) and np.array_equal(gpijk, opijk)
    _pairs_equal(g, o)
    g.step(1)
    o.step(1)
    fo, to = o.get_forces()
    fg, tg = g.get_forces()
    # contact-rich: many spheres overlap at step 1
    oi, on, _ = o.get_pairs()
    rows = np.repeat(np.arange(parts.n), np.diff(np.concatenate([[1], oi])).astype(np.int64))
    d = np.sqrt(((parts.pos[:, rows] - parts.pos[:, on - 1]) ** 2).sum(0))
    contacts = int(np.count_nonzero(parts.radius[rows] + parts.radius[on - 1] > d))
    assert contacts > 300_000, contacts
    fmax, tmax = np.abs(fo).max(), np.abs(to).max()
    assert np.abs(fg - fo).max() <= 1e-12 * fmax
    assert np.abs(tg - to).max() <= 1e-12 * tmax
    # mixed precision: FP32 contact and drag laws, FP64 state
    pm = abi.Params.from_buffer_copy(p)
    pm.precision = abi.MIXED
    m = _run_one_step(pm, parts, gas)
    fm, tm = m.get_forces()
    m.close()
    err_f = np.abs(fm - fo).max() / fmax
    err_t = np.abs(tm - to).max() / tmax
    print(f"1M spheres, {contacts} contacts: mixed FC error {err_f:.3e}, TOW error {err_t:.3e} "
          f"of the largest value")
    assert err_f <= 1e-4 and err_t <= 1e-4
    # ten FP64 steps (no rebuild in between): trajectories stay together
    g.set_debug(0)
    o.set_debug(0)
    g.run(2, 9)
    o.run(2, 9)
    gp, gv, _, _ = g.get_particles()
    op, ov, _, _ = o.get_particles()
    assert np.abs(gp - op).max() <= 1e-12
    assert np.abs(gv - ov).max() <= 1e-9 * np.abs(ov).max()
    g.close()
    o.close()


def test_full_sweep_ensemble_members_match_single_runs():
    ms = workloads.sweep_ensemble(nsys=256, seed=11,
                                  member=lambda *a: workloads.lattice_member(
                                      *a, precision=abi.FP64))
    ns = np.array([p.max_pip for p, _ in ms])
    assert len(ms) == 256 and ns.min() >= 2000 and ns.max() <= 200_000
    e = Ensemble([p for p, _ in ms])
    for s, (_, parts) in enumerate(ms):
        e.set_particles(s, parts)
    e.init()
    e.run(1, 60)   # rebuilds at NN = 1, 25, 50
    rng = np.random.default_rng(4)
    sample = sorted({int(ns.argmin()), int(ns.argmax()), 0, 255,
                     *rng.choice(256, 4, replace=False).tolist()})
    for s in sample:
        p, parts = ms[s]
        g = DemSystem(p)
        g.set_particles(parts)
        g.init()
        g.run(1, 60)
        for a, b in zip(e.get_particles(s), g.get_particles()):
            assert np.array_equal(a, b), f"member {s} (n = {p.max_pip}) differs"
        g.close()
    e.close()

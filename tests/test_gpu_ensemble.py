"""Ensembles (mfixdem_ensemble_*): many independent systems stepped by the
same launches.  Each member must evolve exactly as it does in its own
handle (and so as the oracle): same pair sets, same FP64 trajectory bit for
bit, whatever the other members' sizes, parameters and time steps."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.ensemble import Ensemble
from mfixdem.system import DemSystem
from oracle_py import OracleSystem

pytestmark = pytest.mark.gpu


def members(nsys=6, seed=5):
    out = []
    for s, (p, parts) in enumerate(workloads.sweep_ensemble(nsys=nsys, seed=seed, nmin=300,
                                                            nmax=3000)):
        p.integrator = abi.ADAMS_BASHFORTH if s % 2 else abi.EULER
        p, parts = workloads.jostled(p, parts, seed=seed + s, v0=0.2, w0=100.0)
        out.append((p, parts))
    return out


def test_ensemble_members_match_single_systems_bit_exact():
    ms = members()
    ens = Ensemble([p for p, _ in ms])
    for s, (_, parts) in enumerate(ms):
        ens.set_particles(s, parts)
    ens.init()
    ens.run(1, 60)
    for s, (p, parts) in enumerate(ms):
        g = DemSystem(p)
        g.set_particles(parts)
        g.init()
        g.run(1, 60)
        for a, b in zip(ens.get_particles(s), g.get_particles()):
            assert np.array_equal(a, b), f"member {s} differs"
        g.close()
    p, parts = ms[0]
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    o.run(1, 60)
    for a, b in zip(ens.get_particles(0), o.get_particles()):
        assert np.array_equal(a, b)
    ens.close()


def test_ensemble_step_by_step_equals_graph_replay():
    ms = members(nsys=4, seed=9)
    a, b = Ensemble([p for p, _ in ms]), Ensemble([p for p, _ in ms])
    for e in (a, b):
        for s, (_, parts) in enumerate(ms):
            e.set_particles(s, parts)
        e.init()
    a.run(1, 55)
    for nn in range(1, 56):
        b.step(nn)
    for s in range(len(ms)):
        for x, y in zip(a.get_particles(s), b.get_particles(s)):
            assert np.array_equal(x, y)
    a.close()
    b.close()


def test_ensemble_rejects_mixed_launch_choices():
    ms = members(nsys=2)
    ps = [p for p, _ in ms]
    ps[1].neighbor_search_n = 10
    with pytest.raises(abi.MfixDemError):
        Ensemble(ps)


def test_tuning_choices_give_identical_results():
    """Every (search-cell size, row kernel) choice gives the same state bit
    for bit, for a single system and an ensemble."""
    from mfixdem import autotune
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=(0.01, 0.035, 0.01))
    p, parts = workloads.jostled(p, parts, seed=4)
    ref = None
    for c, r in autotune.candidates():
        g = DemSystem(p)
        g.set_particles(parts)
        g.set_tuning(c, r)
        g.init()
        g.run(1, 55)
        st = g.get_particles()
        g.close()
        if ref is None:
            ref = st
        else:
            for a, b in zip(st, ref):
                assert np.array_equal(a, b), (c, r)
    ms = members(nsys=3, seed=13)
    outs = []
    for c, r in [(1.0, abi.ROWS_THREAD), (1.5, abi.ROWS_WAVE256)]:
        e = Ensemble([q for q, _ in ms])
        for s, (_, pts) in enumerate(ms):
            e.set_particles(s, pts)
        e.set_tuning(c, r)
        e.init()
        e.run(1, 55)
        outs.append([e.get_particles(s) for s in range(len(ms))])
        e.close()
    for s in range(len(ms)):
        for a, b in zip(outs[0][s], outs[1][s]):
            assert np.array_equal(a, b)


def test_autotuner_picks_a_candidate():
    from mfixdem import autotune
    ms = members(nsys=3, seed=17)

    def make():
        e = Ensemble([q for q, _ in ms])
        for s, (_, pts) in enumerate(ms):
            e.set_particles(s, pts)
        return e

    best, timings = autotune.autotune(make, steps=30,
                                      cands=[(1.0, abi.ROWS_WAVE192), (1.25, abi.ROWS_WAVE128)])
    assert best in [(1.0, abi.ROWS_WAVE192), (1.25, abi.ROWS_WAVE128)]
    assert len(timings) == 2 and all(t > 0 for t in timings.values())


def test_skin_tuner_keeps_every_contact():
    """The opt-in skin tuner (FACTOR_RLM with the longest safe rebuild
    interval) on the GPU: its choice follows a run that rebuilds every step
    (complete contact detection) to 1e-12 m over 200 steps; on this bed the
    two are bit-identical (no DES-cell changes reorder the sums)."""
    from mfixdem import autotune
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=(0.01, 0.035, 0.01))
    p, parts = workloads.jostled(p, parts, seed=3)
    vmax = float(np.sqrt((parts.vel ** 2).sum(0)).max())
    r_min = float(parts.radius.min())

    def make(q):
        s = DemSystem(q)
        s.set_particles(parts)
        return s
    (f, n), timings = autotune.autotune_skin(make, p, r_min, vmax, factors=(1.2, 1.35, 1.5),
                                             steps=100)
    assert len(timings) == 3 and n >= 1
    runs = []
    for ff, nn in [(f, n), (p.factor_rlm, 1)]:
        q = abi.Params.from_buffer_copy(p)
        q.factor_rlm = ff
        q.neighbor_search_n = nn
        s = make(q)
        s.init()
        s.run(1, 200)
        runs.append(s.get_particles()[0])
        s.close()
    assert np.max(np.abs(runs[0] - runs[1])) <= 1e-12

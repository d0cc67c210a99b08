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

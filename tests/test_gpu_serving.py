"""Serving loop on the GPU: every request completes, admission respects the
HBM budget, and a request served among others ends in exactly the state it
reaches when run alone (streams do not interfere)."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import serving, workloads
from mfixdem.system import DemSystem

pytestmark = pytest.mark.gpu


def test_served_requests_match_running_alone():
    reqs = serving.poisson_requests(16, rate=200.0, make_member=workloads.lattice_member,
                                    nmin=2000, nmax=20000, smin=40, smax=160, seed=5)
    model = serving.FootprintModel(lambda n: workloads.lattice_member(n, 1e3, 0.9, 0.3, 1))
    for r in reqs:
        r.footprint = model(r.params.max_pip)
    budget = 3 * max(r.footprint for r in reqs)
    srv = serving.SweepServer(budget, workers=8, keep_results=True)
    done = srv.serve(reqs)
    rep = serving.report(done, srv.makespan)
    assert rep["failed"] == 0, [r.error for r in done if r.error]
    assert srv.max_in_use <= budget
    assert rep["p99_latency_s"] >= rep["p50_latency_s"] > 0
    for r in done[:4]:
        s = DemSystem(r.params)
        s.set_particles(r.parts)
        s.init()
        s.run(1, r.nsteps)
        for a, b in zip(s.get_particles(), r.result):
            assert np.array_equal(a, b), f"request {r.rid}"
        s.close()

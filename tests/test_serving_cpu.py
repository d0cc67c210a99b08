"""Host logic of the serving loop (no GPU): admission by footprint in
arrival order, continuous admission as requests finish, latency report."""
import time

import numpy as np

import _paths  # noqa: F401
from mfixdem import abi, serving


class FakeServer(serving.SweepServer):
    """Runs each request as a sleep proportional to its work."""

    def _execute(self, req, t0):
        time.sleep(req.nsteps * 1e-4)
        req.finished = time.perf_counter() - t0
        with self._cv:
            self._in_use -= req.footprint
            self._running -= 1
            self._cv.notify_all()


def reqs(n=30, seed=1):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(0.005, n))
    out = []
    for i in range(n):
        p = abi.default_params()
        p.max_pip = 1000
        r = serving.Request(i, float(t[i]), p, None, int(rng.integers(50, 300)))
        r.footprint = int(rng.integers(1, 5))
        out.append(r)
    return out


def test_budget_and_worker_limits_hold():
    s = FakeServer(hbm_budget=6, workers=3)
    rs = s.serve(reqs())
    assert s.max_in_use <= 6 and s.max_running <= 3
    assert all(r.finished >= r.admitted >= r.arrival - 1e-9 for r in rs)
    # FIFO admission
    adm = [r.admitted for r in sorted(rs, key=lambda r: r.arrival)]
    assert adm == sorted(adm)


def test_report_percentiles():
    s = FakeServer(hbm_budget=100, workers=8)
    rs = s.serve(reqs(n=40, seed=3))
    rep = serving.report(rs, s.makespan)
    lat = sorted(r.finished - r.arrival for r in rs)
    assert rep["requests"] == 40 and rep["failed"] == 0
    assert abs(rep["p50_latency_s"] - float(np.percentile(lat, 50))) < 1e-12
    assert rep["p99_latency_s"] >= rep["p50_latency_s"]


def test_oversized_request_is_refused():
    s = FakeServer(hbm_budget=3, workers=2)
    rs = reqs(n=3)
    rs[1].footprint = 10
    try:
        s.serve(rs)
    except ValueError:
        return
    raise AssertionError("expected refusal")

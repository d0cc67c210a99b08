"""Serving loop for parameter-sweep DEM runs on one GPU.

A request is one DEM system (parameters, particles, number of steps) with
an arrival time.  The server admits queued requests in arrival order while
the HBM footprint of everything running stays within a budget, runs each
admitted request in its own handle and HIP stream from a worker thread
(ctypes releases the GIL, so the requests' launches overlap on the GPU),
and advances every request in segments of `segment` steps (one neighbour
rebuild interval).  A finished request leaves at its segment boundary and
frees its memory, and the next queued request is admitted at once rather
than when a whole batch drains (continuous batching).  Latency = finish -
arrival; the report gives p50 / p99, queueing delay and throughput.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time

import numpy as np

from . import abi
from .system import DemSystem


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    params: abi.Params        # None until built (lazy requests)
    parts: object             # workloads.Particles; None until built
    nsteps: int
    footprint: int = 0        # estimated device bytes
    admitted: float = float("nan")
    finished: float = float("nan")
    error: str = ""
    result: object = None     # final (pos, vel, omega, state) when kept
    n: int = 0                # particles (params.max_pip once built)
    make: object = None       # () -> (params, parts), called by the worker that runs it

    def __post_init__(self):
        if self.params is not None and not self.n:
            self.n = int(self.params.max_pip)

    @property
    def work(self) -> float:
        return float(self.n) * self.nsteps


class FootprintModel:
    """Device bytes of a handle as a + b * N, fitted from two probe handles
    of the request's kind (mfixdem_get_device_bytes after init)."""

    def __init__(self, make_probe, n_small=2000, n_large=20000):
        f = []
        for n in (n_small, n_large):
            p, parts = make_probe(n)
            s = DemSystem(p)
            s.set_particles(parts)
            s.init()
            f.append(s.device_bytes())
            s.close()
        self.b = (f[1] - f[0]) / (n_large - n_small)
        self.a = f[0] - self.b * n_small

    def __call__(self, n: int) -> int:
        return int(1.25 * (self.a + self.b * n))   # room for neighbour-list growth


class SweepServer:
    def __init__(self, hbm_budget: int, workers: int = 8, segment: int = 25, log=None,
                 keep_results: bool = False):
        self.keep_results = keep_results
        self.budget = int(hbm_budget)
        self.workers = workers
        self.segment = segment
        self.log = log
        self._lock = threading.Lock()
        self._cv = threading.Condition(self._lock)
        self._in_use = 0
        self._running = 0
        self.max_running = 0
        self.max_in_use = 0

    # ------------------------------------------------------------ worker side
    def _execute(self, req: Request, t0: float):
        try:
            if req.params is None:   # built on admission: the queue holds no particle data
                req.params, req.parts = req.make()
            s = DemSystem(req.params)
            try:
                # several handles are driven from different threads: no
                # stream capture (a pure-DEM step is one kernel anyway)
                s.set_debug(abi.DEBUG_NO_GRAPH)
                s.set_particles(req.parts)
                s.init()
                nn, left = 1, req.nsteps
                while left > 0:
                    k = min(self.segment, left)
                    s.run(nn, k)
                    nn += k
                    left -= k
                s.synchronize()
                if self.keep_results:
                    req.result = s.get_particles()
            finally:
                s.close()
            if not self.keep_results:
                req.params, req.parts = None, None   # a long stream keeps no particle data
        except Exception as e:  # a failed request is reported, the server goes on
            req.error = str(e)
        req.finished = time.perf_counter() - t0
        with self._cv:
            self._in_use -= req.footprint
            self._running -= 1
            self._cv.notify_all()

    def _worker(self, jobs: "queue.Queue", t0: float):
        while True:
            req = jobs.get()
            if req is None:
                return
            self._execute(req, t0)

    # -------------------------------------------------------------- dispatch
    def serve(self, requests):
        """Run the requests (sorted by arrival); returns them with times."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.footprint > self.budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B > budget {self.budget}")
        jobs: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()
        pool = [threading.Thread(target=self._worker, args=(jobs, t0), daemon=True)
                for _ in range(self.workers)]
        for th in pool:
            th.start()
        for r in reqs:
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            with self._cv:
                # FIFO: the head of the queue waits for memory and a worker
                while (self._in_use + r.footprint > self.budget or
                       self._running >= self.workers):
                    self._cv.wait()
                self._in_use += r.footprint
                self._running += 1
                self.max_running = max(self.max_running, self._running)
                self.max_in_use = max(self.max_in_use, self._in_use)
                r.admitted = time.perf_counter() - t0
            jobs.put(r)
        for _ in pool:
            jobs.put(None)
        for th in pool:
            th.join()
        self.makespan = time.perf_counter() - t0
        return reqs


def report(reqs, makespan):
    ok = [r for r in reqs if not r.error]
    lat = np.array([r.finished - r.arrival for r in ok])
    wait = np.array([r.admitted - r.arrival for r in ok])
    work = sum(r.work for r in ok)
    return {
        "requests": len(reqs), "failed": len(reqs) - len(ok),
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_latency_s": float(lat.mean()) if len(lat) else None,
        "p99_queue_wait_s": float(np.percentile(wait, 99)) if len(wait) else None,
        "particle_steps_per_s": work / makespan if makespan > 0 else None,
        "makespan_s": makespan,
    }


def poisson_requests(n_req, rate, make_member, nmin, nmax, smin, smax, seed=21, lazy=False):
    """n_req requests with Poisson arrivals at `rate` per second; N
    log-uniform in [nmin, nmax], steps uniform in [smin, smax], KN
    log-uniform in [1e2, 1e4], EN ~ U[0.7, 0.95], MEW ~ U[0.1, 0.5].
    lazy: each request's system is built by the worker that runs it (a
    long stream holds only its parameters until then)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for i in range(n_req):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        kn = float(np.exp(rng.uniform(np.log(1e2), np.log(1e4))))
        en = float(rng.uniform(0.7, 0.95))
        mew = float(rng.uniform(0.1, 0.5))
        steps = int(rng.integers(smin, smax + 1))
        args = (n, kn, en, mew, seed * 100000 + i)
        if lazy:
            out.append(Request(i, float(t[i]), None, None, steps, n=n,
                               make=lambda a=args: make_member(*a)))
        else:
            p, parts = make_member(*args)
            out.append(Request(i, float(t[i]), p, parts, steps))
    return out

"""Autotuner for the result-invariant performance choices of a handle
(mfixdem_set_tuning): the neighbour-search cell size (multiple of the
largest cut-off) and the neighbour-row kernel (thread per particle, or one
wavefront per search cell with an LDS candidate tile of 128 / 192 / 256).

The neighbour-list skin of the model, FACTOR_RLM, fixes the pair set the
reference builds and is therefore not a tuning choice: every candidate
here gives bit-identical results (the test suite checks this), so the tuner
may pick freely.  A candidate is timed on a fresh copy of the workload over
a run that contains at least two neighbour rebuilds.
"""
from __future__ import annotations

import json
import os
import time

from . import abi

CELL_SCALES = (1.0, 1.25, 1.5)
ROW_KERNELS = (abi.ROWS_THREAD, abi.ROWS_WAVE128, abi.ROWS_WAVE192, abi.ROWS_WAVE256)


def candidates(cell_scales=CELL_SCALES, row_kernels=ROW_KERNELS):
    return [(c, r) for c in cell_scales for r in row_kernels]


def time_candidate(make, cell_scale, row_kernel, steps):
    """Seconds for `steps` DES steps (from NN = 1) of a fresh handle."""
    h = make()
    try:
        h.set_tuning(cell_scale, row_kernel)
        h.init()
        h.synchronize()
        t0 = time.perf_counter()
        h.run(1, steps)
        h.synchronize()
        return time.perf_counter() - t0
    finally:
        h.close()


def autotune(make, steps=None, cands=None, cache_path=None, key=None, log=None):
    """Pick the fastest (cell_scale, row_kernel) for the workload built by
    make() (a callable returning a DemSystem or Ensemble with its particles
    set, not yet initialised).  steps defaults to two rebuild intervals plus
    one.  With cache_path and key, a previous choice for the same key is
    returned without timing, and a new choice is stored."""
    cache = {}
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            c = cache[key]
            return (c["cell_scale"], c["row_kernel"]), c.get("timings", {})
    probe = make()
    nsn = getattr(probe, "params", None)
    nsn = (nsn[0] if isinstance(nsn, list) else nsn).neighbor_search_n if nsn is not None else 25
    probe.close()
    steps = steps or 2 * nsn + 1
    timings = {}
    best, best_t = None, float("inf")
    for c, r in (cands or candidates()):
        t = time_candidate(make, c, r, steps)
        timings[f"{c}/{r}"] = t
        if log:
            log(f"autotune: cell_scale {c} row_kernel {r}: {t * 1e3:.2f} ms / {steps} steps")
        if t < best_t:
            best, best_t = (c, r), t
    if cache_path and key:
        cache[key] = {"cell_scale": best[0], "row_kernel": best[1], "timings": timings}
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, timings


# --- skin (opt-in, changes results) -----------------------------------------
#
# FACTOR_RLM (the neighbour-list cut-off, FACTOR_RLM * (ri + rj)) and
# NEIGHBOR_SEARCH_N (the rebuild interval) are model parameters of the
# reference: they fix which pairs are listed and the order in which contact
# forces are summed, so changing them changes results in the last bits, and
# the reference's fixed schedule can miss a contact when particles close a
# gap larger than the skin between rebuilds.  The skin tuner below is for
# callers who want guaranteed contact detection and the fastest schedule
# that keeps it: for each candidate FACTOR_RLM it uses the longest interval
# the kinematic bound allows and times the pair.

def safe_interval(factor_rlm, r_min, v_max, dt, safety=2.0):
    """Longest rebuild interval that cannot miss a contact.  A pair is
    listed when its centre distance is at most FACTOR_RLM (ri + rj); two
    spheres close their gap by at most 2 v_max dt per step, so N steps are
    safe while 2 v_max dt N <= (FACTOR_RLM - 1) (ri + rj), whose smallest
    right-hand side is (FACTOR_RLM - 1) 2 r_min.  safety divides the bound
    (velocities can grow between the probe and the run)."""
    if v_max <= 0.0:
        return 1 << 30
    n = (factor_rlm - 1.0) * r_min / (safety * v_max * dt)
    return max(1, int(n))


def autotune_skin(make, params, r_min, v_max, factors=(1.1, 1.2, 1.35, 1.5), steps=100,
                  safety=2.0, log=None):
    """Pick (FACTOR_RLM, NEIGHBOR_SEARCH_N) for the fastest run that keeps
    every contact.  make(params) returns a handle (DemSystem or Ensemble)
    with its particles set, not initialised, for a copy of params with the
    candidate skin; params.dtsolid must be set.  Each candidate runs `steps`
    steps after init.  Returns ((factor_rlm, neighbor_search_n), timings)."""
    timings = {}
    best, best_t = None, float("inf")
    for f in factors:
        n = min(safe_interval(f, r_min, v_max, params.dtsolid, safety), 1 << 20)
        p = abi.Params.from_buffer_copy(params)
        p.factor_rlm = float(f)
        p.neighbor_search_n = int(n)
        h = make(p)
        try:
            h.init()
            h.synchronize()
            t0 = time.perf_counter()
            h.run(1, steps)
            h.synchronize()
            t = time.perf_counter() - t0
        finally:
            h.close()
        timings[f"{f}/{n}"] = t
        if log:
            log(f"skin: FACTOR_RLM {f} NEIGHBOR_SEARCH_N {n}: {t * 1e3:.2f} ms / {steps} steps")
        if t < best_t:
            best, best_t = (float(f), int(n)), t
    return best, timings

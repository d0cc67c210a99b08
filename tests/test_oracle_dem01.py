"""Pins the oracle's integrator (Euler and Adams-Bashforth), LSD wall
contact, parameter derivation (DES_ETAN_WALL, DTSOLID = TCOLL/50) and default
wall facets against the reference's own DEM01 regression data
(tests/dem/DEM01/AUTOTEST/POST_POS.dat, POST_VEL.dat: 38 runs of a sphere
bouncing on the south wall), printed there to 8 decimals."""
import json
import os

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from oracle_py import OracleSystem

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dem01_post.json")))
RUNS = GOLD["runs"]


def run_oracle(run):
    integ = abi.EULER if run["scheme"] == "EULER" else abi.ADAMS_BASHFORTH
    p, parts, _ = workloads.dem01(kn=run["kn"], en=run["en"], integrator=integ)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    dt = o.dtsolid()
    times = np.array([r[0] for r in run["rows"]])
    steps = np.rint(times / dt).astype(int)
    ys, vs = {0: 0.5}, {0: 0.0}
    for nn in range(1, steps.max() + 1):
        o.step(nn)
        if nn in set(steps.tolist()):
            pos, vel, _, _ = o.get_particles()
            ys[nn], vs[nn] = pos[1, 0], vel[1, 0]
    return dt, steps, ys, vs


@pytest.mark.parametrize("k", range(len(RUNS)),
                         ids=[f"{r['scheme']}-kn{r['kn']:g}-en{r['en']:g}" for r in RUNS])
def test_dem01_matches_reference_output(k):
    run = RUNS[k]
    dt, steps, ys, vs = run_oracle(run)
    rows = np.array(run["rows"])
    # the output times are whole steps of DTSOLID = TCOLL/50
    assert np.allclose(steps * dt, rows[:, 0], atol=6e-9)
    y = np.array([ys[s] for s in steps])
    v = np.array([vs[s] for s in steps])
    # the reference prints 8 decimals: agreement to the last printed digit
    assert np.max(np.abs(y - rows[:, 1])) <= 1.0e-8
    assert np.max(np.abs(v - rows[:, 2])) <= 1.0e-8

"""Pins the oracle's LINEAR_HAT deposition (calc_interp_weights.f:351-528,
comp_mean_fields1.f:7-116) and CALC_EPG_DES against the reference's CGDEM03
output (tests/cgdem/CGDEM03/POST_VOL.dat, t = 0): the 480 spheres of its
particle_input.dat, several within a radius of the walls, deposit a solids
volume of 0.25132741E-03 m^3 and leave a fluid volume of 0.99974867 m^3,
which sum to the 1 m^3 cube to 1e-15."""
import json
import os

import numpy as np

import _paths  # noqa: F401
from cgdem03_case import CELLS, cgdem03_case, volumes
from oracle_py import OracleSystem

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "cgdem03_post_vol.json")))["rows"]


def test_cgdem03_linear_hat_deposition_matches_reference_output():
    p, parts, gas = cgdem03_case()
    o = OracleSystem(p)
    o.set_particles(parts)
    o.set_gas(gas)
    o.init()
    ep, rop = o.get_gas()
    solids, fluid = volumes(ep, rop)
    t0, ref_solids, ref_fluid, ref_err = GOLD[0]
    assert t0 == 0.0
    # printed to 8 significant digits
    assert abs(solids - ref_solids) <= 5e-12
    assert abs(fluid - ref_fluid) <= 5e-9
    assert abs(solids + fluid - 1.0) <= 1e-15
    assert ref_err <= 1e-15
    # the closed form: the spheres' own volume
    assert abs(solids - np.sum(4.0 / 3.0 * np.pi * parts.radius ** 3)) <= 1e-18
    # spheres in the first and last cell layers: their hats reach the walls
    assert parts.pos.min() < 1.0 / CELLS and parts.pos.max() > 1.0 - 1.0 / CELLS
    o.close()

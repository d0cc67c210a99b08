"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on
the same seeded inputs.

Bar: DES cells, fluid cells, neighbour pair sets (and their row order) are
bit-exact.  FP64 trajectories are bit-exact wherever no libm pow enters
(LSD/Hertz contact, walls, integrators; the library is built with
-ffp-contract=off and the oracle keeps the reference's expression order).
With gas coupling the Gidaspow/Wen-Yu laws call pow (Re^0.687, eps^-2.65),
whose ocml and glibc results may differ in the last ulp: those cases are
held to the tolerances written in each test.
"""
import os
import subprocess

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem, pair_table
from oracle_py import OracleSystem

pytestmark = pytest.mark.gpu


def both(p, parts, gas=None, debug=0):
    g, o = DemSystem(p), OracleSystem(p)
    if debug:
        g.set_debug(debug)
    for s in (g, o):
        s.set_particles(parts)
        if gas is not None:
            s.set_gas(gas)
        s.init()
    return g, o


def settling(n=2000, seed=3, jostle=True, integrator=abi.EULER):
    p, parts, _ = workloads.settling_bed(n=n, seed=seed, box=(0.01, 0.035, 0.01))
    p.integrator = integrator
    if jostle:
        p, parts = workloads.jostled(p, parts, seed=seed + 1)
    return p, parts, None


def fluidised(n=4000, seed=2, coupled=True, precision=abi.FP64, jostle=True):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=0.02, ly=0.04, lz=0.02,
                                            cells=(10, 20, 10), dg=(0, 0, 0),
                                            precision=precision)
    if jostle:
        p, parts = workloads.jostled(p, parts, seed=seed + 1, v0=0.2)
    p.coupled = 1 if coupled else 0
    return p, parts, (gas if coupled else None)


def assert_same_pairs(g, o):
    gi, gn, gp = g.get_pairs()
    oi, on, op = o.get_pairs()
    assert np.array_equal(gi, oi), "NEIGHBOR_INDEX differs"
    assert np.array_equal(gn, on), "NEIGHBORS differs (set or row order)"
    return gp, op


def test_derived_parameters_bit_exact():
    for p, _, _ in (settling(), fluidised()):
        g, o = DemSystem(p), OracleSystem(p)
        dg, do = g.derived(), o.derived()
        for name, _t in abi.Derived._fields_:
            a, b = getattr(dg, name), getattr(do, name)
            if hasattr(a, "__len__"):
                assert np.array_equal(np.ctypeslib.as_array(a), np.ctypeslib.as_array(b)), name
            else:
                assert a == b, name
        g.close()


@pytest.mark.parametrize("case", ["settling", "fluidised"])
def test_init_cells_and_pairs_bit_exact(case):
    p, parts, gas = settling() if case == "settling" else fluidised()
    g, o = both(p, parts, gas)
    gd, gpijk = g.get_cells()
    od, opijk = o.get_cells()
    assert np.array_equal(gd, od), "DG_PIJK differs"
    if gas is not None:
        assert np.array_equal(gpijk, opijk), "PIJK differs"
    gp, op = assert_same_pairs(g, o)
    assert not gp.any() and not op.any()
    g.close()


@pytest.mark.parametrize("integrator", [abi.EULER, abi.ADAMS_BASHFORTH], ids=["euler", "ab2"])
def test_settling_trajectory_bit_exact(integrator):
    """LSD, walls, friction: 60 steps (rebuilds at 1, 25, 50), FP64."""
    p, parts, _ = settling(integrator=integrator)
    g, o = both(p, parts, debug=abi.DEBUG_STORE_FORCES)
    for nn in range(1, 61):
        g.step(nn)
        o.step(nn)
        if nn in (1, 2, 24, 25, 26, 60):
            gf, gt = g.get_forces()
            of, ot = o.get_forces()
            assert np.array_equal(gf, of), f"FC differs at step {nn}"
            assert np.array_equal(gt, ot), f"TOW differs at step {nn}"
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    gp, op = assert_same_pairs(g, o)
    assert np.array_equal(gp, op), "PFT history differs"
    assert o.status().n_contacts > 50
    g.close()


def test_hertz_polydisperse_walls_bit_exact():
    """Hertz, polydisperse 0.5-2 mm, walls, no gas: 60 steps bit-exact."""
    p, parts, _ = fluidised(coupled=False)
    g, o = both(p, parts)
    g.run(1, 60)
    o.run(1, 60)
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    gp, op = assert_same_pairs(g, o)
    assert np.array_equal(gp, op)
    g.close()


def test_reference_order_deposition_bit_exact():
    """With MFIXDEM_DEBUG_REFERENCE_DEPOSITION the node sums follow the
    reference's serial order (fluid cells ascending, particles ascending id
    within a cell): ROP_S and EP_G equal the oracle's bit for bit."""
    p, parts, gas = fluidised()
    g, o = both(p, parts, gas, debug=abi.DEBUG_REFERENCE_DEPOSITION)
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.array_equal(ge, oe) and np.array_equal(gr, orr)
    g.close()


def test_coupled_gas_deposition_and_drag():
    """GARG_2012 deposition (order-independent fixed-point node sums) +
    Gidaspow drag.  At set-up EP_G and ROP_S agree with the oracle's serial
    FP64 sums to 1e-14 relative; after 60 steps the trajectory differs only
    through pow ulps and deposition rounding: |dx| <= 1e-12 m, EP_G <= 1e-12."""
    p, parts, gas = fluidised()
    g, o = both(p, parts, gas)
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-14
    assert np.max(np.abs(gr - orr)) <= 1e-14 * np.max(np.abs(orr))
    g.run(1, 60)
    o.run(1, 60)
    gp, gv, gw, gs = g.get_particles()
    op, ov, ow, os_ = o.get_particles()
    assert np.array_equal(gs, os_)
    assert np.max(np.abs(gp - op)) <= 1e-12
    assert np.max(np.abs(gv - ov)) <= 1e-9 * max(1.0, np.max(np.abs(ov)))
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-12
    assert np.max(np.abs(gr - orr)) <= 1e-12 * np.max(np.abs(orr))
    gd, gpijk = g.get_cells()
    od, opijk = o.get_cells()
    assert np.array_equal(gpijk, opijk)
    g.close()


def test_unfused_entry_points_match_fused():
    """calc_force_dem / calc_drag_des / cfnewvalues / desgrid_pic /
    neighbour / particles_in_cell / comp_mean_fields called one by one give
    the same state as the fused step."""
    p, parts, gas = fluidised(n=2000)
    a, _ = both(p, parts, gas)
    b = DemSystem(p)
    b.set_particles(parts)
    b.set_gas(gas)
    b.init()
    for nn in range(1, 31):
        a.step(nn)
        b.calc_force_dem()
        b.calc_drag_des()
        b.cfnewvalues()
        if nn == 1 or nn % p.neighbor_search_n == 0:
            b.desgrid_pic()
            b.neighbour()
        b.particles_in_cell()
        b.comp_mean_fields()
    for x, y in zip(a.get_particles(), b.get_particles()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.get_gas()[0], b.get_gas()[0])
    a.close()
    b.close()


def test_graph_replay_matches_eager_steps():
    """mfixdem_run (hipGraph replay between rebuilds) == step-by-step."""
    p, parts, gas = fluidised(n=2000)
    a, _ = both(p, parts, gas)
    b = DemSystem(p)
    b.set_particles(parts)
    b.set_gas(gas)
    b.init()
    b.set_debug(abi.DEBUG_STORE_FORCES)  # forces the eager path
    a.run(1, 55)
    for nn in range(1, 56):
        b.step(nn)
    for x, y in zip(a.get_particles(), b.get_particles()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("k", [0, 5, 19, 24], ids=lambda k: f"run{k}")
def test_dem01_on_gpu_matches_reference_output(k):
    import json
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dem01_post.json")))
    run = gold["runs"][k]
    integ = abi.EULER if run["scheme"] == "EULER" else abi.ADAMS_BASHFORTH
    p, parts, _ = workloads.dem01(kn=run["kn"], en=run["en"], integrator=integ)
    g = DemSystem(p)
    g.set_particles(parts)
    g.init()
    dt = g.derived().dtsolid
    rows = np.array(run["rows"])
    steps = np.rint(rows[:, 0] / dt).astype(int)
    ys = {0: 0.5}
    for nn in range(1, steps.max() + 1):
        g.step(nn)
        if nn in set(steps.tolist()):
            ys[nn] = g.get_particles()[0][1, 0]
    y = np.array([ys[s] for s in steps])
    assert np.max(np.abs(y - rows[:, 1])) <= 1.0e-8
    g.close()


@pytest.mark.parametrize("k", [0, 3, 11], ids=lambda k: f"run{k}")
def test_dem02_on_gpu_matches_reference_output(k):
    """DEM02 bounce apexes (tests/dem/DEM02/AUTOTEST/POST_HEIGHT.dat) from
    the HIP library, stepped one DES_TIME_STEP at a time for the whole 10 s
    run: same apex count, heights to the last printed digit, and the same
    FP64 heights as the oracle."""
    import json
    from dem02_usr import bounce_heights, dem02_steps
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dem02_post.json")))
    run = gold["runs"][k]
    p, parts, _ = workloads.dem01(kn=run["kn"], en=run["en"], integrator=abi.ADAMS_BASHFORTH)
    g = DemSystem(p)
    g.set_particles(parts)
    g.init()
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    nsteps = dem02_steps(g.derived().dtsolid)
    assert nsteps == dem02_steps(o.dtsolid())
    h = bounce_heights(g, nsteps)
    rows = np.array(run["rows"])
    assert len(h) == len(rows)
    assert np.max(np.abs(h - rows[:, 2])) <= 1.0e-8
    assert np.array_equal(h, bounce_heights(o, nsteps))
    g.close()


def test_cgdem03_linear_hat_deposition_on_gpu():
    """LINEAR_HAT deposition of the reference's CGDEM03 spheres by the HIP
    library: EP_G and ROP_S equal the oracle's to 1e-14 (FP64 atomics add in
    another order), and the solids and fluid volumes are those of
    tests/cgdem/CGDEM03/POST_VOL.dat at t = 0 to the printed digits."""
    import json
    from cgdem03_case import cgdem03_case, volumes
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "cgdem03_post_vol.json")))["rows"][0]
    p, parts, gas = cgdem03_case()
    g, o = both(p, parts, gas)
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-14
    assert np.max(np.abs(gr - orr)) <= 1e-14 * np.max(np.abs(orr))
    solids, fluid = volumes(ge, gr)
    assert abs(solids - gold[1]) <= 5e-12 and abs(fluid - gold[2]) <= 5e-9
    assert abs(solids + fluid - 1.0) <= 1e-14
    g.close()


def test_fortran_driver_through_c_abi():
    """The Fortran ISO_C_BINDING shim drives DEM01 to the same trajectory."""
    exe = os.path.join(_paths.PKG, "lib", "dem01_driver")
    out = subprocess.run([exe, "10000", "0.9", "0", "300"], capture_output=True, text=True,
                         timeout=120, check=True).stdout.split("\n")
    fy = np.array([float(l.split()[2]) for l in out if l.strip()])
    p, parts, _ = workloads.dem01(kn=1e4, en=0.9)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    ys = []
    for nn in range(1, 301):
        o.step(nn)
        ys.append(o.get_particles()[0][1, 0])
    assert np.array_equal(fy, np.array(ys))


def test_udf_overrides_run_dem01():
    """The MFiX UDF overrides (bridge, CALC_FORCE_DEM, CFNEWVALUES,
    NEIGHBOUR, ...; built against stand-ins of MFiX's modules by
    tests/fortran/Makefile) run DEM01 through MFiX's variables and routine
    calls to the oracle's trajectory bit for bit."""
    exe = os.path.join(_paths.ROOT, "tests", "fortran", "_build", "udf_dem01")
    out = subprocess.run([exe, "10000", "0.9", "300"], capture_output=True, text=True,
                         timeout=120, check=True).stdout.split("\n")
    fy = np.array([float(l.split()[1]) for l in out if l.strip()])
    p, parts, _ = workloads.dem01(kn=1e4, en=0.9)
    o = OracleSystem(p)
    o.set_particles(parts)
    o.init()
    ys = []
    for nn in range(1, 301):
        o.step(nn)
        ys.append(o.get_particles()[0][1, 0])
    assert np.array_equal(fy, np.array(ys))


def test_mixed_precision_force_error():
    """FP32 contact and drag arithmetic: max relative force error against
    the FP64 path on the same state <= 1e-4 (error / max |F|)."""
    p64, parts, gas = fluidised(n=4000)
    pmx, _, _ = fluidised(n=4000, precision=abi.MIXED)
    a, _ = both(p64, parts, gas, debug=abi.DEBUG_STORE_FORCES)
    b = DemSystem(pmx)
    b.set_particles(parts)
    b.set_gas(gas)
    b.init()
    b.set_debug(abi.DEBUG_STORE_FORCES)
    a.step(1)
    b.step(1)
    fa, ta = a.get_forces()
    fb, tb = b.get_forces()
    scale = np.max(np.abs(fa))
    assert np.max(np.abs(fa - fb)) / scale <= 1e-4
    a.close()
    b.close()


def _golden(name):
    import json
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", name)))


@pytest.mark.parametrize("k", [0, 5], ids=["en1.0", "en0.5"])
def test_dem03_on_gpu_matches_reference_output(k):
    """Two stacked spheres of two phases between two walls
    (tests/dem/DEM03): GPU positions at the reference's output times agree
    with its printed MFIX columns to the last printed digit (1e-8)."""
    run = _golden("dem03_post.json")["runs"][k]
    p, parts, _ = workloads.dem03(en=run["en"])
    g = DemSystem(p)
    g.set_particles(parts)
    g.init()
    dt = g.derived().dtsolid
    rows = np.array(run["rows"])
    steps = np.rint(rows[:, 0] / dt).astype(int)
    want = set(steps.tolist())
    y = {0: (0.00045, 0.00135)}
    for nn in range(1, steps.max() + 1):
        g.step(nn)
        if nn in want:
            pos = g.get_particles()[0]
            y[nn] = (pos[1, 0], pos[1, 1])
    got = np.array([y[s] for s in steps])
    assert np.max(np.abs(got - rows[:, 1:3])) <= 1.0e-8
    g.close()


@pytest.mark.parametrize("mew", [0.2, 1.0])
def test_dem04_on_gpu_slip_to_roll_bit_exact(mew):
    """Sphere slipping on a rough wall (tests/dem/DEM04): the GPU reaches
    the end of slip at the same step as the oracle with identical velocity
    and spin, and that step matches the reference's printed time (1e-8)."""
    run = next(r for r in _golden("dem04_post.json")["runs"] if abs(r["mew"] - mew) < 1e-12)
    p, parts, _ = workloads.dem04(mew=mew)
    g, o = both(p, parts)
    dt = g.derived().dtsolid
    t = 0.0
    for nn in range(1, 200000):
        g.step(nn)
        o.step(nn)
        t += dt
        _, vel, om, _ = g.get_particles()
        slip = vel[0, 0] + om[2, 0] * 5.0e-4
        if slip < 0.0 or abs(1.0e-6 / abs(slip) - 1.0) <= 1.0e-4:
            break
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    assert abs(abs(mew * 9.80665 / 0.1) * t - run["time_nd"]) <= 1.0e-8
    g.close()


@pytest.mark.parametrize("model", ["LSD", "HERTZ"])
def test_dem05_on_gpu_bit_exact_and_wall_impacts(model):
    """Oblique impacts (tests/dem/DEM05), all 93 spheres for the full 2.5 ms
    (14,423 LSD / 338,195 Hertz steps, graph replay): every state and the
    pair history equal the oracle's bit for bit, and the 31 wall impacts
    give the reference's printed rebound angle, restitution and spin
    (4 decimals)."""
    gold = _golden("dem05_post.json")
    run = next(r for r in gold["runs"] if r["model"] == model)
    p, parts, _ = workloads.dem05(gold["particles"], model=model)
    g, o = both(p, parts)
    n = int(np.ceil(np.float32(0.0025 / g.derived().dtsolid)))
    g.run(1, n)
    o.run(1, n)
    gs, os_ = g.get_particles(), o.get_particles()
    for a, b in zip(gs, os_):
        assert np.array_equal(a, b)
    gp, op = assert_same_pairs(g, o)
    assert np.array_equal(gp, op)
    _, vel, om, _ = gs
    v0 = parts.vel
    vt0 = np.sqrt(v0[0, :31] ** 2 + v0[2, :31] ** 2)
    vt = np.sqrt(vel[0, :31] ** 2 + vel[2, :31] ** 2)
    alpha = np.arctan(vt / vel[1, :31]) * 180.0 / np.pi
    spin = np.sqrt(om[0, :31] ** 2 + om[2, :31] ** 2)
    assert np.max(np.abs(alpha - np.array(run["alpha"])[:, 2])) <= 5.1e-5
    assert np.max(np.abs(spin - np.array(run["omega"])[:, 2])) <= 5.1e-5
    assert np.max(np.abs(vt[1:] / vt0[1:] - np.array(run["coeff"])[:, 2])) <= 5.1e-5
    g.close()


FILTER_CASES = [
    (abi.INTERP_NONE, 0, None), (abi.INTERP_SQUARE_DPVM, 1, 1.6e-3),
    (abi.INTERP_GAUSS_DPVM, 1, 1.6e-3), (abi.INTERP_LINEAR_HAT, 1, None),
    (abi.INTERP_LINEAR_HAT, 0, None),
]


@pytest.mark.parametrize("scheme,interp_on,width", FILTER_CASES,
                         ids=["none", "square_dpvm", "gauss_dpvm", "linear_hat",
                              "linear_hat_cellvalues"])
def test_filter_scheme_drag_and_deposition(scheme, interp_on, width):
    """DES_INTERP_SCHEME filter path (CALC_INTERP_WEIGHTS, DRAG_GS_DES1,
    COMP_MEAN_FIELDS1) against the oracle.  ROP_S / EP_G at set-up agree to
    1e-14 (fixed-point vs serial sums); after 60 steps the trajectory differs
    only through pow ulps and deposition rounding: |dx| <= 1e-12 m,
    |dEP_G| <= 1e-12."""
    p, parts, gas = fluidised()
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    g, o = both(p, parts, gas)
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-14
    assert np.max(np.abs(gr - orr)) <= 1e-14 * np.max(np.abs(orr))
    g.run(1, 60)
    o.run(1, 60)
    gp, gv, gw, gs = g.get_particles()
    op, ov, ow, os_ = o.get_particles()
    assert np.array_equal(gs, os_)
    assert np.max(np.abs(gp - op)) <= 1e-12
    assert np.max(np.abs(gv - ov)) <= 1e-9 * max(1.0, np.max(np.abs(ov)))
    ge, gr = g.get_gas()
    oe, orr = o.get_gas()
    assert np.max(np.abs(ge - oe)) <= 1e-12
    assert np.max(np.abs(gr - orr)) <= 1e-12 * np.max(np.abs(orr))
    g.close()


def test_filter_scheme_mixed_precision_force_error():
    """FP32 drag law on the LINEAR_HAT path: relative force error <= 1e-4."""
    p64, parts, gas = fluidised(n=4000)
    pmx, _, _ = fluidised(n=4000, precision=abi.MIXED)
    for p in (p64, pmx):
        p.interp_scheme = abi.INTERP_LINEAR_HAT
        p.des_interp_on = 1
    a, _ = both(p64, parts, gas, debug=abi.DEBUG_STORE_FORCES)
    b = DemSystem(pmx)
    b.set_particles(parts)
    b.set_gas(gas)
    b.init()
    b.set_debug(abi.DEBUG_STORE_FORCES)
    a.step(1)
    b.step(1)
    fa, _ = a.get_forces()
    fb, _ = b.get_forces()
    assert np.max(np.abs(fa - fb)) / np.max(np.abs(fa)) <= 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("integrator", [abi.EULER, abi.ADAMS_BASHFORTH], ids=["euler", "ab2"])
def test_nsquare_search_bit_exact(integrator):
    """DES_NEIGHBOR_SEARCH = 1 (NSQUARE): pair lists (ascending partner id),
    FC/TOW and the 60-step trajectory equal the oracle's bit for bit."""
    p, parts, _ = settling(integrator=integrator)
    p.des_neighbor_search = abi.NSEARCH_NSQUARE
    g, o = both(p, parts, debug=abi.DEBUG_STORE_FORCES)
    gp, op = assert_same_pairs(g, o)
    for nn in range(1, 61):
        g.step(nn)
        o.step(nn)
        if nn in (1, 25, 26, 60):
            gf, gt = g.get_forces()
            of, ot = o.get_forces()
            assert np.array_equal(gf, of), f"FC differs at step {nn}"
            assert np.array_equal(gt, ot), f"TOW differs at step {nn}"
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    gp, op = assert_same_pairs(g, o)
    assert np.array_equal(gp, op)
    g.close()


def test_pressure_gradient_and_coupled_run_bit_exact():
    """mfixdem_set_gas_pressure (CALC_PG_GRAD) gives the oracle's P_FORCE bit
    for bit; a coupled run driven by it matches as the GARG test does."""
    p, parts, gas = fluidised()
    g, o = DemSystem(p), OracleSystem(p)
    nc = g.ncell
    rng = np.random.default_rng(11)
    p_g = 1.0e5 - 2500.0 * rng.uniform(0.0, 0.04, nc) + rng.normal(0.0, 5.0, nc)
    for s in (g, o):
        s.set_particles(parts)
        s.set_gas(gas)
        s.set_gas_pressure(p_g)
        s.init()
    assert np.array_equal(g.get_p_force(), o.get_p_force())
    g.run(1, 30)
    o.run(1, 30)
    gp, op = g.get_particles()[0], o.get_particles()[0]
    assert np.max(np.abs(gp - op)) <= 1e-12
    g.close()


@pytest.mark.parametrize("name", ["NONE", "GARG_2012", "SQUARE_DPVM"])
def test_dem06_on_gpu_terminal_velocity(name):
    """Single sphere settling against a frozen upflow (tests/dem/DEM06, the
    three one-way coupled runs): the GPU follows the oracle (pow in the drag
    law: 1e-12 relative on y and v at every printed time) and the
    reference's printed MFIX columns to the bounds the oracle meets (3e-5
    m/s, 1e-5 m; the reference's drag ignores the cell voidage)."""
    from test_oracle_dem06 import SCHEMES, golden, trajectory
    scheme, width = SCHEMES[name]
    p, parts, gas, p_g, sub = workloads.dem06(scheme, width)
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        s.set_gas(gas)
        s.set_gas_pressure(p_g)
        s.init()
    tg = trajectory(g, sub, 350)
    to = trajectory(o, sub, 350)
    g.close()
    o.close()
    assert np.max(np.abs(tg[:, 1:] - to[:, 1:]) / np.abs(to[:, 1:])) <= 1e-12
    rows = np.array(next(r for r in golden() if r["scheme"] == name)["rows"])
    assert np.max(np.abs(tg[:, 2] - rows[1:, 2])) <= 3.0e-5
    assert np.max(np.abs(tg[:, 1] - rows[1:, 1])) <= 1.0e-5


@pytest.mark.parametrize("rows", [abi.ROWS_THREAD, abi.ROWS_WAVE192], ids=["thread", "wave192"])
@pytest.mark.parametrize("skin", [1.2, 3.0], ids=["skin1.2", "skin3.0-long-rows"])
def test_row_kernels_bit_exact(rows, skin):
    """Every neighbour-row kernel gives the oracle's pair set, row order and
    trajectory: the half-stencil thread kernel (atomic slots + row sort) and
    the wavefront kernel; FACTOR_RLM = 3 makes rows longer than the initial
    row capacity (48), so the capacity-growth path runs too."""
    p, parts, _ = settling()
    p.factor_rlm = skin
    g, o = DemSystem(p), OracleSystem(p)
    g.set_tuning(1.0, rows)
    for s in (g, o):
        s.set_particles(parts)
        s.init()
    gp, op = assert_same_pairs(g, o)
    if skin > 2.0:
        idx = o.get_pairs()[0]
        assert np.max(np.diff(np.concatenate([[1], idx]))) > 48
    for s in (g, o):
        s.run(1, 30)
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    assert_same_pairs(g, o)
    g.close()
    o.close()


def test_kernel_timing_mode_matches_plain_run():
    """MFIXDEM_DEBUG_TIME_KERNELS (eager launches bracketed by pooled HIP
    events) steps exactly like a plain run with graph replay, coupled gas
    included, and reports every kernel group."""
    import ctypes as C
    p, parts, gas = fluidised(n=2000)
    a, _ = both(p, parts, gas)
    c = DemSystem(p)
    c.set_particles(parts)
    c.set_gas(gas)
    c.init()
    c.set_debug(abi.DEBUG_TIME_KERNELS)
    a.run(1, 80)
    c.run(1, 80)
    for x, y in zip(a.get_particles(), c.get_particles()):
        assert np.array_equal(x, y)
    ea, _ = a.get_gas()
    ec, _ = c.get_gas()
    assert np.array_equal(ea, ec)
    ms = (C.c_double * abi.TIME_GROUPS)()
    calls = (C.c_int64 * abi.TIME_GROUPS)()
    assert c.lib.mfixdem_get_kernel_times(c.h, ms, calls, abi.TIME_GROUPS) == 0
    assert calls[abi.TIME_STEP] == 80 and calls[abi.TIME_GAS] == 80
    assert calls[abi.TIME_REBUILD] == 4   # NN = 1, 25, 50, 75
    assert min(ms[abi.TIME_STEP], ms[abi.TIME_GAS], ms[abi.TIME_REBUILD]) > 0.0
    a.close()
    c.close()


def test_gas_side_drag_matches_oracle():
    """DRAG_GS_GAS0 (mfixdem_drag_gs_gas): F_GDS, DRAG_AM and DRAG_BM against
    the oracle.  At set-up with the reference-order deposition EP_G is
    bit-exact, so the node sums (gathered in the reference's cell and
    particle order) differ only where Wen-Yu's pow differs by an ulp:
    1e-13 of the largest value.  After 30 coupled steps the particles
    differ by <= 1e-12 m, and so do the drag fields, to 1e-9."""
    p, parts, gas = fluidised(n=3000)
    g, o = both(p, parts, gas, debug=abi.DEBUG_REFERENCE_DEPOSITION)
    f = np.zeros(g.ncell)
    rc = g.lib.mfixdem_get_gas_drag(g.h, abi.dptr(f), None, None)
    assert rc == abi.ERRORS_BY_NAME["NOT_READY"]
    assert np.array_equal(g.get_gas()[0], o.get_gas()[0])
    for tol, steps in ((1e-13, 0), (1e-9, 30)):
        if steps:
            g.run(1, steps)
            o.run(1, steps)
        fg, ag, bg = g.drag_gs_gas()
        fo, ao, bo = o.drag_gs_gas()
        assert fo.max() > 0.0 and ao.max() > 0.0
        assert np.max(np.abs(fg - fo)) <= tol * fo.max()
        assert np.max(np.abs(ag - ao)) <= tol * ao.max()
        assert np.max(np.abs(bg - bo)) <= tol * np.max(np.abs(bo))
        assert np.array_equal(fg == 0.0, fo == 0.0)
    g.close()
    o.close()


@pytest.mark.parametrize("scheme,interp_on,width", FILTER_CASES,
                         ids=["none", "square_dpvm", "gauss_dpvm", "linear_hat",
                              "linear_hat_cellvalues"])
def test_filter_gas_side_drag_matches_oracle(scheme, interp_on, width):
    """DRAG_GS_GAS1 (filter schemes): F_GDS and DRAG_BM against the oracle.
    The reference and the kernel both add with atomics in no fixed order,
    and EP_G at set-up agrees to 1e-14 (fixed point against serial sums), so
    the cell sums agree to 1e-12 of the largest value; after 20 coupled
    steps (velocities within 1e-9) to 1e-9.  DRAG_AM is 0."""
    p, parts, gas = fluidised()
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    g, o = both(p, parts, gas)
    for tol, steps in ((1e-12, 0), (1e-9, 20)):
        if steps:
            g.run(1, steps)
            o.run(1, steps)
        fg, ag, bg = g.drag_gs_gas()
        fo, ao, bo = o.drag_gs_gas()
        assert fo.max() > 0.0
        assert not ag.any() and not ao.any()
        assert np.max(np.abs(fg - fo)) <= tol * fo.max()
        assert np.max(np.abs(bg - bo)) <= tol * np.max(np.abs(bo))
    g.close()
    o.close()


def _rows(idx, nb, pft):
    """{(row id, partner id): PFT} of a half list in reference layout."""
    out, start = {}, 0
    for q in range(len(idx)):
        for e in range(start, idx[q] - 1):
            out[(q, int(nb[e]) - 1)] = pft[e]
        start = idx[q] - 1
    return out


def test_restart_pair_history():
    """mfixdem_set_pairs: a restart from the NEIGHBOR_INDEX, NEIGHBORS and
    PFT that mfixdem_get_pairs writes (what a DES restart file holds), given
    before init, carries every pair's tangential history into the new
    neighbour list bit for bit; pairs the new list gains start at zero.
    Without it the restarted run starts with no history."""
    p, parts, _ = settling()
    a = DemSystem(p)
    a.set_particles(parts)
    a.init()
    a.run(1, 60)
    pos, vel, om, state = a.get_particles()
    assert np.all(state == 1)
    ai, an, ap = a.get_pairs()
    old = _rows(ai, an, ap)
    assert sum(bool(np.any(v != 0.0)) for v in old.values()) > 0
    rp = workloads.make_particles(pos, parts.radius, parts.ro_sol, vel=vel, omega=om,
                                  phase=parts.phase)
    b = DemSystem(p)
    b.set_particles(rp)
    b.set_pairs(ai, an, ap)
    b.init()
    new = _rows(*b.get_pairs())
    carried = 0
    for key, h in new.items():
        if key in old:
            assert np.array_equal(h, old[key]), key
            carried += bool(np.any(h != 0.0))
        else:
            assert not np.any(h), key
    assert carried > 0
    with pytest.raises(abi.MfixDemError):  # the history is restarted before init only
        b.set_pairs(ai, an, ap)
    c = DemSystem(p)
    c.set_particles(rp)
    c.init()
    assert not c.get_pairs()[2].any()
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("scheme,interp_on,width", [
    (abi.INTERP_NONE, 0, None), (abi.INTERP_SQUARE_DPVM, 1, 1.6e-3),
    (abi.INTERP_LINEAR_HAT, 1, None)], ids=["none", "square_dpvm", "linear_hat"])
def test_explicit_coupling_matches_oracle(scheme, interp_on, width):
    """DES_EXPLICITLY_COUPLED over three fluid steps: CALC_DES_2FLUID
    (PARTICLES_IN_CELL, weights, mean fields, DRAG_GS_GAS1 keeping F_GP),
    a new gas field, DES_TIME_INIT (DRAG_FC) and 15 DEM steps that add
    DRAG_FC and do not rebin.  Against the oracle: F_GDS within 1e-12 of its
    largest value (atomics), positions within 1e-12 m and velocities within
    1e-9 (pow ulps and deposition rounding), PIJK bit for bit."""
    p, parts, gas = fluidised()
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    p.des_explicitly_coupled = 1
    g, o = both(p, parts, gas)
    nn = 1
    for k, scale in enumerate((1.0, 1.15, 0.9)):
        fo = None
        for s in (g, o):
            s.particles_in_cell()
            s.calc_interp_weights()
            s.comp_mean_fields()
            f, _, _ = s.drag_gs_gas()
            if fo is None:
                fg = f
            fo = f
        assert fo.max() > 0.0
        assert np.max(np.abs(fg - fo)) <= 1e-12 * fo.max()
        gas_k = workloads.Gas(scale * gas.u_g, scale * gas.v_g, scale * gas.w_g, gas.p_force)
        for s in (g, o):
            s.set_gas(gas_k)
            s.des_time_init()
            s.run(nn, 15)
        nn += 15
        assert np.array_equal(g.get_cells()[1], o.get_cells()[1]), f"PIJK, fluid step {k}"
    gp, gv, _, gs = g.get_particles()
    op, ov, _, os_ = o.get_particles()
    assert np.array_equal(gs, os_)
    assert np.max(np.abs(gp - op)) <= 1e-12
    assert np.max(np.abs(gv - ov)) <= 1e-9 * max(1.0, np.max(np.abs(ov)))
    g.close()
    o.close()


def _with_holes(n, frac, seed):
    st = np.ones(n, dtype=np.int8)
    st[np.random.default_rng(seed).random(n) < frac] = 0
    return st


def test_particle_state_holes_match_oracle():
    """PARTICLE_STATE holes at set-up (mfixdem_set_particle_state): 15 % of
    the slots nonexistent; cells, pair lists, PFT and the trajectory over
    60 steps (two rebuilds) bit for bit against the oracle, holes kept."""
    p, parts, _ = settling()
    st = _with_holes(parts.n, 0.15, 9)
    g, o = DemSystem(p), OracleSystem(p)
    for s in (g, o):
        s.set_particles(parts)
        s.set_particle_state(st)
        s.init()
    gp_, op_ = assert_same_pairs(g, o)
    assert np.array_equal(gp_, op_)
    for s in (g, o):
        s.run(1, 60)
    a, b = g.get_particles(), o.get_particles()
    assert np.array_equal(a[3], st) and np.array_equal(b[3], st)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    gp_, op_ = assert_same_pairs(g, o)
    assert np.array_equal(gp_, op_)
    with pytest.raises(abi.MfixDemError):   # after init
        g.set_particle_state(st)
    bad = st.copy()
    bad[0] = 2                              # entering particles are not supported
    h = DemSystem(p)
    h.set_particles(parts)
    with pytest.raises(abi.MfixDemError):
        h.set_particle_state(bad)
    for s in (g, o, h):
        s.close()


def test_des_res_restart_continues_the_run_bit_exact(tmp_path):
    """A run with holes, saved at step 60 to <RUN_NAME>_DES.RES (positions,
    velocities, omega, radius, density, phase, global ids, states and the
    pair histories in the WRITE_RES0_DES layout), read back into a fresh
    handle (particles at their global ids, holes, mfixdem_set_pairs) and
    continued from NN = 61, ends bit for bit where the uninterrupted run
    ends once that run rebuilds its list at the same moment (the restart
    builds its list at set-up).  Spheres start 2 mm inside the walls and
    move < 0.2 mm, so no wall history (which the restart file does not
    hold) is lost; Euler, so no Adams-Bashforth history is either."""
    from mfixdem import restart
    p, parts, _ = workloads.settling_bed(n=600, seed=4, box=(0.014, 0.024, 0.014))
    r = parts.radius
    pos = workloads.rsa_positions(4, r, [0.002, 0.002, 0.002], [0.012, 0.022, 0.012])
    parts = workloads.make_particles(pos, r, parts.ro_sol)
    p, parts = workloads.jostled(p, parts, seed=5, shrink=0.93, v0=0.1, w0=100.0)
    st = _with_holes(parts.n, 0.05, 6)
    a = DemSystem(p)
    a.set_particles(parts)
    a.set_particle_state(st)
    a.init()
    a.run(1, 60)
    res = restart.from_system(a, parts.radius, parts.ro_sol, parts.phase)
    assert res.n == int(st.sum()) and np.any(res.pft != 0.0)
    path = str(tmp_path / "RUN_DES.RES")
    restart.write_des_res(path, res)
    back = restart.read_des_res(path)
    # rebuild the slots: particle with global id g in slot g-1, holes elsewhere
    n = p.max_pip
    full = lambda v, k: np.zeros((k, n)) if k else np.zeros(n)  # noqa: E731
    P, V, W, R, RHO = full(0, 3), full(0, 3), full(0, 3), np.full(n, 1e-4), np.full(n, 1.0)
    PH, ST = np.ones(n, dtype=np.int32), np.zeros(n, dtype=np.int8)
    slot = back.gid - 1
    P[:, slot], V[:, slot], W[:, slot] = back.pos, back.vel, back.omega
    R[slot], RHO[slot], PH[slot], ST[slot] = back.radius, back.ro_sol, back.phase, back.state
    rp = workloads.make_particles(P, R, RHO, vel=V, omega=W, phase=PH)
    ends = 1 + np.cumsum(np.bincount(back.pair_owner - 1, minlength=n))
    b = DemSystem(p)
    b.set_particles(rp)
    b.set_particle_state(ST)
    b.set_pairs(ends.astype(np.int32), back.pair_partner, back.pft)
    b.init()
    a.desgrid_pic()
    a.neighbour()
    for s in (a, b):
        s.run(61, 40)
    for x, y in zip(a.get_particles(), b.get_particles()):
        assert np.array_equal(x, y)
    ai, an, ap = a.get_pairs()
    bi, bn, bp = b.get_pairs()
    assert np.array_equal(ai, bi) and np.array_equal(an, bn) and np.array_equal(ap, bp)
    a.close()
    b.close()

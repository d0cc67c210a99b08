"""particle_input.dat reader and writer (mfixdem.particle_input) against the
reference's own input files (tests/golden/particle_input/, copied from its
DEM03, DEM05, DEM07, DEM08 and CGDEM03 tests).  DEM05's 93 particles were also
extracted independently into dem05_post.json, whose DEM05 oracle runs match
the reference's POST files, so the Version 1.0 columns are pinned by that
regression."""
import json
import os

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import workloads
from mfixdem.particle_input import read_particle_input, write_particle_input

HERE = os.path.dirname(os.path.abspath(__file__))
INP = os.path.join(HERE, "golden", "particle_input")


def test_version1_matches_the_dem05_fixture():
    parts = read_particle_input(os.path.join(INP, "DEM05.dat"), particles=93)
    ref = np.array(json.load(open(os.path.join(HERE, "golden", "dem05_post.json")))["particles"])
    assert parts.n == 93
    assert np.array_equal(parts.pos, ref[:, 0:3].T)
    assert np.array_equal(parts.radius, ref[:, 3])
    assert np.array_equal(parts.ro_sol, ref[:, 4])
    assert np.array_equal(parts.vel, ref[:, 5:8].T)
    assert np.all(parts.phase == 1) and not parts.omega.any()


def test_version1_record_count():
    parts = read_particle_input(os.path.join(INP, "DEM03.dat"))
    assert parts.n == 2
    assert np.array_equal(parts.pos[:, 0], [0.5, 0.00045, 0.5])
    assert parts.radius[0] == 0.0005 and np.array_equal(parts.ro_sol, [20000.0, 10000.0])
    assert read_particle_input(os.path.join(INP, "DEM03.dat"), particles=1).n == 1
    with pytest.raises(ValueError, match="end of file"):
        read_particle_input(os.path.join(INP, "DEM03.dat"), particles=3)


def test_version2_defaults_and_columns():
    one = read_particle_input(os.path.join(INP, "DEM07.dat"))
    assert one.n == 1
    assert np.array_equal(one.pos[:, 0], [0.0, 0.005, 0.0])
    assert one.phase[0] == 1 and one.radius[0] == 0.005 and one.ro_sol[0] == 2500.0
    assert np.array_equal(one.vel[:, 0], [1.0, 0.0, 0.0])
    bed = read_particle_input(os.path.join(INP, "DEM08.dat"))
    assert bed.n == 1154
    assert np.array_equal(bed.pos[:, 0], [0.48914431e-02, 0.15488297e+00, 0.50084351e-02])
    assert np.all(bed.radius == 0.5 * 0.01) and np.all(bed.ro_sol == 2500.0)
    assert not bed.vel.any() and np.all(bed.phase == 1)


def test_version2_tabs_and_user_scalars():
    """CGDEM03's file: tab-separated columns, velocities read, three user
    scalars after them (not kept), diameter and density from the defaults."""
    cg = read_particle_input(os.path.join(INP, "CGDEM03.dat"))
    assert cg.n == 480
    assert np.array_equal(cg.pos[:, 0], [9.92e-01, 1.31e-02, 1.49e-01])
    assert np.array_equal(cg.pos[:, 2], [9.33e-03, 2.87e-02, 1.47e-01])
    assert np.all(cg.radius == 0.005) and np.all(cg.ro_sol == 2700.0)
    assert not cg.vel.any() and np.all(cg.phase == 1)


def test_version2_read_columns(tmp_path):
    """Variables marked T take columns in the reference's order (coordinates,
    phase, diameter, density, velocity); Fortran D exponents and commas."""
    src = open(os.path.join(INP, "DEM07.dat")).read().splitlines()
    src[19] = "Particles:    2"
    src[24] = "Phase_ID:        T          1"
    src[25] = "Diameter:        T          0.010"
    src[27] = "Velocity:        T          1.0 0.0 0.0"
    data = ["0.1 0.2 0.3 2 1.0D-3 0.5 -1.5 0.25", "0.4, 0.5, 0.6, 1, 2.0d-3, 0.0, 0.0, 3.0"]
    p = tmp_path / "particle_input.dat"
    p.write_text("\n".join(src[:34] + data) + "\n")
    parts = read_particle_input(str(p))
    assert parts.n == 2
    assert np.array_equal(parts.phase, [2, 1])
    assert np.array_equal(parts.radius, [0.5e-3, 1.0e-3])
    assert np.array_equal(parts.ro_sol, [2500.0, 2500.0])
    assert np.array_equal(parts.vel[:, 0], [0.5, -1.5, 0.25])
    assert np.array_equal(parts.pos[:, 1], [0.4, 0.5, 0.6])


def test_unsupported_versions_and_missing_coordinates(tmp_path):
    p = tmp_path / "v3.dat"
    p.write_text("Version 3.0\n")
    with pytest.raises(ValueError, match="Version 3.0"):
        read_particle_input(str(p))
    src = open(os.path.join(INP, "DEM07.dat")).read().splitlines()
    src[23] = "Coordinates:     F          Must always be T"
    p = tmp_path / "nocoord.dat"
    p.write_text("\n".join(src) + "\n")
    with pytest.raises(ValueError, match="Coordinates"):
        read_particle_input(str(p))


def test_write_read_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(7)
    n = 50
    parts = workloads.make_particles(rng.uniform(-1, 1, (3, n)), rng.uniform(1e-4, 2e-3, n),
                                     rng.uniform(900.0, 8000.0, n),
                                     vel=rng.normal(0.0, 2.0, (3, n)))
    p = tmp_path / "particle_input.dat"
    write_particle_input(str(p), parts)
    back = read_particle_input(str(p))
    for a in ("pos", "vel", "radius", "ro_sol"):
        assert np.array_equal(getattr(back, a), getattr(parts, a)), a

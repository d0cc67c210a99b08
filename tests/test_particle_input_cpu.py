"""Particle input files (mfixdem.particle_input) against the reference's own
input files (tests/golden/particle_input/, copied from its DEM03, DEM05,
DEM07, DEM08 and CGDEM03 tests and from the particle_input.csv files of its
screwfeeder_3d and mixer_3d tutorials).  DEM05's 93 particles were also
extracted independently into dem05_post.json, whose DEM05 oracle runs match
the reference's POST files, so the Version 1.0 columns are pinned by that
regression.  Version 3.0 and the PART_IN filters are checked on small
files written here; the DES.RES writer and reader (mfixdem.restart) on a
round trip and on the record positions of WRITE_RES0_DES's layout (no
restart file written by the reference exists to compare against)."""
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
    p = tmp_path / "v4.dat"
    p.write_text("Version 4.0\n")
    with pytest.raises(ValueError, match="Version 4.0"):
        read_particle_input(str(p))
    p = tmp_path / "v3.dat"
    p.write_text("Version 3.0\n")
    with pytest.raises(ValueError, match="Version 3.0 header"):
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


# ---------------------------------------------------------------- Version 4.0
from mfixdem.particle_input import PartIn, read_particle_input_csv  # noqa: E402


def test_version4_csv_screwfeeder_two_phases():
    """tutorials/dem/screwfeeder_3d: every column given, two phases,
    PARTICLES = 3964 = the file's rows."""
    opt = PartIn(particles=3964, des_mmax=2, d_p0=(0.005, 0.005), ro_s0=(1000.0, 1000.0),
                 domain=(-0.205, 0.055, 0.0, 0.3, -0.055, 0.055))
    parts, usr = read_particle_input_csv(os.path.join(INP, "screwfeeder_3d.csv"), opt)
    assert parts.n == 3964 and usr is None
    assert np.array_equal(parts.pos[:, 0], [0.002532, 0.1069, 0.002444])
    assert np.array_equal(parts.pos[:, 1], [0.002534, 0.1069, -0.002443])
    assert set(np.unique(parts.phase)) == {1, 2}
    assert np.all(parts.radius == 0.0025) and np.all(parts.ro_sol == 1000.0)
    assert not parts.vel.any()


def test_version4_csv_mixer_first_particles_and_user_scalar():
    """tutorials/dem/mixer_3d: PARTICLES = 2000 of 2873 rows are read; the
    usr_var1 column is DES_USR_VAR_SIZE = 1's scalar; without it declared
    the column count does not match (the reference stops)."""
    opt = PartIn(particles=2000, d_p0=(0.01,), ro_s0=(1000.0,), des_usr_var_size=1,
                 domain=(-0.1, 0.1, -0.1, 0.1, -0.1, 0.1))
    parts, usr = read_particle_input_csv(os.path.join(INP, "mixer_3d.csv"), opt)
    assert parts.n == 2000 and usr.shape == (1, 2000)
    assert np.array_equal(parts.pos[:, 0], [0.00928839, -0.0945781, -0.09293118])
    assert set(np.unique(usr)) <= {1.0, 2.0}
    with pytest.raises(ValueError, match="number of columns"):
        read_particle_input_csv(os.path.join(INP, "mixer_3d.csv"),
                                PartIn(particles=2000, d_p0=(0.01,), ro_s0=(1000.0,)))
    with pytest.raises(ValueError, match="only 2873"):
        read_particle_input_csv(os.path.join(INP, "mixer_3d.csv"),
                                PartIn(particles=3000, des_usr_var_size=1))


def test_version4_defaults_domain_and_filters(tmp_path):
    """Missing columns take phase 1, D_P0 / RO_S0 of the phase and zero
    velocity; particles on or outside the domain bounds are dropped; a
    PART_IN filter keeps min < v < max (exclude=False) or drops
    min <= v <= max (exclude=True); PART_IN_PHASE drops a phase."""
    p = tmp_path / "particle_input.csv"
    p.write_text("X, Y ,Z,Phase_ID\n0.1,0.2,0.3,1\n0.0,0.5,0.5,2\n0.4,0.5,0.6,2\n"
                 "0.7,0.8,0.9,1\n")
    opt = PartIn(des_mmax=2, d_p0=(1e-3, 2e-3), ro_s0=(2500.0, 3000.0),
                 domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0))
    parts, _ = read_particle_input_csv(str(p), opt)
    assert parts.n == 3   # x = 0.0 lies on X_MIN
    assert np.array_equal(parts.phase, [1, 2, 1])
    assert np.array_equal(parts.radius, [0.5e-3, 1e-3, 0.5e-3])
    assert np.array_equal(parts.ro_sol, [2500.0, 3000.0, 2500.0])
    assert not parts.vel.any()
    f = dataclasses.replace(opt, filters={"y": (0.4, 0.85, False)})
    assert np.array_equal(read_particle_input_csv(str(p), f)[0].pos[1], [0.5, 0.8])
    f = dataclasses.replace(opt, filters={"y": (0.4, 0.85, True)})
    assert np.array_equal(read_particle_input_csv(str(p), f)[0].pos[1], [0.2])
    f = dataclasses.replace(opt, part_in_phase=(True, False))
    assert np.array_equal(read_particle_input_csv(str(p), f)[0].phase, [1, 1])


def test_version3_layout_and_filters(tmp_path):
    """Version 3.0: 19 text lines, Dimension, Particles, 3 lines, the eight
    variable lines (name token, T/F, defaults), 3 lines, data."""
    head = ["Version 3.0"] + [f"! comment {k}" for k in range(18)]
    head += ["Dimension: 3", "Particles: 3", "!", "!", "!",
             "Coordinates:  T", "Phase_ID:  F  1", "Diameter:  T", "Density:  F  2500.0",
             "Velocity:  T", "Temperature:  F  300.0", "Species:  F  1.0", "User_scalar:  T",
             "!", "!", "!"]
    data = ["0.1 0.2 0.3 1.0D-3 0.0 -1.0 0.0 5.0", "0.4 0.5 0.6 2.0D-3 1.0 0.0 0.0 6.0",
            "0.7 0.8 0.9 3.0D-3 0.0 0.0 2.0 7.0"]
    p = tmp_path / "particle_input.dat"
    p.write_text("\n".join(head + data) + "\n")
    opt = PartIn(des_usr_var_size=1, filters={"diameter": (2.5e-3, 1.0, True)})
    parts, usr = read_particle_input(str(p), opt=opt)
    assert parts.n == 2
    assert np.array_equal(parts.radius, [0.5e-3, 1.0e-3])
    assert np.all(parts.ro_sol == 2500.0) and np.all(parts.phase == 1)
    assert np.array_equal(parts.vel[:, 1], [1.0, 0.0, 0.0])
    assert np.array_equal(usr, [[5.0, 6.0]])


# ---------------------------------------------------------------- DES.RES
import dataclasses  # noqa: E402

from mfixdem import restart  # noqa: E402


def _res(n=300, m=700, seed=3):
    rng = np.random.default_rng(seed)
    return restart.DesRes(
        1.2, 1.25e-5, rng.random((3, n)), np.arange(1, n + 1, dtype=np.int32),
        np.ones(n, dtype=np.int32), rng.normal(size=(3, n)), rng.normal(size=(3, n)),
        rng.uniform(1e-4, 2e-4, n), np.full(n, 2500.0), rng.integers(1, 3, n).astype(np.int32),
        np.zeros((0, n)), np.zeros(n), rng.integers(1, n + 1, m).astype(np.int32),
        rng.integers(1, n + 1, m).astype(np.int32), rng.normal(size=(m, 3)))


def test_des_res_round_trip_and_record_layout(tmp_path):
    r = _res()
    path = str(tmp_path / "RUN_DES.RES")
    restart.write_des_res(path, r)
    raw = open(path, "rb").read()
    # 7 header records; 300 doubles = 5 records, 300 ints = 3 records; 700
    # collisions: 6 records of ints, 11 of doubles; one record per scalar
    per = 3 * 5 + 3 + 3 + 3 * 5 + 3 * 5 + 5 + 5 + 3 + 1 + 5
    assert len(raw) == 512 * (7 + per + 2 * 6 + 3 * 11 + 1)
    assert np.frombuffer(raw[:8], "<f8")[0] == 1.2
    assert np.frombuffer(raw[512:516], "<i4")[0] == 300
    assert np.frombuffer(raw[3 * 512:3 * 512 + 4], "<i4")[0] == 700
    assert np.frombuffer(raw[6 * 512:6 * 512 + 8], "<f8")[0] == 1.25e-5
    assert np.array_equal(np.frombuffer(raw[7 * 512:8 * 512], "<f8"), r.pos[0, :64])
    back = restart.read_des_res(path)
    for f in dataclasses.fields(r):
        a, b = getattr(r, f.name), getattr(back, f.name)
        assert np.array_equal(np.asarray(a), np.asarray(b)), f.name


def test_des_res_empty_arrays_take_one_record(tmp_path):
    """OUT_BIN_512 writes one record for N <= 64 values, N = 0 included."""
    r = _res(n=10, m=0)
    path = str(tmp_path / "RUN_DES.RES")
    restart.write_des_res(path, r)
    back = restart.read_des_res(path)
    assert back.pair_owner.size == 0 and back.pft.shape == (0, 3)
    assert len(open(path, "rb").read()) == 512 * (7 + 3 + 2 + 3 + 3 + 2 + 1 + 1 + 1 + 2 + 3 + 1)

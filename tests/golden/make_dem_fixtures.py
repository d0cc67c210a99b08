"""Extract the reference's DEM regression data into compact JSON fixtures.

Reads the AUTOTEST files the reference ships for its DEM tests (data, not
code) and keeps, per run, the solver columns the tests compare:
  DEM01  tests/dem/DEM01/AUTOTEST/POST_POS.dat, POST_VEL.dat
         (scheme, KN, EN) -> rows of (time, y_mfix, v_mfix)
  DEM02  tests/dem/DEM02/AUTOTEST/POST_HEIGHT.dat
         (KN, EN; Adams-Bashforth) -> rows of (bounce, h_closed_form, h_mfix)
  DEM03  tests/dem/DEM03/AUTOTEST/POST_POS1.dat, POST_POS2.dat
         (scheme, KN, EN) -> rows of (time, y1_mfix, y2_mfix)
  DEM04  tests/dem/DEM04/AUTOTEST/POST_TIME.dat, POST_TVEL.dat, POST_AVEL.dat
         MEW -> (t_end*MEW*g/u0, v/u0, omega*r/u0) at the end of slip
  DEM05  tests/dem/DEM05/particle_input.dat (the 93 particles) and
         AUTOTEST/POST_ALPHA.dat, POST_COEFF.dat, POST_OMEGA.dat
         (LSD, HERTZ) -> rows of (impact angle, part-part, part-wall)
  DEM06  tests/dem/DEM06/AUTOTEST/POST_POS.dat, POST_VEL.dat: the three
         one-way coupled runs (frozen gas; no interpolation, GARG_2012,
         SQUARE_DPVM width 2 mm) -> rows of (time, y_mfix, v_mfix)
  CGDEM03 tests/cgdem/CGDEM03/POST_VOL.dat -> rows of (time, solids volume,
         fluid volume, |1 - sum|) of the LINEAR_HAT deposition of 480 spheres
  particle_input/DEMnn.dat  the particle_input.dat files of DEM03, DEM05
         (Version 1.0), DEM07, DEM08 and CGDEM03 (Version 2.0), copied as they are
         (input data of the reference's tests, read by the particle-input
         reader's tests)
Run from the repository root when /root/reference is mounted:
  python tests/golden/make_dem_fixtures.py
"""
import json
import os
import re

REF = "/root/reference/tests/dem"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_blocks(path, ncol=6, col=3):
    blocks, cur = [], None
    for line in open(path):
        m = re.search(r"Time Stepping Scheme:\s*(\S+)", line)
        if m:
            cur = {"scheme": m.group(1), "rows": []}
            blocks.append(cur)
            continue
        m = re.search(r"^\s*KN =\s*(\S+)", line)
        if m and cur is not None:
            cur["kn"] = float(m.group(1))
            continue
        m = re.search(r"DES_EN_INPUT =\s*(\S+)", line)
        if m and cur is not None:
            cur["en"] = float(m.group(1))
            continue
        parts = line.split()
        if cur is not None and len(parts) == ncol:
            try:
                vals = [float(x) for x in parts]
            except ValueError:
                continue
            cur["rows"].append([vals[0], vals[col]])  # time, MFIX column
    return blocks


def parse_table(path, ncol=4):
    """Rows of `ncol` numbers in a 4-column AUTOTEST table (key, analytic,
    MFIX, %err), in file order."""
    rows = []
    for line in open(path):
        parts = line.split()
        if len(parts) != ncol:
            continue
        try:
            rows.append([float(x) for x in parts])
        except ValueError:
            continue
    return rows


def dem02():
    """POST_HEIGHT.dat: per run a header naming KN and DES_EN_INPUT, then one
    line per bounce apex (bounce, closed-form height, MFiX height, %diff)."""
    runs, cur = [], None
    for line in open(os.path.join(REF, "DEM02", "AUTOTEST", "POST_HEIGHT.dat")):
        m = re.match(r"\s*KN =\s*(\S+)", line)
        if m:
            cur = {"scheme": "ADAMS_BASHFORTH", "kn": float(m.group(1)), "rows": []}
            runs.append(cur)
            continue
        m = re.match(r"\s*DES_EN_INPUT =\s*(\S+)", line)
        if m:
            cur["en"] = float(m.group(1))
            continue
        f = line.split()
        if cur is not None and len(f) == 4 and re.fullmatch(r"\d+", f[0]):
            cur["rows"].append([int(f[0]), float(f[1]), float(f[2])])
    with open(os.path.join(HERE, "dem02_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM02/AUTOTEST/POST_HEIGHT.dat",
                   "columns": ["bounce", "h_closed_form", "h_mfix"], "runs": runs}, f,
                  separators=(",", ":"))
    print("DEM02:", len(runs), "runs,", sum(len(r["rows"]) for r in runs), "rows")


def dem03():
    b1 = parse_blocks(os.path.join(REF, "DEM03", "AUTOTEST", "POST_POS1.dat"), ncol=4, col=2)
    b2 = parse_blocks(os.path.join(REF, "DEM03", "AUTOTEST", "POST_POS2.dat"), ncol=4, col=2)
    assert len(b1) == len(b2)
    out = []
    for a, b in zip(b1, b2):
        assert a["en"] == b["en"] and len(a["rows"]) == len(b["rows"])
        rows = [[p[0], p[1], q[1]] for p, q in zip(a["rows"], b["rows"])]
        out.append({"scheme": a["scheme"], "kn": a["kn"], "en": a["en"], "rows": rows})
    with open(os.path.join(HERE, "dem03_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM03/AUTOTEST/POST_POS1.dat, POST_POS2.dat",
                   "columns": ["time", "y1_mfix", "y2_mfix"], "runs": out}, f,
                  separators=(",", ":"))
    print("DEM03:", len(out), "runs,", sum(len(b["rows"]) for b in out), "rows")


def dem04():
    d = os.path.join(REF, "DEM04", "AUTOTEST")
    t = parse_table(os.path.join(d, "POST_TIME.dat"))
    v = parse_table(os.path.join(d, "POST_TVEL.dat"))
    w = parse_table(os.path.join(d, "POST_AVEL.dat"))
    assert len(t) == len(v) == len(w)
    runs = [{"mew": a[0], "time_nd": a[2], "tvel_nd": b[2], "avel_nd": c[2]}
            for a, b, c in zip(t, v, w)]
    with open(os.path.join(HERE, "dem04_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM04/AUTOTEST/POST_TIME.dat, POST_TVEL.dat, "
                             "POST_AVEL.dat (MFIX column)", "runs": runs}, f, indent=0)
    print("DEM04:", len(runs), "runs")


def dem05_table(path):
    """Blocks of (angle, part-part, part-wall) rows, one per collision
    model, skipping the experimental rows at the top of the file."""
    blocks, cur = [], None
    for line in open(path):
        if "Collision model:" in line:
            cur = {"model": "HERTZ" if "HERTZ" in line.upper() else "LSD", "rows": []}
            blocks.append(cur)
            continue
        parts = line.split()
        if cur is None or len(parts) != 3:
            continue
        try:
            cur["rows"].append([float(x) for x in parts])
        except ValueError:
            continue
    return blocks


def dem05():
    d = os.path.join(REF, "DEM05")
    parts = [[float(x) for x in line.split()] for line in open(os.path.join(d, "particle_input.dat"))
             if line.strip()]
    tabs = {k: dem05_table(os.path.join(d, "AUTOTEST", f"POST_{k.upper()}.dat"))
            for k in ("alpha", "coeff", "omega")}
    runs = []
    for i, blk in enumerate(tabs["alpha"]):
        runs.append({"model": blk["model"], "alpha": blk["rows"],
                     "coeff": tabs["coeff"][i]["rows"], "omega": tabs["omega"][i]["rows"]})
        assert tabs["coeff"][i]["model"] == blk["model"] == tabs["omega"][i]["model"]
    with open(os.path.join(HERE, "dem05_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM05/particle_input.dat (x y z r rho u v w); "
                             "AUTOTEST/POST_ALPHA.dat, POST_COEFF.dat, POST_OMEGA.dat",
                   "columns": ["impact_angle", "part_part", "part_wall"],
                   "particles": parts, "runs": runs}, f, separators=(",", ":"))
    print("DEM05:", len(parts), "particles,", [(r["model"], len(r["alpha"])) for r in runs])


def dem06_blocks(path):
    """Per run: the rows (time, analytic, MFIX, %diff) of one AUTOTEST block."""
    out, cur = [], None
    for line in open(path):
        if "Time" in line and "MFIX" in line:
            cur = []
            out.append(cur)
            continue
        parts = line.split()
        if cur is not None and len(parts) == 4:
            try:
                cur.append([float(x) for x in parts])
            except ValueError:
                pass
    return out


def dem06():
    d = os.path.join(REF, "DEM06", "AUTOTEST")
    pos = dem06_blocks(os.path.join(d, "POST_POS.dat"))
    vel = dem06_blocks(os.path.join(d, "POST_VEL.dat"))
    # runtest order: the first three runs are DES_ONEWAY_COUPLED
    runs = []
    for (scheme, width), bp, bv in zip([("NONE", None), ("GARG_2012", None),
                                        ("SQUARE_DPVM", 2.0e-3)], pos[:3], vel[:3]):
        assert len(bp) == len(bv)
        rows = [[a[0], a[2], b[2]] for a, b in zip(bp, bv)]
        runs.append({"scheme": scheme, "width": width, "rows": rows})
    with open(os.path.join(HERE, "dem06_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM06/AUTOTEST/POST_POS.dat, POST_VEL.dat "
                             "(runs 1-3, DES_ONEWAY_COUPLED)",
                   "columns": ["time", "y_mfix", "vy_mfix"], "runs": runs}, f,
                  separators=(",", ":"))
    print("DEM06:", len(runs), "runs,", sum(len(r["rows"]) for r in runs), "rows")


def main():
    dem02()
    dem06()
    dem03()
    dem05()
    dem04()
    pos = parse_blocks(os.path.join(REF, "DEM01", "AUTOTEST", "POST_POS.dat"))
    vel = parse_blocks(os.path.join(REF, "DEM01", "AUTOTEST", "POST_VEL.dat"))
    assert len(pos) == len(vel)
    out = []
    for bp, bv in zip(pos, vel):
        assert bp["scheme"] == bv["scheme"] and len(bp["rows"]) == len(bv["rows"])
        rows = [[p[0], p[1], v[1]] for p, v in zip(bp["rows"], bv["rows"])]
        out.append({"scheme": bp["scheme"], "kn": bp["kn"], "en": bp["en"], "rows": rows})
    with open(os.path.join(HERE, "dem01_post.json"), "w") as f:
        json.dump({"source": "tests/dem/DEM01/AUTOTEST/POST_POS.dat, POST_VEL.dat",
                   "columns": ["time", "y_mfix", "vy_mfix"], "runs": out}, f,
                  separators=(",", ":"))
    print(len(out), "runs,", sum(len(b["rows"]) for b in out), "rows")


def particle_inputs():
    import shutil
    out = os.path.join(HERE, "particle_input")
    os.makedirs(out, exist_ok=True)
    for t in ("DEM03", "DEM05", "DEM07", "DEM08"):
        shutil.copyfile(os.path.join(REF, t, "particle_input.dat"), os.path.join(out, f"{t}.dat"))
    cg = os.path.join(os.path.dirname(REF), "cgdem", "CGDEM03")
    shutil.copyfile(os.path.join(cg, "particle_input.dat"), os.path.join(out, "CGDEM03.dat"))
    print("particle_input: DEM03 DEM05 DEM07 DEM08 CGDEM03")


def cgdem03():
    rows = parse_table(os.path.join(os.path.dirname(REF), "cgdem", "CGDEM03", "POST_VOL.dat"))
    with open(os.path.join(HERE, "cgdem03_post_vol.json"), "w") as f:
        json.dump({"source": "tests/cgdem/CGDEM03/POST_VOL.dat",
                   "columns": ["time", "solids_volume", "fluid_volume", "abs_error"],
                   "rows": rows}, f, separators=(",", ":"))
    print("CGDEM03:", len(rows), "rows")


if __name__ == "__main__":
    main()
    particle_inputs()
    cgdem03()

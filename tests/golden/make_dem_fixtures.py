"""Extract the reference's DEM regression data into compact JSON fixtures.

Reads the AUTOTEST files the reference ships for its DEM tests (data, not
code) and keeps, per run, the solver columns the tests compare:
  DEM01  tests/dem/DEM01/AUTOTEST/POST_POS.dat, POST_VEL.dat
         (scheme, KN, EN) -> rows of (time, y_mfix, v_mfix)
Run from the repository root when /root/reference is mounted:
  python tests/golden/make_dem_fixtures.py
"""
import json
import os
import re

REF = "/root/reference/tests/dem"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_blocks(path):
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
        if cur is not None and len(parts) == 6:
            try:
                vals = [float(x) for x in parts]
            except ValueError:
                continue
            cur["rows"].append([vals[0], vals[3]])  # time, MFIX column
    return blocks


def main():
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


if __name__ == "__main__":
    main()

"""MFiX's particle_input.dat, read and written on the host.

Reads the two formats of read_particle_input.f for 3D runs:

* Version 1.0 (no version line): one particle per record, list-directed
  ``x y z r rho u v w`` -- radius, not diameter (READ_PART_INPUT_V1P0,
  read_particle_input.f:3517-3624).  The reference takes PARTICLES records
  from mfix.dat; without it every record of the file is read.
* Version 2.0: 18 lines of text, ``Dimension:`` and ``Particles:``, 3
  lines, one line per variable with a Read flag and a default
  (Coordinates, Phase_ID, Diameter, Density, Velocity, Temperature,
  Species, User_Scalar), 3 lines, then the data, one column per variable
  read (READ_PART_INPUT_V2P0, :156-631; parse_scalar_option and
  parse_vector_option, :3366-3507).  Temperature, species and user scalars
  take columns only when their equations are solved, which this library
  does not do, so they never take columns here.

Version 4.0 (particle_input.csv) is not read.  Fortran list-directed
numbers (``1.0D-3``, commas between values) are accepted.
"""
from __future__ import annotations

import re

import numpy as np

from .workloads import Particles, make_particles

_SPLIT = re.compile(r"[\s,]+")


def _numbers(line: str, count: int | None = None) -> list[float]:
    toks = [t for t in _SPLIT.split(line.strip()) if t]
    vals = [float(t.replace("D", "E").replace("d", "e")) for t in toks]
    if count is not None:
        if len(vals) < count:
            raise ValueError(f"expected {count} values, found {len(vals)}: {line!r}")
        vals = vals[:count]
    return vals


def _v1(lines: list[str], particles: int | None) -> Particles:
    rows = [ln for ln in lines if ln.strip()]
    if particles is not None:
        if len(rows) < particles:
            raise ValueError(f"end of file found for particle {len(rows) + 1} and "
                             f"{particles} entries are expected")
        rows = rows[:particles]
    data = np.array([_numbers(ln, 8) for ln in rows], dtype=np.float64).reshape(-1, 8)
    return make_particles(data[:, 0:3].T, data[:, 3], data[:, 4], vel=data[:, 5:8].T)


def _option(line: str, size: int):
    """(read?, default values) of one variable line of a Version 2.0 header."""
    toks = [t for t in _SPLIT.split(line.strip()) if t]
    flag = toks[1].upper() if len(toks) > 1 else ""
    if flag == "T":
        return True, None
    if flag == "F":
        return False, np.array([float(t.replace("D", "E")) for t in toks[2:2 + size]])
    raise ValueError(f"Read flag must be T or F: {line!r}")


def _v2(lines: list[str], particles: int | None) -> Particles:
    dim = int(_numbers(lines[18].split(":", 1)[1])[0])
    if dim != 3:
        raise ValueError(f"particle_input.dat is {dim}D; this library runs 3D")
    n_file = int(_numbers(lines[19].split(":", 1)[1])[0])
    n = n_file if particles is None else particles
    specs = [("Coordinates", 3), ("Phase_ID", 1), ("Diameter", 1), ("Density", 1),
             ("Velocity", 3)]
    start, ncol = {}, 0
    defaults = {}
    for k, (name, size) in enumerate(specs):
        line = lines[23 + k]
        if not line.strip().lower().startswith(name.lower()):
            raise ValueError(f"expected the {name} line, found {line!r}")
        if name == "Coordinates":
            toks = [t for t in _SPLIT.split(line.strip()) if t]
            if len(toks) < 2 or toks[1].upper() != "T":
                raise ValueError("Coordinates must always be read (T)")
        read, default = _option(line, size)
        if read:
            start[name] = ncol
            ncol += size
        else:
            defaults[name] = default
    rows = [ln for ln in lines[34:] if ln.strip()]
    if len(rows) < n:
        raise ValueError(f"end of file found for particle {len(rows) + 1} and {n} entries "
                         "are expected")
    data = np.array([_numbers(ln, ncol) for ln in rows[:n]], dtype=np.float64).reshape(-1, ncol)

    def column(name, size):
        if name in start:
            return data[:, start[name]:start[name] + size]
        return np.broadcast_to(defaults[name], (n, size))
    pos = column("Coordinates", 3).T
    phase = np.rint(column("Phase_ID", 1)[:, 0]).astype(np.int32)
    radius = 0.5 * column("Diameter", 1)[:, 0]
    rho = column("Density", 1)[:, 0]
    vel = column("Velocity", 3).T
    return make_particles(pos, radius, rho, vel=vel, phase=phase)


def read_particle_input(path: str, particles: int | None = None) -> Particles:
    """Particles of a particle_input.dat (Version 1.0 or 2.0, 3D).
    ``particles`` is mfix.dat's PARTICLES (None: as many as the file holds;
    Version 2.0 then uses its own ``Particles:`` count)."""
    with open(path) as f:
        lines = f.read().splitlines()
    if lines and lines[0][:11].strip().startswith("Version"):
        version = lines[0][:11].strip()
        if version == "Version 2.0":
            return _v2(lines, particles)
        raise ValueError(f"{version} of particle_input is not read (Version 1.0 and 2.0 are)")
    return _v1(lines, particles)


def write_particle_input(path: str, parts: Particles) -> None:
    """Version 1.0 records ``x y z r rho u v w``, every value with 17
    significant digits, so that reading the file back gives the same
    doubles."""
    with open(path, "w") as f:
        for q in range(parts.n):
            vals = (*parts.pos[:, q], parts.radius[q], parts.ro_sol[q], *parts.vel[:, q])
            f.write(" ".join(f"{v:.17g}" for v in vals) + "\n")

"""MFiX's particle input files, read and written on the host.

Reads the formats of read_particle_input.f for 3D runs:

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
* Version 3.0: the Version 2.0 layout with 19 lines of text before
  ``Dimension:`` and the PART_IN_* filters applied to every particle
  (READ_PART_INPUT_V3P0, :641-1181); user scalars take columns when
  DES_USR_VAR_SIZE > 0.
* Version 4.0 (particle_input.csv, read_particle_input.f:2056-3360): a
  header line of comma-separated, case-insensitive names (x, y, z,
  phase_id, diameter, density, u, v, w, t_s / temperature, species*,
  usr_var*) and one comma-separated row per particle; a missing phase_id
  is 1, a missing diameter / density is D_P0 / RO_S0 of the phase, a
  missing velocity 0.  The number of columns must equal the number the
  header implies (temperature and species count only when solved, user
  scalars when DES_USR_VAR_SIZE > 0).  Particles outside the domain
  (x <= X_MIN or x >= X_MAX, likewise y and z) and those the PART_IN_*
  filters reject are dropped.

Fortran list-directed numbers (``1.0D-3``, commas between values) are
accepted.
"""
from __future__ import annotations

import re

import numpy as np

import dataclasses

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


UNDEFINED = 9.87654321e31   # param1_mod.f:20


@dataclasses.dataclass
class PartIn:
    """The mfix.dat keywords the Version 3.0 / 4.0 readers use.  Filters
    (des_init_namelist.f:541-640): PART_IN_<var> = (min, max, exclude) for
    var in x, y, z, diameter, density, u, v, w; a variable is filtered only
    when min or max is defined; exclude=False keeps min < v < max, True
    drops min <= v <= max (filter_single_particle_based_on_min_max_var,
    read_particle_input.f:5676-5714).  part_in_phase[m-1] False drops phase
    m."""
    particles: int | None = None
    des_mmax: int = 1
    d_p0: tuple = ()
    ro_s0: tuple = ()
    des_usr_var_size: int = 0
    energy_eq: bool = False
    species: int = 0                          # DIMENSION_N_S when solids species are solved
    domain: tuple = (0.0, UNDEFINED, 0.0, UNDEFINED, 0.0, UNDEFINED)  # X_MIN, X_MAX, ...
    filters: dict = dataclasses.field(default_factory=dict)
    part_in_phase: tuple = ()


def _keep(v, lo, hi, exclude, keep):
    """filter_single_particle_based_on_min_max_var on arrays."""
    if lo == -UNDEFINED and hi == UNDEFINED:
        return keep
    if exclude:
        return keep & ~((lo <= v) & (v <= hi))
    return keep & ~((v <= lo) | (v >= hi))


def _filtered(pos, radius, rho, vel, phase, usr, opt: PartIn, domain=False):
    keep = np.ones(radius.shape[0], dtype=bool)
    f = opt.filters
    for k, v in (("x", pos[0]), ("y", pos[1]), ("z", pos[2])):
        keep = _keep(v, *f.get(k, (-UNDEFINED, UNDEFINED, False)), keep)
    if opt.part_in_phase:
        ok = np.array([bool(opt.part_in_phase[m - 1]) if 1 <= m <= len(opt.part_in_phase)
                       else True for m in phase])
        keep &= ok
    keep = _keep(2.0 * radius, *f.get("diameter", (-UNDEFINED, UNDEFINED, False)), keep)
    if domain:
        lo_hi = opt.domain
        for d in range(3):
            keep = _keep(pos[d], lo_hi[2 * d], lo_hi[2 * d + 1], False, keep)
    keep = _keep(rho, *f.get("density", (-UNDEFINED, UNDEFINED, False)), keep)
    for k, d in (("u", 0), ("v", 1), ("w", 2)):
        keep = _keep(vel[d], *f.get(k, (-UNDEFINED, UNDEFINED, False)), keep)
    parts = make_particles(pos[:, keep], radius[keep], rho[keep], vel=vel[:, keep],
                           phase=phase[keep])
    return parts, (usr[:, keep] if usr is not None else None)


def _v3(lines: list[str], opt: PartIn):
    if len(lines) < 35:
        raise ValueError(f"Version 3.0 header is incomplete ({len(lines)} lines, 35 expected)")
    dim = int(_numbers(lines[19].split(":", 1)[1])[0])
    if dim != 3:
        raise ValueError(f"particle_input.dat is {dim}D; this library runs 3D")
    n_file = int(_numbers(lines[20].split(":", 1)[1])[0])
    n = n_file if opt.particles is None else opt.particles
    nu = opt.des_usr_var_size
    specs = [("Coordinates", 3, True), ("Phase_ID", 1, True), ("Diameter", 1, True),
             ("Density", 1, True), ("Velocity", 3, True),
             ("Temperature", 1, opt.energy_eq), ("Species", max(opt.species, 1), opt.species > 0),
             ("User scalar", max(nu, 1), nu > 0)]
    start, ncol, defaults = {}, 0, {}
    for k, (name, size, active) in enumerate(specs):
        # list-directed: a name token, a flag token, the defaults
        # (parse_scalar_option / parse_vector_option, :3366-3507)
        line = lines[24 + k]
        toks = [t for t in _SPLIT.split(line.strip()) if t]
        if not toks or not toks[0].lower().startswith(name.split()[0].lower()):
            raise ValueError(f"expected the {name} line, found {line!r}")
        if not active:
            continue
        flag = toks[1][:1].upper() if len(toks) > 1 else ""
        if flag == "T":
            start[name] = ncol
            ncol += size
        elif flag == "F":
            defaults[name] = np.array([float(t.replace("D", "E")) for t in toks[2:2 + size]])
        else:
            raise ValueError(f"Read flag must be T or F: {line!r}")
        if name == "Coordinates" and flag != "T":
            raise ValueError("Coordinates must always be read (T)")
    rows = [ln for ln in lines[35:] if ln.strip()]
    if len(rows) < n:
        raise ValueError(f"end of file found for particle {len(rows) + 1} and {n} entries "
                         "are expected")
    data = np.array([_numbers(ln, ncol) for ln in rows[:n]], dtype=np.float64).reshape(-1, ncol)

    def column(name, size):
        if name in start:
            return data[:, start[name]:start[name] + size]
        return np.broadcast_to(defaults[name], (n, size))
    pos = np.ascontiguousarray(column("Coordinates", 3).T)
    phase = np.trunc(column("Phase_ID", 1)[:, 0]).astype(np.int32)
    radius = 0.5 * column("Diameter", 1)[:, 0]
    rho = np.ascontiguousarray(column("Density", 1)[:, 0])
    vel = np.ascontiguousarray(column("Velocity", 3).T)
    usr = np.ascontiguousarray(column("User scalar", nu).T) if nu > 0 else None
    return _filtered(pos, radius, rho, vel, phase, usr, opt)


def read_particle_input_csv(path: str, opt: PartIn):
    """particle_input.csv (Version 4.0).  Returns (Particles, user scalars
    (DES_USR_VAR_SIZE, n) or None)."""
    with open(path) as f:
        lines = f.read().splitlines()
    header = [h.strip().lower() for h in lines[0].split(",")]
    rows = [ln for ln in lines[1:] if ln.strip()]
    n = len(rows) if opt.particles is None else opt.particles
    if n > len(rows):
        raise ValueError(f"Expecting {n} particles in particle_input.csv, but there are only "
                         f"{len(rows)}.")
    col = {}
    expect = 3
    sp_end = uv_end = -1
    for i, h in enumerate(header):
        if h == "x":
            col["x"] = i
        elif "phase_id" in h:
            col["phase"] = i
            expect += 1
        elif "diameter" in h:
            col["diameter"] = i
            expect += 1
        elif "density" in h:
            col["density"] = i
            expect += 1
        elif h == "u":
            col["u"] = i
            expect += 3
        elif "t_s" in h or "temperature" in h:
            col["t_s"] = i
        elif "species" in h:
            sp_end = max(sp_end, i)
        elif "usr_var" in h:
            uv_end = max(uv_end, i)
    if opt.energy_eq and "t_s" in col:
        expect += 1
    if opt.species > 0 and sp_end >= 0:
        expect += opt.species
    nu = opt.des_usr_var_size
    if nu > 0 and uv_end >= 0:
        expect += nu
    if expect != len(header):
        raise ValueError(f"The expected number of columns does not match the actual number of "
                         f"columns in particle_input.csv, expected number of columns = {expect}, "
                         f"number of columns in file = {len(header)}.")
    if "x" not in col:
        raise ValueError("Missing particle coordinates info from particle_input.csv")
    if "phase" not in col and opt.des_mmax > 1:
        raise ValueError("Missing phase info from particle_input.csv when number of discrete "
                         "solids phases is larger than 1.")
    data = np.array([_numbers(ln, len(header)) for ln in rows[:n]],
                    dtype=np.float64).reshape(-1, len(header))
    x0 = col["x"]
    pos = np.ascontiguousarray(data[:, x0:x0 + 3].T)
    phase = (np.trunc(data[:, col["phase"]]).astype(np.int32) if "phase" in col
             else np.ones(n, dtype=np.int32))
    if np.any(phase < 1) or np.any(phase > opt.des_mmax):
        raise ValueError("phase_id outside 1..DES_MMAX")
    if "diameter" in col:
        radius = 0.5 * data[:, col["diameter"]]
    else:
        if not opt.d_p0:
            raise ValueError("Missing diameter info from particle_input.csv and the keyword "
                             "D_P0 undefined.")
        radius = 0.5 * np.array([opt.d_p0[m - 1] for m in phase])
    if "density" in col:
        rho = np.ascontiguousarray(data[:, col["density"]])
    else:
        if not opt.ro_s0:
            raise ValueError("Missing density info from particle_input.csv and the keyword "
                             "RO_S0 undefined.")
        rho = np.array([opt.ro_s0[m - 1] for m in phase])
    vel = (np.ascontiguousarray(data[:, col["u"]:col["u"] + 3].T) if "u" in col
           else np.zeros((3, n)))
    usr = None
    if nu > 0:
        usr = (np.ascontiguousarray(data[:, uv_end - nu + 1:uv_end + 1].T) if uv_end >= 0
               else np.ones((nu, n)))
    return _filtered(pos, radius, rho, vel, phase, usr, opt, domain=True)


def read_particle_input(path: str, particles: int | None = None,
                        opt: PartIn | None = None):
    """Particles of a particle_input.dat (Version 1.0, 2.0 or 3.0, 3D).
    ``particles`` is mfix.dat's PARTICLES (None: as many as the file holds;
    Versions 2.0 and 3.0 then use their own ``Particles:`` count).  A
    Version 3.0 file returns (Particles, user scalars or None) and takes its
    keywords from ``opt``."""
    with open(path) as f:
        lines = f.read().splitlines()
    if lines and lines[0][:11].strip().startswith("Version"):
        version = lines[0][:11].strip()
        if version == "Version 2.0":
            return _v2(lines, particles)
        if version == "Version 3.0":
            o = dataclasses.replace(opt or PartIn(), particles=particles if particles is not None
                                    else (opt.particles if opt else None))
            return _v3(lines, o)
        raise ValueError(f"{version} of particle_input.dat is not read (Versions 1.0, 2.0 and "
                         f"3.0 are; 4.0 is particle_input.csv)")
    return _v1(lines, particles)


def write_particle_input(path: str, parts: Particles) -> None:
    """Version 1.0 records ``x y z r rho u v w``, every value with 17
    significant digits, so that reading the file back gives the same
    doubles."""
    with open(path, "w") as f:
        for q in range(parts.n):
            vals = (*parts.pos[:, q], parts.radius[q], parts.ro_sol[q], *parts.vel[:, q])
            f.write(" ".join(f"{v:.17g}" for v in vals) + "\n")

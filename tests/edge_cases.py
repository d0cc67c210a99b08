"""Edge-case workloads shared by the CPU (oracle) and GPU edge tests:
empty and partly filled systems, a coincident pair, an overfilled fluid
cell, a rogue particle and a particle leaving the domain."""
import numpy as np

import _paths  # noqa: F401
from mfixdem import abi, workloads


def small_box(n_cap, coupled=False, remove_rogue=1):
    p, _, _ = workloads.settling_bed(n=n_cap, seed=1, box=(0.01, 0.02, 0.01))
    p.remove_rogue = remove_rogue
    if coupled:
        p.coupled = 1
        p.imax, p.jmax, p.kmax = 5, 10, 5
        p.drag_type = abi.DRAG_WEN_YU
        p.interp_scheme = abi.INTERP_GARG_2012
        p.des_interp_on = p.des_interp_mean_fields = 1
        p.d_p0[0] = 1.0e-3
        p.ro_s0[0] = 2500.0
    return p


def uniform_gas(p, v=0.3):
    n = workloads.n_cells(p)
    return workloads.Gas(np.zeros(n), np.full(n, v), np.zeros(n), np.zeros((n, 3)))


def empty():
    p = small_box(100)
    return p, workloads.make_particles(np.zeros((3, 0)), np.zeros(0), 2500.0)


def partly_filled(seed=3):
    """Capacity 3000, 2000 particles (the rest of the slots nonexistent)."""
    p, parts, _ = workloads.settling_bed(n=2000, seed=seed, box=(0.01, 0.035, 0.01))
    p.max_pip = 3000
    return workloads.jostled(p, parts, seed=seed)


def coincident():
    """Two spheres at the same point (calc_force_dem.f:267-271)."""
    p = small_box(2)
    pos = np.array([[0.005, 0.005], [0.01, 0.01], [0.005, 0.005]])
    return p, workloads.make_particles(pos, 0.5e-3, 2500.0)


def overfilled_cell():
    """Solids volume above the volume of a fluid cell: EP_G < 0
    (calc_epg_des.f:146-203).  32 overlapping spheres (16.8 mm^3) inside
    one 2 mm cell (8 mm^3); no interpolation, so all of it stays there."""
    p = small_box(32, coupled=True)
    p.interp_scheme = abi.INTERP_NONE
    p.des_interp_on = p.des_interp_mean_fields = 0
    rng = np.random.default_rng(4)
    pos = np.array([0.005, 0.011, 0.005])[:, None] + rng.uniform(-0.5e-3, 0.5e-3, (3, 32))
    return p, workloads.make_particles(pos, 0.5e-3, 2500.0)


def rogue(remove_rogue):
    """One sphere placed outside the box (in a ghost DES cell), the others
    inside: removed at the first rebuild, or ROGUE when removal is off
    (desgrid_mod.f:1074-1104)."""
    p = small_box(3, remove_rogue=remove_rogue)
    pos = np.array([[0.003, 0.007, 0.005], [0.005, 0.005, 0.021], [0.005, 0.005, 0.005]])
    return p, workloads.make_particles(pos, 0.5e-3, 2500.0)


def leaving():
    """A coupled system where one sphere hits the top wall at 60 m/s, goes
    through it (the soft wall cannot stop it within its radius) and is
    removed as a rogue particle at the next rebuild."""
    p = small_box(2, coupled=True)
    p.grav[:] = [0.0, 0.0, 0.0]
    pos = np.array([[0.003, 0.007], [0.005, 0.0199], [0.005, 0.005]])
    vel = np.array([[0.0, 0.0], [0.0, 60.0], [0.0, 0.0]])
    return p, workloads.make_particles(pos, 0.5e-3, 2500.0, vel)

"""The deposition case of the reference's CGDEM03 test
(tests/cgdem/CGDEM03/advection.mfx, particle_input.dat): 480 spheres, d =
0.01 m, rho = 2700, at rest in a 1 m cube of 32^3 fluid cells, some within
a radius of the walls; LINEAR_HAT interpolation with DES_INTERP_MEAN_FIELDS,
desgrid 32^3.  The gas arrays come from the fluidised-bed builder at the
same grid (a gas at rest); only the deposition is compared."""
import os

from mfixdem import abi, workloads
from mfixdem.particle_input import read_particle_input

HERE = os.path.dirname(os.path.abspath(__file__))
CELLS = 32


def cgdem03_case():
    parts = read_particle_input(os.path.join(HERE, "golden", "particle_input", "CGDEM03.dat"))
    p, _, gas = workloads.fluidised_bed(n=100, seed=1, lx=1.0, ly=1.0, lz=1.0,
                                        cells=(CELLS, CELLS, CELLS), dg=(0, 0, 0))
    p.interp_scheme = abi.INTERP_LINEAR_HAT
    p.des_interp_on = 1
    p.des_interp_width = abi.UNDEFINED
    p.max_pip = parts.n
    p.max_radius = 0.005
    p.d_p0[0] = 0.01
    p.ro_s0[0] = 2700.0
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = CELLS
    return p, parts, gas


def volumes(ep, rop):
    """Solids volume (ROP_S / RO_S0 over the fluid cells) and fluid volume
    (EP_G over the fluid cells) of the 1 m cube, as POST_VOL.dat prints."""
    n = CELLS + 2
    vol = (1.0 / CELLS) ** 3
    e = ep.reshape(n, n, n)[1:-1, 1:-1, 1:-1]
    r = rop.reshape(-1, n, n, n)[:, 1:-1, 1:-1, 1:-1]
    return r.sum() * vol / 2700.0, e.sum() * vol

"""The reference's DEM02 user hooks restated for any system with step(nn)
and get_particles(): USR0 records the release height
(tests/dem/DEM02/usr0.f), USR2_DES, called after every DEM step, records
max(previous y, current y) each time the vertical velocity turns from
positive to negative (tests/dem/DEM02/usr2_des.f), for at most MAX_BOUNCE =
50 bounces (tests/dem/DEM02/usr_mod.f).  The run lasts FACTOR =
CEILING(REAL((TSTOP-TIME)/DTSOLID)) steps, TSTOP = 10 s
(model/des/des_time_march.f:130, tests/dem/DEM02/mfix.dat)."""
import math

import numpy as np

MAX_BOUNCE = 50
TSTOP = 10.0


def dem02_steps(dtsolid):
    return math.ceil(float(np.float32(TSTOP / dtsolid)))


def bounce_heights(system, nsteps, y0=0.5):
    heights = [y0]
    y_old, v_old = y0, 0.0
    for nn in range(1, nsteps + 1):
        system.step(nn)
        pos, vel = system.get_particles()[:2]
        y, v = pos[1, 0], vel[1, 0]
        if v < 0.0 and v_old > 0.0 and len(heights) <= MAX_BOUNCE:
            heights.append(max(y_old, y))
        y_old, v_old = y, v
    return np.array(heights)

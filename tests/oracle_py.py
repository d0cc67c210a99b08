"""ctypes driver of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: same method names as mfixdem.system.DemSystem so a
parity test can run the same call sequence on both.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import _paths  # noqa: F401  (puts the package on sys.path)
from mfixdem import abi
from mfixdem.workloads import Gas, Particles, n_cells

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

_libs = {}


def load_oracle(omp: bool = False):
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int32, C.POINTER(C.c_double)
    IP, BP, L = C.POINTER(C.c_int32), C.POINTER(C.c_int8), C.c_int64
    sig = {
        "orc_create": ([C.POINTER(abi.Params)], P),
        "orc_destroy": ([P], None),
        "orc_get_derived": ([P, C.POINTER(abi.Derived)], C.c_int),
        "orc_set_particles": ([P, L, D, D, D, D, D, IP, L], C.c_int),
        "orc_get_particles": ([P, D, D, D, BP, L], C.c_int),
        "orc_set_gas": ([P, D, D, D, D], C.c_int),
        "orc_get_gas": ([P, D, D], C.c_int),
        "orc_init": ([P], C.c_int),
        "orc_time_step": ([P, I], C.c_int),
        "orc_run": ([P, I, I], C.c_int),
        "orc_desgrid_pic": ([P], C.c_int),
        "orc_neighbour": ([P], C.c_int),
        "orc_calc_force_dem": ([P], C.c_int),
        "orc_calc_drag_des": ([P], C.c_int),
        "orc_cfnewvalues": ([P], C.c_int),
        "orc_particles_in_cell": ([P], C.c_int),
        "orc_comp_mean_fields": ([P], C.c_int),
        "orc_get_forces": ([P, D, D, L], C.c_int),
        "orc_get_cells": ([P, IP, IP, L], C.c_int),
        "orc_get_pairs": ([P, IP, IP, D, L, C.POINTER(C.c_int64)], C.c_int),
        "orc_get_status": ([P, C.POINTER(abi.Status)], C.c_int),
        "orc_dtsolid": ([P], C.c_double),
        "orc_nsquare_pairs": ([P, IP, L], C.c_int64),
        "orc_drag_fgp": ([I, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double], C.c_double),
        "orc_omp_threads": ([], C.c_int),
        "orc_set_omp_threads": ([C.c_int], None),
        "orc_user_set_particles": ([P, L, L, D, D, D], C.c_int),
        "orc_get_filter": ([P, L, IP, D], C.c_int),
        "orc_get_ugc": ([P, D], C.c_int),
        "orc_set_gas_pressure": ([P, D], C.c_int),
        "orc_get_p_force": ([P, D], C.c_int),
        "orc_drag_gs_gas": ([P, D, D, D], C.c_int),
        "orc_des_time_init": ([P], C.c_int),
        "orc_set_particle_state": ([P, C.POINTER(C.c_int8), L], C.c_int),
        "orc_set_facets": ([P, C.c_int32, D, D, C.c_int32], C.c_int),
        "orc_set_wall_motion": ([P, C.c_int32, C.c_int32, D, D, D], C.c_int),
        "orc_set_dem_bcs": ([P, C.c_int32, C.POINTER(abi.DemBc), C.c_uint64], C.c_int),
        "orc_particles_in_cell": ([P], C.c_int),
        "orc_calc_interp_weights": ([P], C.c_int),
        "orc_comp_mean_fields": ([P], C.c_int),
    }
    for n, (a, r) in sig.items():
        f = getattr(lib, n)
        f.argtypes = a
        f.restype = r
    _libs[name] = lib
    return lib


class OracleSystem:
    def __init__(self, params: abi.Params, omp: bool = False):
        self.lib = load_oracle(omp)
        self.params = params
        self.h = self.lib.orc_create(C.byref(params))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.n = params.max_pip
        self.ncell = n_cells(params)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            raise abi.MfixDemError(f"oracle {what}: status {rc} ({abi.ERRORS.get(rc, '?')})", rc)

    def derived(self):
        d = abi.Derived()
        self._chk(self.lib.orc_get_derived(self.h, C.byref(d)), "derived")
        return d

    def dtsolid(self):
        return self.lib.orc_dtsolid(self.h)

    def set_particles(self, parts: Particles):
        n = parts.n
        self._chk(self.lib.orc_set_particles(
            self.h, n, abi.dptr(parts.pos), abi.dptr(parts.vel), abi.dptr(parts.omega),
            abi.dptr(parts.radius), abi.dptr(parts.ro_sol), abi.iptr(parts.phase), n),
            "set_particles")

    def set_gas(self, gas: Gas):
        self._chk(self.lib.orc_set_gas(self.h, abi.dptr(gas.u_g), abi.dptr(gas.v_g),
                                       abi.dptr(gas.w_g), abi.dptr(gas.p_force)), "set_gas")

    def init(self):
        self._chk(self.lib.orc_init(self.h), "init")

    def step(self, nn):
        self._chk(self.lib.orc_time_step(self.h, nn), f"step {nn}")

    def run(self, nn_first, nsteps):
        self._chk(self.lib.orc_run(self.h, nn_first, nsteps), "run")

    def set_debug(self, flags):
        pass

    def desgrid_pic(self):
        self._chk(self.lib.orc_desgrid_pic(self.h), "desgrid_pic")

    def neighbour(self):
        self._chk(self.lib.orc_neighbour(self.h), "neighbour")

    def calc_force_dem(self):
        self._chk(self.lib.orc_calc_force_dem(self.h), "calc_force_dem")

    def get_particles(self):
        n = self.n
        pos, vel, om = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        st = np.zeros(n, dtype=np.int8)
        self._chk(self.lib.orc_get_particles(self.h, abi.dptr(pos), abi.dptr(vel),
                                             abi.dptr(om), abi.bptr(st), n), "get_particles")
        return pos, vel, om, st

    def get_forces(self):
        n = self.n
        fc, tow = np.zeros((3, n)), np.zeros((3, n))
        self._chk(self.lib.orc_get_forces(self.h, abi.dptr(fc), abi.dptr(tow), n), "get_forces")
        return fc, tow

    def get_cells(self):
        n = self.n
        dg = np.zeros(n, dtype=np.int32)
        pijk = np.zeros((4, n), dtype=np.int32)
        self._chk(self.lib.orc_get_cells(self.h, abi.iptr(dg), abi.iptr(pijk), n), "get_cells")
        return dg, pijk

    def get_pairs(self):
        npairs = C.c_int64(0)
        self._chk(self.lib.orc_get_pairs(self.h, abi.iptr(None), abi.iptr(None), abi.dptr(None),
                                         0, C.byref(npairs)), "get_pairs")
        m = max(int(npairs.value), 1)
        idx = np.zeros(self.n, dtype=np.int32)
        nb = np.zeros(m, dtype=np.int32)
        pft = np.zeros(3 * m)
        self._chk(self.lib.orc_get_pairs(self.h, abi.iptr(idx), abi.iptr(nb), abi.dptr(pft), m,
                                         C.byref(npairs)), "get_pairs")
        k = int(npairs.value)
        return idx, nb[:k], pft[:3 * k].reshape(k, 3)

    def get_gas(self):
        ep = np.zeros(self.ncell)
        rops = np.zeros(self.ncell * self.params.nphases)
        self._chk(self.lib.orc_get_gas(self.h, abi.dptr(ep), abi.dptr(rops)), "get_gas")
        return ep, rops.reshape(self.params.nphases, self.ncell)

    def status(self):
        s = abi.Status()
        self._chk(self.lib.orc_get_status(self.h, C.byref(s)), "status")
        return s

    def user_set_particles(self, q0, pos=None, vel=None, omega=None):
        """Edit particles q0.. (arrays of shape (3, count)) between steps,
        like a reference USR2_DES hook."""
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
               for a in (pos, vel, omega)]
        count = next(a.shape[1] for a in arr if a is not None)
        self._chk(self.lib.orc_user_set_particles(self.h, q0, count,
                                                  *(abi.dptr(a) for a in arr)),
                  "user_set_particles")

    def filter(self, q):
        """(cells 0-based, weights) of particle q after the last weights pass."""
        cells = np.zeros(27, dtype=np.int32)
        w = np.zeros(27)
        nf = self.lib.orc_get_filter(self.h, q, abi.iptr(cells), abi.dptr(w))
        return cells[:nf], w[:nf]

    def drag_gs_gas(self):
        """CALC_DES_2FLUID's drag (DRAG_GS_GAS0 for GARG_2012, else
        DRAG_GS_GAS1) on the current particles: (F_GDS, DRAG_AM, DRAG_BM[3])."""
        f = np.zeros(self.ncell)
        am = np.zeros(self.ncell)
        bm = np.zeros(3 * self.ncell)
        self._chk(self.lib.orc_drag_gs_gas(self.h, abi.dptr(f), abi.dptr(am), abi.dptr(bm)),
                  "drag_gs_gas")
        return f, am, bm.reshape(3, self.ncell)

    def particles_in_cell(self):
        self._chk(self.lib.orc_particles_in_cell(self.h), "particles_in_cell")

    def calc_interp_weights(self):
        self._chk(self.lib.orc_calc_interp_weights(self.h), "calc_interp_weights")

    def comp_mean_fields(self):
        self._chk(self.lib.orc_comp_mean_fields(self.h), "comp_mean_fields")

    def set_particle_state(self, state):
        st = np.ascontiguousarray(state, dtype=np.int8)
        self._chk(self.lib.orc_set_particle_state(self.h, st.ctypes.data_as(C.POINTER(C.c_int8)),
                                                  st.size), "set_particle_state")

    def set_facets(self, vertex, normal, default_walls=True):
        """STL facets before init: vertex (nf, 3 coords, 3 vertices) = MFiX's
        VERTEX(3,3,nf) memory order, normal (nf, 3)."""
        v = np.ascontiguousarray(vertex, dtype=np.float64)
        nv = np.ascontiguousarray(normal, dtype=np.float64)
        self._chk(self.lib.orc_set_facets(self.h, v.shape[0], abi.dptr(v), abi.dptr(nv),
                                          1 if default_walls else 0), "set_facets")

    def set_wall_motion(self, first, count, vel, omega=(0.0, 0.0, 0.0),
                        rot_center=(0.0, 0.0, 0.0)):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (vel, omega, rot_center)]
        self._chk(self.lib.orc_set_wall_motion(self.h, first, count, *[abi.dptr(x) for x in a]),
                  "set_wall_motion")

    def set_dem_bcs(self, bcs, seed=1):
        arr = (abi.DemBc * max(len(bcs), 1))(*bcs)
        self._chk(self.lib.orc_set_dem_bcs(self.h, len(bcs), arr, seed), "set_dem_bcs")

    def des_time_init(self):
        self._chk(self.lib.orc_des_time_init(self.h), "des_time_init")

    def set_gas_pressure(self, p_g):
        self._chk(self.lib.orc_set_gas_pressure(self.h, abi.dptr(np.ascontiguousarray(p_g))),
                  "set_gas_pressure")

    def get_p_force(self):
        out = np.zeros(3 * self.ncell)
        self.lib.orc_get_p_force(self.h, abi.dptr(out))
        return out.reshape(self.ncell, 3)

    def ugc(self):
        out = np.zeros(3 * self.ncell)
        self.lib.orc_get_ugc(self.h, abi.dptr(out))
        return out.reshape(self.ncell, 3)

    def nsquare_pairs(self):
        cap = 8 * self.n + 16
        buf = np.zeros(2 * cap, dtype=np.int32)
        k = self.lib.orc_nsquare_pairs(self.h, abi.iptr(buf), cap)
        if k > cap:
            buf = np.zeros(2 * k, dtype=np.int32)
            k = self.lib.orc_nsquare_pairs(self.h, abi.iptr(buf), k)
        return {(int(buf[2 * q]), int(buf[2 * q + 1])) for q in range(k)}

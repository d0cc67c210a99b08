"""Python driver of one DEM system on the GPU through the C ABI.

Mirrors the call sequence of the reference's DES time loop
(model/des/des_time_march.f:293-460): set_particles / set_gas (state in),
init (MAKE_ARRAYS_DES tail), step(NN) or run(NN0, K) (DES_TIME_STEP), and
getters that return results in global-id order and reference layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .workloads import Gas, Particles, n_cells


class DemSystem:
    def __init__(self, params: abi.Params):
        self.lib = abi.load_library()
        self.params = params
        self.h = C.c_void_p()
        abi.check(self.lib.mfixdem_create(C.byref(params), C.byref(self.h)), "mfixdem_create")
        self.n = params.max_pip
        self.ncell = n_cells(params)

    # lifetime -------------------------------------------------------------
    def close(self):
        if self.h:
            abi.check(self.lib.mfixdem_destroy(self.h), "mfixdem_destroy")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ------------------------------------------------------------------
    def derived(self) -> abi.Derived:
        d = abi.Derived()
        abi.check(self.lib.mfixdem_get_derived(self.h, C.byref(d)), "get_derived")
        return d

    def set_particles(self, parts: Particles):
        n = parts.n
        abi.check(self.lib.mfixdem_set_particles(
            self.h, n, abi.dptr(parts.pos), abi.dptr(parts.vel), abi.dptr(parts.omega),
            abi.dptr(parts.radius), abi.dptr(parts.ro_sol), abi.iptr(parts.phase), n),
            "set_particles")

    def set_particle_state(self, state):
        """PARTICLE_STATE of the first len(state) slots (0: hole), before init()."""
        st = np.ascontiguousarray(state, dtype=np.int8)
        abi.check(self.lib.mfixdem_set_particle_state(self.h, abi.bptr(st), st.size),
                  "set_particle_state")

    def set_facets(self, vertex, normal, default_walls=True):
        """STL walls before init(): vertex[f, axis, v] (VERTEX(v, axis, f)),
        normal[f, axis]; the box walls follow as facets nf..nf+5."""
        v = np.ascontiguousarray(vertex, dtype=np.float64).reshape(-1)
        nrm = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1)
        abi.check(self.lib.mfixdem_set_facets(self.h, nrm.size // 3, abi.dptr(v), abi.dptr(nrm),
                                              int(bool(default_walls))), "set_facets")

    def set_wall_motion(self, first, count, vel, omega=(0.0, 0.0, 0.0),
                        rot_center=(0.0, 0.0, 0.0)):
        """Facets [first, first+count) move at vel + omega x (centre - rot_center)."""
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (vel, omega, rot_center)]
        abi.check(self.lib.mfixdem_set_wall_motion(self.h, int(first), int(count),
                                                   *[abi.dptr(a) for a in arr]),
                  "set_wall_motion")

    def set_dem_bcs(self, bcs, seed=1):
        """DEM mass inlets and outlets (abi.dem_bc) before init()."""
        arr = (abi.DemBc * max(len(bcs), 1))(*bcs)
        abi.check(self.lib.mfixdem_set_dem_bcs(self.h, len(bcs), arr, seed), "set_dem_bcs")

    def set_gas(self, gas: Gas):
        abi.check(self.lib.mfixdem_set_gas(
            self.h, abi.dptr(gas.u_g), abi.dptr(gas.v_g), abi.dptr(gas.w_g),
            abi.dptr(gas.p_force)), "set_gas")

    def set_tuning(self, cell_scale=1.0, row_kernel=abi.ROWS_THREAD):
        abi.check(self.lib.mfixdem_set_tuning(self.h, float(cell_scale), int(row_kernel)),
                  "set_tuning")

    def device_bytes(self) -> int:
        b = C.c_int64(0)
        abi.check(self.lib.mfixdem_get_device_bytes(self.h, C.byref(b)), "get_device_bytes")
        return int(b.value)

    def init(self):
        abi.check(self.lib.mfixdem_init(self.h), "init")

    def step(self, nn: int):
        abi.check(self.lib.mfixdem_step(self.h, nn), f"step {nn}")

    def run(self, nn_first: int, nsteps: int):
        abi.check(self.lib.mfixdem_run(self.h, nn_first, nsteps), "run")

    def synchronize(self):
        abi.check(self.lib.mfixdem_synchronize(self.h), "synchronize")

    def set_debug(self, flags: int):
        abi.check(self.lib.mfixdem_set_debug(self.h, flags), "set_debug")

    # reference routines, unfused ---------------------------------------------
    def calc_force_dem(self):
        abi.check(self.lib.mfixdem_calc_force_dem(self.h), "calc_force_dem")

    def calc_drag_des(self):
        abi.check(self.lib.mfixdem_calc_drag_des(self.h), "calc_drag_des")

    def cfnewvalues(self):
        abi.check(self.lib.mfixdem_cfnewvalues(self.h), "cfnewvalues")

    def desgrid_pic(self):
        abi.check(self.lib.mfixdem_desgrid_pic(self.h), "desgrid_pic")

    def neighbour(self):
        abi.check(self.lib.mfixdem_neighbour(self.h), "neighbour")

    def particles_in_cell(self):
        abi.check(self.lib.mfixdem_particles_in_cell(self.h), "particles_in_cell")

    def set_gas_pressure(self, p_g):
        abi.check(self.lib.mfixdem_set_gas_pressure(self.h, abi.dptr(np.ascontiguousarray(p_g))),
                  "set_gas_pressure")

    def get_p_force(self):
        out = np.zeros(3 * self.ncell)
        abi.check(self.lib.mfixdem_get_p_force(self.h, abi.dptr(out)), "get_p_force")
        return out.reshape(self.ncell, 3)

    def drag_gs_gas(self):
        """DRAG_GS_GAS0 (drag_gs_des0.f:235-497) on the current particles:
        returns (F_GDS[ncell], DRAG_AM[ncell], DRAG_BM[3, ncell])."""
        abi.check(self.lib.mfixdem_drag_gs_gas(self.h), "drag_gs_gas")
        f = np.zeros(self.ncell)
        am = np.zeros(self.ncell)
        bm = np.zeros(3 * self.ncell)
        abi.check(self.lib.mfixdem_get_gas_drag(self.h, abi.dptr(f), abi.dptr(am), abi.dptr(bm)),
                  "get_gas_drag")
        return f, am, bm.reshape(3, self.ncell)

    def des_time_init(self):
        """DES_TIME_INIT's coupled part (explicit coupling: DRAG_FC)."""
        abi.check(self.lib.mfixdem_des_time_init(self.h), "des_time_init")

    def calc_interp_weights(self):
        abi.check(self.lib.mfixdem_calc_interp_weights(self.h), "calc_interp_weights")

    def comp_mean_fields(self):
        abi.check(self.lib.mfixdem_comp_mean_fields(self.h), "comp_mean_fields")

    # results -------------------------------------------------------------------
    def get_particles(self):
        n = self.n
        pos, vel, om = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        st = np.zeros(n, dtype=np.int8)
        abi.check(self.lib.mfixdem_get_particles(
            self.h, abi.dptr(pos), abi.dptr(vel), abi.dptr(om), abi.bptr(st), n), "get_particles")
        return pos, vel, om, st

    def get_forces(self):
        n = self.n
        fc, tow = np.zeros((3, n)), np.zeros((3, n))
        abi.check(self.lib.mfixdem_get_forces(self.h, abi.dptr(fc), abi.dptr(tow), n), "get_forces")
        return fc, tow

    def get_cells(self):
        n = self.n
        dg = np.zeros(n, dtype=np.int32)
        pijk = np.zeros((4, n), dtype=np.int32)
        abi.check(self.lib.mfixdem_get_cells(self.h, abi.iptr(dg), abi.iptr(pijk), n), "get_cells")
        return dg, pijk

    def get_pairs(self):
        npairs = C.c_int64(0)
        abi.check(self.lib.mfixdem_get_pairs(self.h, abi.iptr(None), abi.iptr(None),
                                             abi.dptr(None), 0, C.byref(npairs)), "get_pairs")
        m = max(int(npairs.value), 1)
        idx = np.zeros(self.n, dtype=np.int32)
        nb = np.zeros(m, dtype=np.int32)
        pft = np.zeros(3 * m)
        abi.check(self.lib.mfixdem_get_pairs(self.h, abi.iptr(idx), abi.iptr(nb), abi.dptr(pft),
                                             m, C.byref(npairs)), "get_pairs")
        k = int(npairs.value)
        return idx, nb[:k], pft[:3 * k].reshape(k, 3)

    def set_pairs(self, neighbor_index, neighbors, pft):
        """Restart of the pair history (get_pairs layout) before init()."""
        idx = np.ascontiguousarray(neighbor_index, dtype=np.int32)
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        pf = np.ascontiguousarray(np.asarray(pft, dtype=np.float64).reshape(-1))
        abi.check(self.lib.mfixdem_set_pairs(self.h, abi.iptr(idx), abi.iptr(nb), abi.dptr(pf),
                                             nb.size), "set_pairs")

    def get_gas(self):
        ep = np.zeros(self.ncell)
        rops = np.zeros(self.ncell * self.params.nphases)
        abi.check(self.lib.mfixdem_get_gas(self.h, abi.dptr(ep), abi.dptr(rops)), "get_gas")
        return ep, rops.reshape(self.params.nphases, self.ncell)

    def status(self) -> abi.Status:
        s = abi.Status()
        abi.check(self.lib.mfixdem_get_status(self.h, C.byref(s)), "get_status")
        return s


def pair_set(neighbor_index, neighbors):
    """Reference CSR half-list -> set of (i, j) 1-based id pairs."""
    out = set()
    start = 1
    for i, end in enumerate(neighbor_index, start=1):
        for c in range(start, end):
            out.add((i, int(neighbors[c - 1])))
        start = end
    return out


def pair_table(neighbor_index, neighbors, pft):
    """CSR half-list -> {(i, j): pft row} (1-based ids)."""
    out = {}
    start = 1
    for i, end in enumerate(neighbor_index, start=1):
        for c in range(start, end):
            out[(i, int(neighbors[c - 1]))] = pft[c - 1]
        start = end
    return out

"""Python driver of an ensemble handle: many independent DEM systems stepped
by the same launches (mfixdem_ensemble_* in include/mfixdem.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .workloads import Particles


class Ensemble:
    def __init__(self, params):
        self.lib = abi.load_library()
        self.params = list(params)
        self.nsys = len(self.params)
        arr = (abi.Params * self.nsys)(*self.params)
        self.h = C.c_void_p()
        abi.check(self.lib.mfixdem_ensemble_create(self.nsys, arr, C.byref(self.h)),
                  "mfixdem_ensemble_create")
        self.cap = [p.max_pip for p in self.params]

    def close(self):
        if self.h:
            abi.check(self.lib.mfixdem_destroy(self.h), "mfixdem_destroy")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, s: int, parts: Particles):
        n = parts.n
        abi.check(self.lib.mfixdem_ensemble_set_particles(
            self.h, s, n, abi.dptr(parts.pos), abi.dptr(parts.vel), abi.dptr(parts.omega),
            abi.dptr(parts.radius), abi.dptr(parts.ro_sol), abi.iptr(parts.phase), n),
            "ensemble_set_particles")

    def set_tuning(self, cell_scale=1.0, row_kernel=abi.ROWS_THREAD):
        abi.check(self.lib.mfixdem_set_tuning(self.h, float(cell_scale), int(row_kernel)),
                  "set_tuning")

    def device_bytes(self) -> int:
        b = C.c_int64(0)
        abi.check(self.lib.mfixdem_get_device_bytes(self.h, C.byref(b)), "get_device_bytes")
        return int(b.value)

    def init(self):
        abi.check(self.lib.mfixdem_init(self.h), "init")

    def step(self, nn):
        abi.check(self.lib.mfixdem_step(self.h, nn), f"step {nn}")

    def run(self, nn_first, nsteps):
        abi.check(self.lib.mfixdem_run(self.h, nn_first, nsteps), "run")

    def synchronize(self):
        abi.check(self.lib.mfixdem_synchronize(self.h), "synchronize")

    def get_particles(self, s: int):
        n = self.cap[s]
        pos, vel, om = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        st = np.zeros(n, dtype=np.int8)
        abi.check(self.lib.mfixdem_ensemble_get_particles(
            self.h, s, abi.dptr(pos), abi.dptr(vel), abi.dptr(om), abi.bptr(st), n),
            "ensemble_get_particles")
        return pos, vel, om, st

    def status(self):
        st = abi.Status()
        abi.check(self.lib.mfixdem_get_status(self.h, C.byref(st)), "status")
        return st

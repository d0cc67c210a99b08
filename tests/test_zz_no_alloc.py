"""mfixdem_drag_gs_gas_noalloc against mfixdem_drag_gs_gas through the C ABI.

The no-allocation variant writes F_GDS, DRAG_AM and DRAG_BM into the
caller's device buffers and keeps its per-particle records and
DES_VOL_NODE in the library's fixed per-GPU scratch.  GARG_2012 results
(node gathers in a fixed order) must equal the allocating entry point's
bit for bit; the filter scheme adds with atomics, so its sums agree to
1e-13 of the largest value.  Two calls in a row, the second on a smaller
system (fewer particles and cells), show that the previous call's scratch
contents do not leak into the next; a handle too large for the scratch is
refused with MFIXDEM_ERR_SCRATCH before any launch."""
import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi, workloads
from mfixdem.system import DemSystem

pytestmark = pytest.mark.gpu


def _bed(n, cells, lx, ly, lz, seed, scheme=abi.INTERP_GARG_2012):
    p, parts, gas = workloads.fluidised_bed(n=n, seed=seed, lx=lx, ly=ly, lz=lz, cells=cells,
                                            dg=(0, 0, 0))
    p, parts = workloads.jostled(p, parts, seed=seed + 1, v0=0.2)
    p.interp_scheme = scheme
    if scheme != abi.INTERP_GARG_2012:
        p.des_interp_on = 1
    return p, parts, gas


def _hip():
    # the HIP runtime the library itself is linked against (device buffers
    # for the caller's side of the call; the library allocates nothing here)
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc(g):
    import ctypes as C
    hip = _hip()
    nc = g.ncell
    bufs = []
    for k in (1, 1, 3):
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), 8 * k * nc) == 0
        assert hip.hipMemset(d, 0xFF, 8 * k * nc) == 0   # NaN in every double
        bufs.append(d)
    assert hip.hipDeviceSynchronize() == 0
    rc = g.lib.mfixdem_drag_gs_gas_noalloc(g.h, *bufs)
    abi.check(rc, "drag_gs_gas_noalloc")
    g.synchronize()
    out = []
    for k, d in zip((1, 1, 3), bufs):
        a = np.zeros(k * nc)
        assert hip.hipMemcpy(a.ctypes.data, d, 8 * k * nc, 2) == 0   # device -> host
        assert hip.hipFree(d) == 0
        out.append(a)
    return out[0], out[1], out[2].reshape(3, nc)


@pytest.mark.parametrize("scheme", [abi.INTERP_GARG_2012, abi.INTERP_LINEAR_HAT],
                         ids=["garg_2012", "linear_hat"])
def test_noalloc_gas_drag_equals_allocating_call(scheme):
    cases = [(_bed(4000, (10, 20, 10), 0.02, 0.04, 0.02, 2, scheme), 12),
             (_bed(1500, (6, 12, 6), 0.012, 0.024, 0.012, 5, scheme), 7)]
    for (p, parts, gas), steps in cases:
        g = DemSystem(p)
        g.set_particles(parts)
        g.set_gas(gas)
        g.init()
        g.run(1, steps)
        f1, am1, bm1 = _noalloc(g)             # before the allocating call
        f0, am0, bm0 = g.drag_gs_gas()
        f2, am2, bm2 = _noalloc(g)             # and after it
        g.close()
        assert f0.max() > 0.0 and np.abs(bm0).max() > 0.0
        for f, am, bm in ((f1, am1, bm1), (f2, am2, bm2)):
            assert not np.isnan(f).any() and not np.isnan(am).any() and not np.isnan(bm).any()
            if scheme == abi.INTERP_GARG_2012:
                assert np.array_equal(f, f0)
                assert np.array_equal(am, am0)
                assert np.array_equal(bm, bm0)
            else:   # atomics in no fixed order
                assert not am.any()
                assert np.max(np.abs(f - f0)) <= 1e-13 * f0.max()
                assert np.max(np.abs(bm - bm0)) <= 1e-13 * np.abs(bm0).max()


def test_noalloc_refuses_a_handle_beyond_its_scratch():
    import ctypes as C
    hip = _hip()
    p, _, _ = workloads.fluidised_bed(n=2000, lx=0.128, ly=0.256, lz=0.128,
                                      cells=(64, 128, 64), dg=(26, 52, 26))
    p.max_pip = 1_300_000   # 8 * 1.3M + 66*130*66 doubles > 10 * 2^20
    g = DemSystem(p)
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), 8 * 3 * g.ncell) == 0
    rc = g.lib.mfixdem_drag_gs_gas_noalloc(g.h, d, d, d)
    assert rc == abi.ERRORS_BY_NAME["SCRATCH"]
    assert g.lib.mfixdem_drag_gs_gas_noalloc(g.h, None, None, None) == abi.ERRORS_BY_NAME["ARG"]
    assert hip.hipFree(d) == 0
    g.close()

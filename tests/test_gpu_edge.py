"""Edge cases on the GPU, through the C ABI, against the oracle's outcome
on the same inputs (tests/test_edge_cpu.py pins the oracle's): empty and
partly filled systems, a coincident pair and an overfilled fluid cell
(fatal, returned at the next call that waits for the device), rogue
particles, a particle leaving through a wall, and a caller buffer that is
too small for the pair list."""
import ctypes as C

import numpy as np
import pytest

import _paths  # noqa: F401
import edge_cases as E
from mfixdem import abi
from mfixdem.system import DemSystem
from oracle_py import OracleSystem

pytestmark = pytest.mark.gpu


def both(p, parts, gas=None):
    out = []
    for cls in (DemSystem, OracleSystem):
        s = cls(p)
        s.set_particles(parts)
        if gas is not None:
            s.set_gas(gas)
        out.append(s)
    return out


def test_empty_system_on_gpu():
    p, parts = E.empty()
    g = DemSystem(p)
    g.set_particles(parts)
    g.init()
    g.run(1, 60)
    g.synchronize()
    st = g.status()
    assert st.n_particles == 0 and st.steps == 60 and st.n_pairs == 0
    assert g.get_particles()[3].max() == 0
    g.close()


def test_partly_filled_capacity_bit_exact():
    g, o = both(*E.partly_filled())
    for s in (g, o):
        s.init()
        s.run(1, 60)
    for a, b in zip(g.get_particles(), o.get_particles()):
        assert np.array_equal(a, b)
    g.close()
    o.close()


def test_coincident_pair_reported():
    g, o = both(*E.coincident())
    g.init()
    with pytest.raises(abi.MfixDemError) as e:  # step 1 rebuilds: it waits for the device
        g.run(1, 1)
        g.synchronize()
    assert e.value.code == abi.ERRORS_BY_NAME["ZERO_DISTANCE"]
    assert g.status().zero_distance == 1
    o.init()
    with pytest.raises(abi.MfixDemError) as eo:
        o.step(1)
    assert eo.value.code == e.value.code
    g.close()
    o.close()


def test_overfilled_cell_reported_at_init():
    p, parts = E.overfilled_cell()
    g, o = both(p, parts, E.uniform_gas(p))
    codes = []
    for s in (g, o):
        with pytest.raises(abi.MfixDemError) as e:
            s.init()
        codes.append(e.value.code)
    assert codes == [abi.ERRORS_BY_NAME["EPG"]] * 2
    assert g.status().epg_errors >= 1
    g.close()
    o.close()


@pytest.mark.parametrize("remove", [1, 0])
def test_rogue_particle_on_gpu(remove):
    g, o = both(*E.rogue(remove))
    if remove:
        for s in (g, o):
            s.init()
            s.run(1, 60)
        g.synchronize()
        assert g.get_particles()[3].tolist() == o.get_particles()[3].tolist() == [1, 1, 0]
        assert g.status().rogue_deleted == 1
        for a, b in zip(g.get_particles()[:3], o.get_particles()[:3]):
            assert np.array_equal(a[:, :2], b[:, :2])
    else:
        for s in (g, o):
            with pytest.raises(abi.MfixDemError) as e:
                s.init()
            assert e.value.code == abi.ERRORS_BY_NAME["ROGUE"]
    g.close()
    o.close()


def test_particle_leaving_through_a_wall_on_gpu():
    p, parts = E.leaving()
    g, o = both(p, parts, E.uniform_gas(p))
    for s in (g, o):
        s.init()
        s.run(1, 60)
    g.synchronize()
    assert g.get_particles()[3].tolist() == o.get_particles()[3].tolist() == [1, 0]
    pg, vg = g.get_particles()[:2]
    po, vo = o.get_particles()[:2]
    assert np.allclose(pg[:, 0], po[:, 0], rtol=1e-12, atol=0)
    assert np.allclose(vg[:, 0], vo[:, 0], rtol=1e-12, atol=1e-15)
    g.close()
    o.close()


def test_pair_buffer_too_small():
    p, parts = E.partly_filled()
    g = DemSystem(p)
    g.set_particles(parts)
    g.init()
    n = C.c_int64(0)
    assert g.lib.mfixdem_get_pairs(g.h, abi.iptr(None), abi.iptr(None), abi.dptr(None), 0,
                                   C.byref(n)) == 0
    assert n.value > 10
    m = n.value - 1
    idx = np.zeros(p.max_pip, dtype=np.int32)
    nb = np.zeros(m, dtype=np.int32)
    pft = np.zeros(3 * m)
    rc = g.lib.mfixdem_get_pairs(g.h, abi.iptr(idx), abi.iptr(nb), abi.dptr(pft), m, C.byref(n))
    assert rc == abi.ERRORS_BY_NAME["CAPACITY"] and n.value == m + 1
    g.close()

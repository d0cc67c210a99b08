"""Edge cases on the CPU: parameter validation of the library (checked on
the host before any device call), and the oracle's
behaviour on the edge-case workloads that the GPU tests repeat: empty and
partly filled systems, a coincident pair (calc_force_dem.f:267-271), an
overfilled fluid cell (calc_epg_des.f:146-203), rogue particles
(desgrid_mod.f:1074-1104) and a particle that leaves through a wall."""
import ctypes as C

import numpy as np
import pytest

import _paths  # noqa: F401
import edge_cases as E
from mfixdem import abi, workloads
from oracle_py import OracleSystem


def create_status(p):
    lib = abi.load_library()
    h = C.c_void_p()
    rc = lib.mfixdem_create(C.byref(p), C.byref(h))
    assert rc != 0 and not h.value   # these cases never reach the device
    return rc


def base():
    p, _, _ = workloads.settling_bed(n=100, seed=1, box=(0.01, 0.02, 0.01))
    return p


def coupled_base():
    return E.small_box(100, coupled=True)


@pytest.mark.parametrize("edit,code", [
    (lambda p: setattr(p, "max_pip", 0), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "nphases", 0), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "neighbor_search_n", 0), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "coll_model", 7), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "des_neighbor_search", 2), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "max_radius", 0.0), abi.ERRORS_BY_NAME["ARG"]),
], ids=["no-capacity", "no-phases", "search-interval", "collision-model", "search-kind",
        "radius"])
def test_invalid_parameters_rejected(edit, code):
    p = base()
    edit(p)
    assert create_status(p) == code


@pytest.mark.parametrize("edit,code", [
    (lambda p: setattr(p, "drag_type", 9), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "model_b", 1), abi.ERRORS_BY_NAME["UNSUPPORTED"]),
    (lambda p: (setattr(p, "interp_scheme", abi.INTERP_NONE)), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "interp_scheme", abi.INTERP_SQUARE_DPVM), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "des_interp_width", 1.0e-3), abi.ERRORS_BY_NAME["ARG"]),
    (lambda p: setattr(p, "imax", 1100), abi.ERRORS_BY_NAME["UNSUPPORTED"]),
], ids=["drag-law", "model-b", "none-with-interp-flags", "dpvm-without-width",
        "garg-with-width", "grid-too-large"])
def test_invalid_coupling_rejected(edit, code):
    p = coupled_base()
    edit(p)
    assert create_status(p) == code


def run_oracle(p, parts, gas=None, steps=60):
    o = OracleSystem(p)
    o.set_particles(parts)
    if gas is not None:
        o.set_gas(gas)
    err = 0
    try:
        o.init()
        o.run(1, steps)
    except abi.MfixDemError as e:
        err = e.code
    return o, err


def test_empty_system_steps():
    o, err = run_oracle(*E.empty())
    assert err == 0 and o.status().n_particles == 0 and o.status().steps == 60


def test_partly_filled_capacity():
    p, parts = E.partly_filled()
    o, err = run_oracle(p, parts)
    _, _, _, st = o.get_particles()
    assert err == 0 and st[:2000].min() == 1 and st[2000:].max() == 0


def test_coincident_pair_is_fatal():
    o, err = run_oracle(*E.coincident())
    assert err == abi.ERRORS_BY_NAME["ZERO_DISTANCE"] and o.status().zero_distance == 1


def test_overfilled_cell_is_fatal():
    p, parts = E.overfilled_cell()
    o, err = run_oracle(p, parts, E.uniform_gas(p))
    assert err == abi.ERRORS_BY_NAME["EPG"] and o.status().epg_errors >= 1


@pytest.mark.parametrize("remove", [1, 0])
def test_rogue_particle(remove):
    o, err = run_oracle(*E.rogue(remove))
    st = o.get_particles()[3]
    if remove:
        assert err == 0 and st.tolist() == [1, 1, 0] and o.status().rogue_deleted == 1
    else:
        assert err == abi.ERRORS_BY_NAME["ROGUE"]


def test_particle_leaving_through_a_wall():
    p, parts = E.leaving()
    o, err = run_oracle(p, parts, E.uniform_gas(p))
    assert err == 0 and o.get_particles()[3].tolist() == [1, 0]

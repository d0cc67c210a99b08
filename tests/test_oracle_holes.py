"""PARTICLE_STATE holes in the oracle (discretelement_mod.f:68-78; every
routine skips a NONEXISTENT slot, e.g. calc_force_dem.f:185,211), pinned
against the same run without the holes.

Removing the holes and renumbering the remaining particles in their order
keeps every ordering the reference's sums depend on: the DES-cell lists
are in ascending id, a pair belongs to its lower id, and a row lists its
partners in ascending id within each stencil cell.  So a run with holes
must equal, bit for bit, the run of the compacted particles: positions,
velocities, spins, and the pair lists with their tangential history once
the ids are mapped.  A routine that forgot to skip a hole (a force from a
stale slot, a hole binned or listed) breaks this at once."""
import numpy as np

import _paths  # noqa: F401
from mfixdem import workloads
from mfixdem.system import pair_table
from oracle_py import OracleSystem


def test_holes_equal_the_compacted_run():
    p, parts, _ = workloads.settling_bed(n=2000, seed=3, box=(0.01, 0.035, 0.01))
    p, parts = workloads.jostled(p, parts, seed=4)
    st = np.ones(parts.n, dtype=np.int8)
    st[np.random.default_rng(9).random(parts.n) < 0.2] = 0
    keep = st == 1
    ids = np.nonzero(keep)[0]
    holes = OracleSystem(p)
    holes.set_particles(parts)
    holes.set_particle_state(st)
    holes.init()
    pc = workloads.make_particles(parts.pos[:, keep], parts.radius[keep], parts.ro_sol[keep],
                                  parts.vel[:, keep], parts.omega[:, keep])
    pk = type(p).from_buffer_copy(p)
    pk.max_pip = int(keep.sum())
    comp = OracleSystem(pk)
    comp.set_particles(pc)
    comp.init()
    for nn in range(1, 61):          # rebuilds at 1, 25 and 50
        holes.step(nn)
        comp.step(nn)
    a, b = holes.get_particles(), comp.get_particles()
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x[:, keep], y)
        assert not np.any(x[:, ~keep])                 # holes stay zero
    assert np.array_equal(a[3], st)
    th = pair_table(*holes.get_pairs())
    tc = pair_table(*comp.get_pairs())
    mapped = {(int(ids[i - 1]) + 1, int(ids[j - 1]) + 1): v for (i, j), v in tc.items()}
    assert set(th) == set(mapped)
    assert len(th) > 1000
    assert all(np.array_equal(th[k], mapped[k]) for k in th)
    holes.close()
    comp.close()

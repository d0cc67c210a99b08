"""Seeded synthetic inputs for the DEM hot path.

Every workload is a (Params, particles, gas) triple built from a seed, so the
HIP library and the CPU oracle can be driven with identical inputs.  Axis
convention follows the reference: gravity along -Y (constant_mod.f:235), so
the long/vertical box axis is Y.

  settling_bed       20,000 monodisperse spheres, d = 1 mm, rho = 2500,
                     0.02 x 0.08 x 0.02 m box, LSD kn = 1e3, e = 0.9,
                     mu = 0.3, dt = 1e-5 s, no gas.
  fluidised_bed      polydisperse spheres (log-normal d, median 1 mm,
                     sigma 0.25, clipped to [0.5, 2] mm), Hertz E = 1e7,
                     nu = 0.3, e = 0.9, mu = 0.3, gas grid 64 x 128 x 64 on
                     0.128 x 0.256 x 0.128 m, frozen upflow 1.2 m/s with a
                     5 % sinusoidal perturbation, Gidaspow drag, GARG_2012.
  dem01              the reference test DEM01 (single particle bouncing on
                     the south wall, tests/dem/DEM01/mfix.dat).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import abi


@dataclasses.dataclass
class Particles:
    pos: np.ndarray      # (3, n) float64, reference SoA (component-major)
    vel: np.ndarray      # (3, n)
    omega: np.ndarray    # (3, n)
    radius: np.ndarray   # (n,)
    ro_sol: np.ndarray   # (n,)
    phase: np.ndarray    # (n,) int32, 1-based

    @property
    def n(self) -> int:
        return int(self.radius.shape[0])


@dataclasses.dataclass
class Gas:
    u_g: np.ndarray
    v_g: np.ndarray
    w_g: np.ndarray
    p_force: np.ndarray  # (ncell, 3) -> P_FORCE(3, IJK)


def n_cells(p: abi.Params) -> int:
    return (p.imax + 2) * (p.jmax + 2) * (p.kmax + 2)


def make_particles(pos, radius, rho, vel=None, omega=None, phase=None) -> Particles:
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[1]
    vel = np.zeros((3, n)) if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
    omega = np.zeros((3, n)) if omega is None else np.ascontiguousarray(omega, dtype=np.float64)
    radius = np.ascontiguousarray(np.broadcast_to(radius, (n,)), dtype=np.float64)
    rho = np.ascontiguousarray(np.broadcast_to(rho, (n,)), dtype=np.float64)
    phase = (np.ones(n, dtype=np.int32) if phase is None
             else np.ascontiguousarray(phase, dtype=np.int32))
    return Particles(pos, vel, omega, radius, rho, phase)


def _rsa_lib():
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "lib", "libmfixdem_workload.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    fn = lib.mfixdem_rsa_place
    fn.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                   C.c_int64, C.c_void_p]
    fn.restype = C.c_int64
    return fn


def rsa_positions(seed, radii, lo, hi, max_tries=1_000_000):
    """Random sequential addition of spheres (in the order given) inside the
    box [lo, hi]; returns (3, n) centres.  Deterministic for a given seed."""
    radii = np.ascontiguousarray(radii, dtype=np.float64)
    n = radii.shape[0]
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    pos = np.zeros((3, n))
    got = _rsa_lib()(n, radii.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                     int(seed), int(max_tries), pos.ctypes.data)
    if got != n:
        raise RuntimeError(f"RSA placed {got} of {n} spheres")
    return pos


def settling_bed(n=20_000, seed=1, box=(0.02, 0.08, 0.02)):
    """Plain FP64 DEM settling bed (the reference's CPU-runnable case):
    n spheres of d = 1 mm placed by RSA in the box (X, Y-up, Z)."""
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = 1
    p.max_pip = n
    p.xlength, p.ylength, p.zlength = box
    p.imax = p.jmax = p.kmax = 1
    p.dtsolid = 1.0e-5
    p.grav[:] = [0.0, -9.81, 0.0]
    p.kn = 1.0e3
    p.kn_w = 1.0e3
    p.mew = 0.3
    p.mew_w = 0.3
    p.des_en_input[0] = 0.9
    p.des_en_wall_input[0] = 0.9
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 2500.0
    p.max_radius = 0.5e-3
    r = np.full(n, 0.5e-3)
    pos = rsa_positions(seed, r, [0, 0, 0], list(box))
    parts = make_particles(pos, r, 2500.0)
    return p, parts, None


def lognormal_diameters(rng, n, median=1.0e-3, sigma=0.25, dmin=0.5e-3, dmax=2.0e-3):
    d = median * np.exp(sigma * rng.standard_normal(n))
    return np.clip(d, dmin, dmax)


def frozen_gas(p: abi.Params, seed=7, u0=1.2, amp=0.05):
    """Uniform upflow V_G = u0 with a seeded 5 % sinusoidal perturbation on
    each component; P_FORCE = -dp/dy of a hydrostatic gas column."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    # cell index arrays in MFiX order: J fastest, then I, then K
    k, i, j = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1),
                          np.arange(1, nj + 1), indexing="ij")
    x = (i - 1.5) / p.imax
    y = (j - 1.5) / p.jmax
    z = (k - 1.5) / p.kmax
    ph = rng.random(9) * 2 * np.pi
    kx, ky, kz = 2 * np.pi * (1 + rng.integers(0, 3, 3))
    def wave(a, b, c):
        return np.sin(kx * x + ph[a]) * np.sin(ky * y + ph[b]) * np.sin(kz * z + ph[c])
    u = amp * u0 * wave(0, 1, 2)
    v = u0 * (1.0 + amp * wave(3, 4, 5))
    w = amp * u0 * wave(6, 7, 8)
    pf = np.zeros((ni * nj * nk, 3))
    pf[:, 1] = p.ro_g0 * 9.81
    return Gas(np.ascontiguousarray(u.ravel()), np.ascontiguousarray(v.ravel()),
               np.ascontiguousarray(w.ravel()), np.ascontiguousarray(pf))


def fluidised_bed(n=1_000_000, seed=2, phi=0.30, precision=abi.FP64,
                  lx=0.128, ly=0.256, lz=0.128, cells=(64, 128, 64),
                  dg=(26, 52, 26)):
    """FP64 CFD-DEM with a frozen gas field (1M polydisperse Hertz spheres at
    full size).  Smaller n keeps the same box and shrinks the bed height."""
    rng = np.random.default_rng(seed)
    d = lognormal_diameters(rng, n)
    order = np.argsort(-d, kind="stable")   # RSA places large spheres first
    d = d[order]
    r = 0.5 * d
    vol = (np.pi / 6.0) * np.sum(d ** 3)
    h = min(ly * 0.9, vol / (phi * lx * lz))
    pos = rsa_positions(seed, r, [0, 0, 0], [lx, h, lz])
    perm = rng.permutation(n)               # global ids in random spatial order
    pos = pos[:, perm]
    r = r[perm]
    p = abi.default_params()
    p.coll_model = abi.HERTZ
    p.integrator = abi.EULER
    p.precision = precision
    p.nphases = 1
    p.max_pip = n
    p.coupled = 1
    p.drag_type = abi.DRAG_GIDASPOW
    p.interp_scheme = abi.INTERP_GARG_2012
    p.xlength, p.ylength, p.zlength = lx, ly, lz
    p.imax, p.jmax, p.kmax = cells
    p.desgridsearch_imax, p.desgridsearch_jmax, p.desgridsearch_kmax = dg
    p.dtsolid = 1.0e-5
    p.grav[:] = [0.0, -9.81, 0.0]
    p.mew = 0.3
    p.mew_w = 0.3
    p.des_en_input[0] = 0.9
    p.des_et_input[0] = 0.9
    p.des_en_wall_input[0] = 0.9
    p.des_et_wall_input[0] = 0.9
    p.e_young[0] = 1.0e7
    p.v_poisson[0] = 0.3
    p.ew_young = 1.0e7
    p.vw_poisson = 0.3
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 2500.0
    p.max_radius = float(r.max())
    p.ro_g0 = 1.2
    p.mu_g0 = 1.8e-5
    parts = make_particles(pos, r, 2500.0)
    gas = frozen_gas(p)
    return p, parts, gas


def jostled(p: abi.Params, parts: Particles, seed=5, shrink=0.98, v0=0.3, w0=200.0):
    """Test variant of a packing: the spheres are grown by 1/shrink after
    placement (small initial overlaps with neighbours and walls) and given
    seeded random translational and angular velocities, so contacts,
    friction and wall history are exercised from the first step."""
    rng = np.random.default_rng(seed)
    n = parts.n
    radius = parts.radius / shrink
    vel = v0 * rng.uniform(-1.0, 1.0, (3, n))
    om = w0 * rng.uniform(-1.0, 1.0, (3, n))
    p.max_radius = float(radius.max())
    return p, make_particles(parts.pos, radius, parts.ro_sol, vel, om, parts.phase)


def dem01(kn=1.0e4, en=0.9, integrator=abi.EULER):
    """tests/dem/DEM01: one particle (d = 0.2 m, rho = 2600) released at
    y = 0.5 in a 1 m box with default walls; KN = KN_W, EN = EN_WALL."""
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = integrator
    p.nphases = 1
    p.max_pip = 1
    p.xlength = p.ylength = p.zlength = 1.0
    p.imax = p.jmax = p.kmax = 1
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = 1
    p.neighbor_search_n = 99
    p.kn = kn
    p.kn_w = kn
    p.mew = 0.0
    p.mew_w = 0.0
    p.des_en_input[0] = en
    p.des_en_wall_input[0] = en
    p.d_p0[0] = 0.20
    p.ro_s0[0] = 2600.0
    p.max_radius = 0.1
    parts = make_particles(np.array([[0.6], [0.5], [0.6]]), 0.1, 2600.0)
    return p, parts, None


def dem03(en=1.0, integrator=abi.ADAMS_BASHFORTH):
    """tests/dem/DEM03: two stacked, compressed spheres (d = 1 mm; rho 2e4
    and 1e4, two phases) between the south and north walls of a 1 x 0.0018 x
    1 m box; KN = KN_W = 1e3, KT_FAC = KT_W_FAC = 1, DES_ETAT_FAC = 1,
    EN = EN_WALL for every pair, no friction."""
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = integrator
    p.nphases = 2
    p.max_pip = 2
    p.xlength, p.ylength, p.zlength = 1.0, 0.0018, 1.0
    p.imax = p.jmax = p.kmax = 1
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = 1
    p.neighbor_search_n = 99
    p.kn = p.kn_w = 1.0e3
    p.kt_fac = p.kt_w_fac = 1.0
    p.des_etat_fac = p.des_etat_w_fac = 1.0
    p.mew = p.mew_w = 0.0
    for i in range(3):
        p.des_en_input[i] = en
    for i in range(2):
        p.des_en_wall_input[i] = en
    p.d_p0[0] = p.d_p0[1] = 1.0e-3
    p.ro_s0[0], p.ro_s0[1] = 2.0e4, 1.0e4
    p.max_radius = 5.0e-4
    pos = np.array([[0.5, 0.5], [0.00045, 0.00135], [0.5, 0.5]])
    parts = make_particles(pos, 5.0e-4, np.array([2.0e4, 1.0e4]),
                           phase=np.array([1, 2], dtype=np.int32))
    return p, parts, None


def dem04(mew=0.2, integrator=abi.ADAMS_BASHFORTH):
    """tests/dem/DEM04: a sphere (d = 1 mm, rho = 1e4) resting on the south
    wall of a 1 m box and launched along +x at 0.1 m/s; KN = KN_W = 1e4,
    KT_FAC = KT_W_FAC = 1, EN = 1, DES_ETAT_FAC = 1, MEW = MEW_W = mew."""
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = integrator
    p.nphases = 1
    p.max_pip = 1
    p.xlength = p.ylength = p.zlength = 1.0
    p.imax = p.jmax = p.kmax = 1
    p.desgridsearch_imax = p.desgridsearch_jmax = p.desgridsearch_kmax = 1
    p.neighbor_search_n = 99
    p.kn = p.kn_w = 1.0e4
    p.kt_fac = p.kt_w_fac = 1.0
    p.des_etat_fac = p.des_etat_w_fac = 1.0
    p.mew = p.mew_w = mew
    p.des_en_input[0] = 1.0
    p.des_en_wall_input[0] = 1.0
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 1.0e4
    p.max_radius = 5.0e-4
    parts = make_particles(np.array([[0.001], [0.0005], [0.5]]), 5.0e-4, 1.0e4,
                           vel=np.array([[0.1], [0.0], [0.0]]))
    return p, parts, None


def dem06(scheme=abi.INTERP_NONE, width=None, u_gas=0.4, fluid_dt=1.0e-3):
    """tests/dem/DEM06: one sphere (d = 0.1 mm, rho = 2000) released at rest
    at (0.005, 0.0925, 0.005) in a 0.01 x 0.10 x 0.01 m column of 5 x 20 x 5
    cells with a frozen upflow V_G = 0.4 m/s (rho_g = 1.2, mu_g = 1.8e-5),
    one-way coupled; KN = KN_W = 0.1, EN = 1, no friction, Euler.  The
    reference case uses a user drag law equal to Wen-Yu at EP_G = 1; here
    the drag is WEN_YU at the cell voidage the deposition gives.  The gas
    pressure is hydrostatic, so P_FORCE carries the buoyancy.

    Returns (params, particles, gas, p_g, substeps): DTSOLID is the
    reference's TCOLL/50 rounded so that `substeps` solid steps make one
    fluid step of `fluid_dt` (des_time_march.f:120-126)."""
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = 1
    p.max_pip = 1
    p.coupled = 1
    p.drag_type = abi.DRAG_WEN_YU
    p.interp_scheme = scheme
    if scheme == abi.INTERP_NONE:
        p.des_interp_on = p.des_interp_mean_fields = 0
    else:
        p.des_interp_on = p.des_interp_mean_fields = 1
    p.des_interp_width = abi.UNDEFINED if width is None else width
    p.xlength, p.ylength, p.zlength = 0.01, 0.10, 0.01
    p.imax, p.jmax, p.kmax = 5, 20, 5
    p.desgridsearch_imax, p.desgridsearch_jmax, p.desgridsearch_kmax = 5, 20, 5
    p.kn = p.kn_w = 0.1
    p.mew = p.mew_w = 0.0
    p.des_en_input[0] = 1.0
    p.des_en_wall_input[0] = 1.0
    p.d_p0[0] = 1.0e-4
    p.ro_s0[0] = 2000.0
    p.max_radius = 5.0e-5
    p.ro_g0 = 1.2
    p.mu_g0 = 1.8e-5
    # TCOLL (check_solids_dem.f:676-694 with e = 1): pi sqrt(m_eff / kn)
    m = 2000.0 * np.pi / 6.0 * 1.0e-12
    tcoll = min(np.pi * np.sqrt(0.5 * m / 0.1), np.pi * np.sqrt(m / 0.1))
    substeps = int(np.ceil(np.float32(fluid_dt / (tcoll / 50.0))))
    p.dtsolid = fluid_dt / substeps
    parts = make_particles(np.array([[0.005], [0.0925], [0.005]]), 5.0e-5, 2000.0)
    nc = n_cells(p)
    gas = Gas(np.zeros(nc), np.full(nc, u_gas), np.zeros(nc), np.zeros((nc, 3)))
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    dy = p.ylength / p.jmax
    j = np.tile(np.arange(1, nj + 1), ni * nk)      # J fastest (MFiX cell order)
    y = (j - 1.5) * dy
    p_g = p.ro_g0 * -p.grav[1] * (p.ylength - y)
    return p, parts, gas, np.ascontiguousarray(p_g), substeps


def dem05(rows, model="LSD", integrator=abi.ADAMS_BASHFORTH):
    """tests/dem/DEM05: oblique impacts, no gravity.  `rows` are the 93
    lines of its particle_input.dat (x y z r rho u v w); rho 4000 is phase
    1, 5000 phase 2.  MEW = MEW_W = 0.092, EN = 1 everywhere.  LSD: KN = KN_W
    = 1.72e7, KT_FAC = KT_W_FAC = 1.48/1.72.  HERTZ: E = 380e9 / 70e9 (wall
    70e9), nu = 0.23 / 0.25 (wall 0.25), ET = 1."""
    a = np.asarray(rows, dtype=np.float64)
    p = abi.default_params()
    p.coll_model = abi.HERTZ if model == "HERTZ" else abi.LSD
    p.integrator = integrator
    p.nphases = 2
    p.max_pip = a.shape[0]
    p.xlength, p.ylength, p.zlength = 0.5, 0.1, 0.5
    p.imax, p.jmax, p.kmax = 5, 1, 5
    p.grav[:] = [0.0, 0.0, 0.0]
    p.mew = p.mew_w = 0.092
    for i in range(3):
        p.des_en_input[i] = 1.0
        p.des_et_input[i] = 1.0
    for i in range(2):
        p.des_en_wall_input[i] = 1.0
        p.des_et_wall_input[i] = 1.0
    p.kn = p.kn_w = 1.72e7
    p.kt_fac = p.kt_w_fac = 1.48 / 1.72
    p.e_young[0], p.e_young[1] = 380.0e9, 70.0e9
    p.v_poisson[0], p.v_poisson[1] = 0.23, 0.25
    p.ew_young, p.vw_poisson = 70.0e9, 0.25
    p.d_p0[0] = p.d_p0[1] = 5.0e-3
    p.ro_s0[0], p.ro_s0[1] = 4000.0, 5000.0
    p.max_radius = float(a[:, 3].max())
    phase = np.where(a[:, 4] > 4500.0, 2, 1).astype(np.int32)
    parts = make_particles(a[:, 0:3].T, a[:, 3], a[:, 4], vel=a[:, 5:8].T, phase=phase)
    return p, parts, None


def ensemble_member(n, kn, en, mew, seed, phi=0.30):
    """One member of a parameter-sweep ensemble: n spheres of d = 1 mm,
    rho = 2500 (as the settling bed) placed by RSA in an X = Z, Y = 4X box
    sized for solids fraction phi over the region the centres can reach;
    LSD with KN = KN_W = kn, EN = EN_WALL = en, MEW = MEW_W = mew,
    DTSOLID = TCOLL/50 (the reference's rule)."""
    d = 1.0e-3
    vol = n * (np.pi / 6.0) * d ** 3 / phi
    lx = (vol / 4.0) ** (1.0 / 3.0) + d
    for _ in range(50):  # (lx - d)^2 (4 lx - d) = vol
        f = (lx - d) ** 2 * (4 * lx - d) - vol
        df = 2 * (lx - d) * (4 * lx - d) + 4 * (lx - d) ** 2
        lx -= f / df
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.integrator = abi.EULER
    p.nphases = 1
    p.max_pip = n
    p.xlength, p.ylength, p.zlength = lx, 4.0 * lx, lx
    p.imax = p.jmax = p.kmax = 1
    p.grav[:] = [0.0, -9.81, 0.0]
    p.kn = p.kn_w = kn
    p.mew = p.mew_w = mew
    p.des_en_input[0] = en
    p.des_en_wall_input[0] = en
    p.d_p0[0] = 1.0e-3
    p.ro_s0[0] = 2500.0
    p.max_radius = 0.5e-3
    r = np.full(n, 0.5e-3)
    pos = rsa_positions(seed, r, [0, 0, 0], [lx, 4.0 * lx, lx])
    return p, make_particles(pos, r, 2500.0)


def sweep_ensemble(nsys=256, seed=11, nmin=2000, nmax=200_000, phi=0.30, member=None):
    """Ensemble of independent settling beds of uneven size: N log-uniform in
    [nmin, nmax], KN log-uniform in [1e2, 1e4], EN ~ U[0.7, 0.95],
    MEW ~ U[0.1, 0.5] (seeded).  member(n, kn, en, mew, seed) builds one
    system (default: RSA packing, ensemble_member)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys)).astype(int)
    kns = np.exp(rng.uniform(np.log(1e2), np.log(1e4), nsys))
    ens = rng.uniform(0.7, 0.95, nsys)
    mews = rng.uniform(0.1, 0.5, nsys)
    if member is None:
        return [ensemble_member(int(ns[s]), float(kns[s]), float(ens[s]), float(mews[s]),
                                seed * 1000 + s, phi) for s in range(nsys)]
    return [member(int(ns[s]), float(kns[s]), float(ens[s]), float(mews[s]), seed * 1000 + s)
            for s in range(nsys)]


def lattice_member(n, kn, en, mew, seed, phi=0.30, precision=abi.MIXED):
    """A sweep member whose spheres (d = 1 mm, rho = 2500) sit on a jittered
    cubic lattice filling an X = Z, Y = 4X box at solids fraction phi (no
    overlaps); O(n) to build, for request streams.  LSD, KN = KN_W = kn,
    EN = EN_WALL = en, MEW = MEW_W = mew, DTSOLID = TCOLL/50."""
    d = 1.0e-3
    a = (np.pi / 6.0 * d ** 3 / phi) ** (1.0 / 3.0)   # lattice spacing
    m = max(1, int(np.ceil((n / 4.0) ** (1.0 / 3.0))))
    nx, ny, nz = m, int(np.ceil(n / (m * m))), m
    lx, ly, lz = nx * a, max(ny * a, 4.0 * nx * a), nz * a
    rng = np.random.default_rng(seed)
    ii = np.arange(nx * ny * nz)[:n]
    ix, iy, iz = ii % nx, (ii // nx) % ny, ii // (nx * ny)
    jit = 0.5 * (a - d) * rng.uniform(-1.0, 1.0, (3, n))
    pos = np.stack([(ix + 0.5) * a, (iy + 0.5) * a, (iz + 0.5) * a]) + jit
    p = abi.default_params()
    p.coll_model = abi.LSD
    p.precision = precision
    p.nphases = 1
    p.max_pip = n
    p.xlength, p.ylength, p.zlength = lx, ly, lz
    p.imax = p.jmax = p.kmax = 1
    p.grav[:] = [0.0, -9.81, 0.0]
    p.kn = p.kn_w = kn
    p.mew = p.mew_w = mew
    p.des_en_input[0] = en
    p.des_en_wall_input[0] = en
    p.d_p0[0] = d
    p.ro_s0[0] = 2500.0
    p.max_radius = 0.5 * d
    return p, make_particles(pos, 0.5 * d, 2500.0)


def funnel_facets(box, y_lo, y_hi, outlet):
    """An STL funnel (a square hopper) inside the box (X, Y-up, Z): four
    inclined faces from the box's side walls at y_hi down to a square outlet
    of half-width `outlet` around the box's vertical axis at y_lo, two
    triangles per face (the shared diagonal is a coplanar edge).  Returns
    (vertex[f, axis, v], normal[f, axis], planes) with the normals pointing
    into the funnel and planes = [(point, normal)] of the four faces."""
    lx, _, lz = box
    cx, cz = 0.5 * lx, 0.5 * lz
    top = [(0.0, 0.0), (lx, 0.0), (lx, lz), (0.0, lz)]
    bot = [(cx - outlet, cz - outlet), (cx + outlet, cz - outlet),
           (cx + outlet, cz + outlet), (cx - outlet, cz + outlet)]
    verts, norms, planes = [], [], []
    for k in range(4):
        a = np.array([top[k][0], y_hi, top[k][1]])
        b = np.array([top[(k + 1) % 4][0], y_hi, top[(k + 1) % 4][1]])
        c = np.array([bot[(k + 1) % 4][0], y_lo, bot[(k + 1) % 4][1]])
        d = np.array([bot[k][0], y_lo, bot[k][1]])
        n = np.cross(b - a, d - a)
        n = n / np.linalg.norm(n)
        if n[1] < 0.0:
            n = -n
        for tri in ((a, c, b), (a, d, c)):   # right-handed about n
            verts.append(np.stack(tri, axis=1))   # [axis, v]
            norms.append(n)
        planes.append((a, n))
    return np.array(verts), np.array(norms), planes


def inside_funnel(pos, radius, planes, margin=1.0):
    """Mask of the spheres on the inner side of every funnel face by more
    than margin * radius."""
    keep = np.ones(pos.shape[1], dtype=bool)
    for a, n in planes:
        keep &= (n @ (pos - a[:, None])) > margin * radius
    return keep

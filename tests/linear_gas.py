"""A gas field linear in space on MFiX's staggered grid, and the closed form
of what the filter-scheme drag (DRAG_GS_DES1, drag_gs_des1.f:50-160) must
see at a particle for it.

U_G lives on east faces (x = XE(I)), V_G on north faces (y = YN(J)), W_G on
top faces (z = ZT(K)); the cell-centre velocity of CALC_CELL_CENTER_GAS_VEL
(drag_gs_des1.f:313-370) averages the two faces of a cell, which gives the
field at the cell centre exactly for a linear field.  P_FORCE(:,IJK) is
given at the cell centres.  Interpolated at a particle:

* no interpolation (DES_INTERP_ON false, or the NONE scheme): the value at
  the centre of the particle's own cell;
* LINEAR_HAT with DES_INTERP_ON: the hat weights about the cell centres
  reproduce a linear field, so the value at the particle itself (in cells
  whose 26 neighbours are fluid);
* SQUARE_DPVM / GAUSS_DPVM with DES_INTERP_ON: per axis, weight w_lo of the
  lower neighbour and w_hi of the upper one are the fractions of a sphere of
  radius DES_INTERP_WIDTH/2 about the particle beyond the two faces; the
  interpolated value is f(cell centre) + sum_a g_a dx_a (w_hi,a - w_lo,a),
  with the sphere fractions found here by quadrature of the sphere's cross
  sections (an independent route to the kernel mass).

A mistake in which face feeds a cell-centre value, in which cell a weight is
applied to or in the axis order shifts the result by a multiple of g dx,
which these closed forms see.
"""
from __future__ import annotations

import warnings

import numpy as np
from scipy import integrate

from mfixdem import abi, workloads


class LinearField:
    def __init__(self, c0, grad):
        self.c0 = np.asarray(c0, dtype=np.float64)      # (3,) value at the origin
        self.grad = np.asarray(grad, dtype=np.float64)  # (3, 3): grad[comp, axis]

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        return self.c0 + self.grad @ x


def cell_sizes(p):
    """Per axis, the sizes of cells 1..N+2 (index cell - 1): the given
    interior sizes (p.dx ...) or the uniform ALENGTH/N, ghosts copying
    their neighbour."""
    out = []
    for ptr, n, length in ((p.dx, p.imax, p.xlength), (p.dy, p.jmax, p.ylength),
                           (p.dz, p.kmax, p.zlength)):
        inner = (np.array([ptr[i] for i in range(n)]) if ptr else
                 np.full(n, length / n))
        out.append(np.concatenate([[inner[0]], inner, [inner[-1]]]))
    return out


def grid(p):
    """Face arrays XE(0..IMAX+2), YN, ZT (0-based index = face index) and
    the cell sizes of each axis (cell_sizes)."""
    dc = cell_sizes(p)
    lo = np.array([p.x_min, p.y_min, p.z_min])
    f = []
    for a in range(3):
        fa = [lo[a] - dc[a][0]]
        for c in dc[a]:
            fa.append(fa[-1] + c)
        f.append(np.array(fa))
    return f, dc


def centre(f, dc, a, cell):
    """Centre of (1-based) cell `cell` along axis a."""
    return f[a][cell - 1] + 0.5 * dc[a][cell - 1]


def gas_fields(p, vel: LinearField, pforce: LinearField):
    """Gas with U_G/V_G/W_G sampled on their staggered faces and P_FORCE at
    the cell centres (MFiX serial layout, J fastest, then I, then K)."""
    f, dc = grid(p)
    ni, nj, nk = p.imax + 2, p.jmax + 2, p.kmax + 2
    K, I, J = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), np.arange(1, nj + 1),
                          indexing="ij")
    xc = centre(f, dc, 0, I)
    yc = centre(f, dc, 1, J)
    zc = centre(f, dc, 2, K)
    u = vel.c0[0] + vel.grad[0, 0] * f[0][I] + vel.grad[0, 1] * yc + vel.grad[0, 2] * zc
    v = vel.c0[1] + vel.grad[1, 0] * xc + vel.grad[1, 1] * f[1][J] + vel.grad[1, 2] * zc
    w = vel.c0[2] + vel.grad[2, 0] * xc + vel.grad[2, 1] * yc + vel.grad[2, 2] * f[2][K]
    pf = np.stack([pforce.c0[c] + pforce.grad[c, 0] * xc + pforce.grad[c, 1] * yc +
                   pforce.grad[c, 2] * zc for c in range(3)], axis=-1)
    return workloads.Gas(u.ravel().copy(), v.ravel().copy(), w.ravel().copy(),
                         pf.reshape(-1, 3).copy())


def sphere_fraction_beyond(dist, r):
    """Fraction of a ball of radius r lying beyond a plane at distance dist
    from its centre, by quadrature of the cross-section areas."""
    if dist >= r:
        return 0.0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        v = integrate.quad(lambda s: np.pi * (r * r - s * s), dist, r,
                           epsabs=1e-22, epsrel=1e-14)[0]
    return v / (4.0 / 3.0 * np.pi * r ** 3)


def expected_at(p, x, cell, field: LinearField):
    """Closed-form interpolated value of `field` at particle position x in
    fluid cell `cell` = (I, J, K) (1-based) for the configured scheme."""
    f, dc = grid(p)
    x = np.asarray(x, dtype=np.float64)
    c = np.array([centre(f, dc, a, cell[a]) for a in range(3)])
    scheme = p.interp_scheme
    if not p.des_interp_on or scheme == abi.INTERP_NONE:
        return field(c)
    if scheme == abi.INTERP_LINEAR_HAT:
        return field(x)
    # SQUARE_DPVM / GAUSS_DPVM: sphere of radius DES_INTERP_WIDTH/2; the
    # weights of an axis sum to 1, so a linear field is read at the
    # weighted mean of the three cell centres
    r = 0.5 * p.des_interp_width
    out = field(c).copy()
    for a in range(3):
        w_lo = sphere_fraction_beyond(x[a] - f[a][cell[a] - 1], r)
        w_hi = sphere_fraction_beyond(f[a][cell[a]] - x[a], r)
        c_lo = centre(f, dc, a, cell[a] - 1)
        c_hi = centre(f, dc, a, cell[a] + 1)
        out = out + field.grad[:, a] * (w_hi * (c_hi - c[a]) - w_lo * (c[a] - c_lo))
    return out


def sparse_particles(p, n, seed, v_scale=0.3):
    """n spheres of the bed's d_p0 in cells whose 26 neighbours are fluid,
    far enough apart that no pair is listed and away from every wall."""
    rng = np.random.default_rng(seed)
    f, _ = grid(p)
    r = 0.5 * p.d_p0[0]
    lo = np.array([f[a][2] for a in range(3)]) + r
    hi = np.array([f[a][[p.imax, p.jmax, p.kmax][a]] for a in range(3)]) - r
    pts = []
    while len(pts) < n:
        x = lo + (hi - lo) * rng.random(3)
        if all(np.linalg.norm(x - q) > 3.0 * (2 * r) for q in pts):
            pts.append(x)
    pos = np.array(pts).T.copy()
    vel = v_scale * (rng.random((3, n)) - 0.5)
    return workloads.make_particles(pos, r, p.ro_s0[0], vel=vel)


def cell_of(p, x):
    f, _ = grid(p)
    return tuple(int(np.searchsorted(f[a], x[a], side="right")) for a in range(3))



def textbook_fgp(drag, eps, mu, rho, vrel, dp, pvol):
    """Drag coefficient times particle volume: Ergun (eps <= 0.8, Gidaspow
    only) or Wen-Yu with Cd = 24/Re (1 + 0.15 Re^0.687), Re = eps rho dp |u|/mu,
    0.44 above Re = 1000; both per unit solids volume (beta / eps_s).  Like
    the reference (drag_gs.f DRAG_GIDASPOW / DRAG_WEN_YU), Re = 0 gives 0 in
    either branch; the force F_gp (U - v) is then zero anyway."""
    if vrel == 0.0:
        return 0.0
    if drag == abi.DRAG_GIDASPOW and eps <= 0.8:
        return (150.0 * (1.0 - eps) * mu / (eps * dp * dp) + 1.75 * rho * vrel / dp) * pvol
    re = dp * vrel * rho * eps / mu
    if re == 0.0:
        return 0.0
    cd = 24.0 / (re + 1e-15) * (1.0 + 0.15 * re ** 0.687) if re <= 1000.0 else 0.44
    return 0.75 * cd * vrel * rho * eps * eps ** (-2.65) / dp * pvol


# beds of the filter-scheme pins (0.02 x 0.04 x 0.02 m, 10 x 20 x 10 cells)
LX, LY, LZ, CELLS = 0.02, 0.04, 0.02, (10, 20, 10)
LINEAR_CASES = [(abi.INTERP_LINEAR_HAT, 1, None), (abi.INTERP_SQUARE_DPVM, 1, 1.6e-3),
                (abi.INTERP_GAUSS_DPVM, 1, 1.2e-3), (abi.INTERP_LINEAR_HAT, 0, None),
                (abi.INTERP_NONE, 0, None)]
LINEAR_IDS = ["lhat-on", "square-on", "gauss-on", "lhat-off", "none"]


def linear_expectation(p, parts, ep, filt, match_gas_velocity=False):
    """(particle velocities, FC the drag must give) for every particle:
    F_gp(lEPg, |VELFP - v|) (VELFP - v) + lPF PVOL (drag_gs_des1.f:96-157)
    with VELFP and lPF in closed form.  filt(q) -> (cells, weights) of particle q for
    the voidage; with match_gas_velocity the particles move with the
    closed-form gas velocity, so only the pressure force remains."""
    n = parts.n
    vel = np.zeros((3, n))
    want = np.zeros((3, n))
    for q in range(n):
        x = parts.pos[:, q]
        cell = cell_of(p, x)
        u = expected_at(p, x, cell, LIN_VEL)
        pf = expected_at(p, x, cell, LIN_PF)
        if p.des_interp_on and p.interp_scheme != abi.INTERP_NONE:
            cells, w = filt(q)
            eps = float(np.sum(ep[cells] * w))
        else:
            own = (cell[1] - 1) + (cell[0] - 1) * (p.jmax + 2) + \
                (cell[2] - 1) * (p.jmax + 2) * (p.imax + 2)
            eps = float(ep[own])
        v = u if match_gas_velocity else parts.vel[:, q]
        vel[:, q] = v
        r = parts.radius[q]
        pvol = (4.0 / 3.0) * np.pi * r ** 3
        fgp = textbook_fgp(p.drag_type, eps, p.mu_g0, p.ro_g0, np.linalg.norm(u - v), 2 * r,
                           pvol)
        want[:, q] = fgp * (u - v) + pf * pvol
    return vel, want


def _lin_fields():
    vel = LinearField([0.1, 1.2, -0.2], [[3.0, -5.0, 7.0], [11.0, -2.0, 4.0],
                                            [-6.0, 9.0, 2.5]])
    pf = LinearField([300.0, 1.2e3, -500.0], [[2.0e4, -1.1e4, 7.0e3],
                                                  [-9.0e3, 1.5e4, 3.0e3],
                                                  [5.0e3, 8.0e3, -1.7e4]])
    return vel, pf


LIN_VEL, LIN_PF = _lin_fields()


def stretched_sizes(n, length, ratio):
    """n cell sizes growing geometrically by `ratio`, summing to length."""
    g = ratio ** np.arange(n)
    d = length * g / g.sum()
    d[-1] = length - d[:-1].sum()
    return d


def linear_case(scheme, interp_on, width, drag=abi.DRAG_GIDASPOW, n=24, seed=9,
                stretch=None):
    """A sparse bed (no pairs, no wall contacts) in the linear gas field;
    stretch = growth ratio of a non-uniform grid (None: uniform)."""
    p, _, _ = workloads.fluidised_bed(n=8, seed=seed, lx=LX, ly=LY, lz=LZ, cells=CELLS,
                                      dg=(0, 0, 0))
    if stretch is not None:
        p.set_cell_sizes(stretched_sizes(p.imax, p.xlength, stretch),
                         stretched_sizes(p.jmax, p.ylength, 1.0 / stretch),
                         stretched_sizes(p.kmax, p.zlength, stretch ** 0.5))
    p.max_pip = n
    p.max_radius = 5.0e-4
    p.drag_type = drag
    p.interp_scheme = scheme
    p.des_interp_on = interp_on
    p.des_interp_mean_fields = interp_on
    p.des_interp_width = abi.UNDEFINED if width is None else width
    parts = sparse_particles(p, n, seed)
    gas = gas_fields(p, LIN_VEL, LIN_PF)
    return p, parts, gas

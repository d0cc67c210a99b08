// host_setup.hpp -- one-time host set-up of a DEM handle: collision
// parameters, fluid and DES grids, and the wall facets binned to DES cells.
//
// Reference routines restated here (paths relative to the reference's model/):
//   CHECK_SOLIDS_DEM_COLL_LSD    check_data/check_solids_dem.f:519-782
//   CHECK_SOLIDS_DEM_COLL_HERTZ  check_data/check_solids_dem.f:798-1122
//   REF_MASS / REF_RADIUS        check_data/check_solids_dem.f:1162-1163
//   CHECK_AXIS (uniform DX)      check_data/check_axis.f:152-156
//   SET_GEOMETRY_DES (faces)     des/set_geometry_des.f:72-87
//   VOL, VOL_SURR                set_geometry1.f:125, des/make_arrays_des.f:78-101
//   DES grid default / init      check_data/check_geometry.f:134-170,
//                                des/desgrid_mod.f:572-574,833-847
//   default walls -> triangles   des/stl_preproc_des_mod.f:973-1047
//   BIN_FACETS_TO_DG / ADD_FACET des/stl_preproc_des_mod.f:135-345
//   edge-collision flags         des/stl_preproc_des_mod.f:362-480
//   TRI_BOX_OVERLAP              des/stl_functions_des_mod.f:280-536
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mfixdem.h"

namespace mfixdem {

constexpr double kPi = 3.141592653589793;  // 4.D0*ATAN(1.D0)
constexpr double kUndef = MFIXDEM_UNDEFINED;

struct Facet {
  double v[3][3];   // v[vertex][axis]
  double n[3];      // NORM_FACE
  int dir;          // axis of smallest extent (ADD_FACET)
  double fmin, fmax;
  uint8_t ignore_edge[4];
  double vt[3];     // STL_TANGENTIAL_VEL
  double vn[3];     // STL_NORMAL_VEL
};

struct Setup {
  mfixdem_params p{};
  mfixdem_derived d{};
  // fluid grid: cells 1..ni (ghosts included), faces 0..ni
  int ni = 0, nj = 0, nk = 0, ncell = 0;
  // cell sizes DX(I), I = 1..ni, at cdx[I-1] (the ghost cells take their
  // interior neighbour's size); smallest interior sizes; largest cell volume
  std::vector<double> cdx, cdy, cdz;
  double dx_min = 0, dy_min = 0, dz_min = 0, vol_max = 0;
  double x_max = 0, y_max = 0, z_max = 0;
  std::vector<double> xe, yn, zt, vol, vol_surr;
  std::vector<uint8_t> fluid;
  // DES grid (ghost layer included)
  int dg_ni = 0, dg_nj = 0, dg_nk = 0, dg_ncell = 0;
  int coarse[3] = {0, 0, 0};
  // facets (user STL first, then the default box walls) and the facets of
  // each DES cell: ids dg_fids[dg_foff[c] .. dg_foff[c+1]), ascending
  std::vector<Facet> facets;
  std::vector<int> dg_foff, dg_fids;
  int user_facets = 0;
  bool default_walls = true;
};

inline int fluid_ijk(const Setup& s, int i, int j, int k) {  // 1-based in, 0-based out
  return (j - 1) + (i - 1) * s.nj + (k - 1) * s.nj * s.ni;
}
inline int dg_ijk(const Setup& s, int i, int j, int k) {
  return (j - 1) + (i - 1) * s.dg_nj + (k - 1) * s.dg_nj * s.dg_ni;
}
inline int floor_sat(double v) {
  if (!(v > -1.0e9)) return -1000000000;
  if (v > 1.0e9) return 1000000000;
  return (int)std::floor(v);
}

inline void derive_lsd(Setup& s) {
  const mfixdem_params& p = s.p;
  double tcoll = kUndef;
  double ktf = p.kt_fac == kUndef ? 2.0 / 7.0 : p.kt_fac;
  double ktwf = p.kt_w_fac == kUndef ? 2.0 / 7.0 : p.kt_w_fac;
  double etf = p.des_etat_fac == kUndef ? 0.5 : p.des_etat_fac;
  double etwf = p.des_etat_w_fac == kUndef ? 0.5 : p.des_etat_w_fac;
  s.d.kt = ktf * p.kn;
  s.d.kt_w = ktwf * p.kn_w;
  int lc = 0;
  for (int m = 0; m < p.nphases; m++) {
    const double dm = p.d_p0[m];
    const double mass_m = (kPi / 6.0) * (dm * dm * dm) * p.ro_s0[m];
    for (int l = m; l < p.nphases; l++) {
      const double en = p.des_en_input[lc++];
      const double dl = p.d_p0[l];
      const double mass_l = (kPi / 6.0) * (dl * dl * dl) * p.ro_s0[l];
      const double meff = mass_m * mass_l / (mass_m + mass_l);
      double etan, tc;
      if (en != 0.0) {
        const double lg = std::log(en);
        etan = 2.0 * std::sqrt(p.kn * meff) * std::fabs(lg);
        etan = etan / std::sqrt(kPi * kPi + lg * lg);
        const double q = etan / meff;
        tc = kPi / std::sqrt(p.kn / meff - (q * q) / 4.0);
      } else {
        etan = 2.0 * std::sqrt(p.kn * meff);
        tc = kPi / std::sqrt(p.kn / meff);
      }
      if (tc < tcoll) tcoll = tc;
      etan = etan / std::sqrt(meff);
      s.d.des_etan[m][l] = s.d.des_etan[l][m] = etan;
      s.d.des_etat[m][l] = s.d.des_etat[l][m] = etf * etan;
    }
    const double en = p.des_en_wall_input[m];
    double etan;
    if (en != 0.0) {
      const double lg = std::log(en);
      etan = 2.0 * std::sqrt(p.kn_w * mass_m) * std::fabs(lg);
      etan = etan / std::sqrt(kPi * kPi + lg * lg);
    } else {
      etan = 2.0 * std::sqrt(p.kn_w * mass_m);
    }
    etan = etan / std::sqrt(mass_m);
    s.d.des_etan_wall[m] = etan;
    s.d.des_etat_wall[m] = etwf * etan;
  }
  s.d.tcoll = tcoll;
}

// Hertz damping / stiffness in the particle-agnostic form the force routines use.
inline void hertz_eta(double k, double meff, double red_meff, double e, bool tangential,
                      double* eta) {
  double v;
  const double mm = tangential ? red_meff : meff;
  if (e != 0.0) {
    const double lg = std::log(e);
    v = std::sqrt(5.0) * std::sqrt(k * mm) * std::fabs(lg);
    v = v / std::sqrt(kPi * kPi + lg * lg);
  } else {
    v = std::sqrt(5.0) * std::sqrt(k * mm);
  }
  *eta = v;
}

inline void derive_hertz(Setup& s) {
  const mfixdem_params& p = s.p;
  double tcoll = kUndef;
  double g[MFIXDEM_MAX_PHASES];
  const double gw = 0.5 * p.ew_young / (1.0 + p.vw_poisson);
  for (int m = 0; m < p.nphases; m++) g[m] = 0.5 * p.e_young[m] / (1.0 + p.v_poisson[m]);
  int lc = 0;
  for (int m = 0; m < p.nphases; m++) {
    const double dm = p.d_p0[m];
    const double mass_m = (kPi / 6.0) * (dm * dm * dm) * p.ro_s0[m];
    const double rm = 0.5 * dm;
    for (int l = m; l < p.nphases; l++) {
      const double en = p.des_en_input[lc], et = p.des_et_input[lc];
      lc++;
      const double dl = p.d_p0[l];
      const double mass_l = (kPi / 6.0) * (dl * dl * dl) * p.ro_s0[l];
      const double rl = 0.5 * dl;
      const double meff = (mass_m * mass_l) / (mass_m + mass_l);
      const double red = (2.0 / 7.0) * meff;
      const double reff = rm * rl / (rm + rl);
      const double vm = p.v_poisson[m], vl = p.v_poisson[l];
      const double eeff = p.e_young[m] * p.e_young[l] /
                          (p.e_young[m] * (1.0 - vl * vl) + p.e_young[l] * (1.0 - vm * vm));
      const double geff = g[m] * g[l] / (g[m] * (2.0 - vl) + g[l] * (2.0 - vm));
      double hkn = (4.0 / 3.0) * std::sqrt(reff) * eeff;
      double hkt = 8.0 * std::sqrt(reff) * geff;
      double etan, etat;
      hertz_eta(hkn, meff, red, en, false, &etan);
      hertz_eta(hkt, meff, red, et, true, &etat);
      double tc;
      if (en != 0.0) {
        const double q = etan / meff;
        tc = kPi / std::sqrt(hkn / meff - (q * q) / 5.0);
      } else {
        tc = kPi / std::sqrt(hkn / meff);
      }
      if (tc < tcoll) tcoll = tc;
      etan = etan / std::sqrt(hkn * meff);
      etat = etat / std::sqrt(hkt * meff);
      hkn = hkn / std::sqrt(reff);
      hkt = hkt / std::sqrt(reff);
      s.d.hert_kn[m][l] = s.d.hert_kn[l][m] = hkn;
      s.d.hert_kt[m][l] = s.d.hert_kt[l][m] = hkt;
      s.d.des_etan[m][l] = s.d.des_etan[l][m] = etan;
      s.d.des_etat[m][l] = s.d.des_etat[l][m] = etat;
    }
    const double en = p.des_en_wall_input[m], et = p.des_et_wall_input[m];
    const double meff = mass_m, red = (2.0 / 7.0) * meff, reff = rm;
    const double vm = p.v_poisson[m], vw = p.vw_poisson;
    const double eeff = p.e_young[m] * p.ew_young /
                        (p.e_young[m] * (1.0 - vw * vw) + p.ew_young * (1.0 - vm * vm));
    const double geff = g[m] * gw / (g[m] * (2.0 - vw) + gw * (2.0 - vm));
    double hkwn = (4.0 / 3.0) * std::sqrt(reff) * eeff;
    double hkwt = 8.0 * std::sqrt(reff) * geff;
    double etan, etat;
    hertz_eta(hkwn, meff, red, en, false, &etan);
    hertz_eta(hkwt, meff, red, et, true, &etat);
    s.d.des_etan_wall[m] = etan / std::sqrt(hkwn * meff);
    s.d.des_etat_wall[m] = etat / std::sqrt(hkwt * meff);
    s.d.hert_kwn[m] = hkwn / std::sqrt(reff);
    s.d.hert_kwt[m] = hkwt / std::sqrt(reff);
  }
  s.d.tcoll = tcoll;
}

inline void setup_grids(Setup& s) {
  const mfixdem_params& p = s.p;
  s.ni = p.imax + 2;
  s.nj = p.jmax + 2;
  s.nk = p.kmax + 2;
  s.ncell = s.ni * s.nj * s.nk;
  s.x_max = p.x_min + p.xlength;
  s.y_max = p.y_min + p.ylength;
  s.z_max = p.z_min + p.zlength;
  // SHIFT_DXYZ (shift_dxyz.f:36-44): interior sizes given or uniform, the
  // ghost cells copy their neighbour
  auto sizes = [](std::vector<double>& c, int n, const double* in, double len, double& mn) {
    c.assign(n + 2, 0.0);
    for (int i = 1; i <= n; i++) c[i] = in ? in[i - 1] : len / (double)n;
    c[0] = c[1];
    c[n + 1] = c[n];
    mn = c[1];
    for (int i = 1; i <= n; i++) mn = std::min(mn, c[i]);
  };
  sizes(s.cdx, p.imax, p.dx, p.xlength, s.dx_min);
  sizes(s.cdy, p.jmax, p.dy, p.ylength, s.dy_min);
  sizes(s.cdz, p.kmax, p.dz, p.zlength, s.dz_min);
  // SET_GEOMETRY_DES (set_geometry_des.f:72-74): XE(0) = X_MIN - DX(1),
  // XE(I) = XE(I-1) + DX(I)
  auto faces = [](std::vector<double>& f, int n, double lo, const std::vector<double>& c) {
    f.assign(n + 1, 0.0);
    f[0] = lo - c[0];
    for (int i = 1; i <= n; i++) f[i] = f[i - 1] + c[i - 1];
  };
  faces(s.xe, s.ni, p.x_min, s.cdx);
  faces(s.yn, s.nj, p.y_min, s.cdy);
  faces(s.zt, s.nk, p.z_min, s.cdz);
  s.vol_max = 0.0;
  s.vol.assign(s.ncell, 0.0);
  s.vol_surr.assign(s.ncell, 0.0);
  s.fluid.assign(s.ncell, 0);
  for (int k = 1; k <= s.nk; k++)
    for (int i = 1; i <= s.ni; i++)
      for (int j = 1; j <= s.nj; j++) {
        const int c = fluid_ijk(s, i, j, k);
        // set_geometry1.f:125, X(I) = 1 (Cartesian)
        s.vol[c] = s.cdx[i - 1] * s.cdy[j - 1] * (1.0 * s.cdz[k - 1]);
        s.vol_max = std::max(s.vol_max, s.vol[c]);
        s.fluid[c] = (i >= 2 && i <= p.imax + 1 && j >= 2 && j <= p.jmax + 1 && k >= 2 &&
                      k <= p.kmax + 1);
      }
  for (int k = 1; k <= p.kmax + 1; k++)
    for (int j = 1; j <= p.jmax + 1; j++)
      for (int i = 1; i <= p.imax + 1; i++) {
        const int c = fluid_ijk(s, i, j, k);
        for (int kk = k; kk <= k + 1; kk++)
          for (int jj = j; jj <= j + 1; jj++)
            for (int ii = i; ii <= i + 1; ii++) {
              const int c2 = fluid_ijk(s, ii, jj, kk);
              if (s.fluid[c2]) s.vol_surr[c] = s.vol_surr[c] + s.vol[c2];
            }
      }
  // DES grid
  const double width = 3.0 * (2.0 * p.max_radius);
  int gi = p.desgridsearch_imax, gj = p.desgridsearch_jmax, gk = p.desgridsearch_kmax;
  // check_geometry.f:134-148: WIDTH_X = min(WIDTH, minval(DX))
  if (gi <= 0) gi = std::max(1, (int)(p.xlength / std::min(width, s.dx_min)));
  if (gj <= 0) gj = std::max(1, (int)(p.ylength / std::min(width, s.dy_min)));
  if (gk <= 0) gk = std::max(1, (int)(p.zlength / std::min(width, s.dz_min)));
  const int isz = (int)std::lround((s.xe[p.imax + 1] - s.xe[1]) / (p.xlength / gi));
  const int jsz = (int)std::lround((s.yn[p.jmax + 1] - s.yn[1]) / (p.ylength / gj));
  const int ksz = (int)std::lround((s.zt[p.kmax + 1] - s.zt[1]) / (p.zlength / gk));
  s.dg_ni = isz + 2;
  s.dg_nj = jsz + 2;
  s.dg_nk = ksz + 2;
  s.dg_ncell = s.dg_ni * s.dg_nj * s.dg_nk;
  s.d.dg_imax = isz;
  s.d.dg_jmax = jsz;
  s.d.dg_kmax = ksz;
  s.d.dg_xstart = s.xe[1];
  s.d.dg_ystart = s.yn[1];
  s.d.dg_zstart = s.zt[1];
  s.d.dg_dxinv = (double)isz / (s.xe[p.imax + 1] - s.d.dg_xstart);
  s.d.dg_dyinv = (double)jsz / (s.yn[p.jmax + 1] - s.d.dg_ystart);
  s.d.dg_dzinv = (double)ksz / (s.zt[p.kmax + 1] - s.d.dg_zstart);
  // coarse half-stencil switch (desgrid_mod.f:1117-1119)
  s.coarse[0] = 1.0 > 2.0 * p.max_radius * p.factor_rlm * s.d.dg_dxinv;
  s.coarse[1] = 1.0 > 2.0 * p.max_radius * p.factor_rlm * s.d.dg_dyinv;
  s.coarse[2] = 1.0 > 2.0 * p.max_radius * p.factor_rlm * s.d.dg_dzinv;
}

// Separating-axis triangle/box overlap (Akenine-Moller), axis order and
// comparisons as in TRI_BOX_OVERLAP.
inline bool tri_box_overlap(const double c[3], const double h[3], const double vt[3][3]) {
  double v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
  for (int i = 0; i < 3; i++) {
    v0[i] = vt[0][i] - c[i];
    v1[i] = vt[1][i] - c[i];
    v2[i] = vt[2][i] - c[i];
  }
  for (int i = 0; i < 3; i++) {
    e0[i] = v1[i] - v0[i];
    e1[i] = v2[i] - v1[i];
    e2[i] = v0[i] - v2[i];
  }
  auto sep = [](double pa, double pb, double rad) {
    const double lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
    return lo > rad || hi < -rad;
  };
  // x-axis tests use (y,z); y-axis tests (x,z); z-axis tests (x,y)
  auto ax_x = [&](double a, double b, const double* pa, const double* pb) {
    return sep(a * pa[1] - b * pa[2], a * pb[1] - b * pb[2],
               std::fabs(a) * h[1] + std::fabs(b) * h[2]);
  };
  auto ax_y = [&](double a, double b, const double* pa, const double* pb) {
    return sep(-a * pa[0] + b * pa[2], -a * pb[0] + b * pb[2],
               std::fabs(a) * h[0] + std::fabs(b) * h[2]);
  };
  auto ax_z = [&](double a, double b, const double* pa, const double* pb) {
    return sep(a * pa[0] - b * pa[1], a * pb[0] - b * pb[1],
               std::fabs(a) * h[0] + std::fabs(b) * h[1]);
  };
  if (ax_x(e0[2], e0[1], v0, v2) || ax_y(e0[2], e0[0], v0, v2) || ax_z(e0[1], e0[0], v1, v2))
    return false;
  if (ax_x(e1[2], e1[1], v0, v2) || ax_y(e1[2], e1[0], v0, v2) || ax_z(e1[1], e1[0], v0, v1))
    return false;
  if (ax_x(e2[2], e2[1], v0, v1) || ax_y(e2[2], e2[0], v0, v1) || ax_z(e2[1], e2[0], v1, v2))
    return false;
  for (int i = 0; i < 3; i++) {
    const double mn = std::min(v0[i], std::min(v1[i], v2[i]));
    const double mx = std::max(v0[i], std::max(v1[i], v2[i]));
    if (mn > h[i] || mx < -h[i]) return false;
  }
  double nrm[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                   e0[0] * e1[1] - e0[1] * e1[0]};
  double vmin[3], vmax[3];
  for (int q = 0; q < 3; q++) {
    if (nrm[q] > 0.0) {
      vmin[q] = -h[q] - v0[q];
      vmax[q] = h[q] - v0[q];
    } else {
      vmin[q] = h[q] - v0[q];
      vmax[q] = -h[q] - v0[q];
    }
  }
  const double dmin = nrm[0] * vmin[0] + nrm[1] * vmin[1] + nrm[2] * vmin[2];
  if (dmin > 0.0) return false;
  const double dmax = nrm[0] * vmax[0] + nrm[1] * vmax[1] + nrm[2] * vmax[2];
  return dmax >= 0.0;
}

// DES_STL_PREPROCESSING (stl_preproc_des_mod.f:80-124): the user's STL
// facets (vertex[(f*3 + axis)*3 + v], MFiX's VERTEX(3,3,N) memory order;
// normal[f*3 + axis], NORM_FACE(3,N)) first, then with default walls
// CONVERT_DEFAULT_WALLS_TO_STL (:973-1047), one large triangle per box face,
// W E S N B T; BIN_FACETS_TO_DG (:135-203) with the +-1 cell halo and
// TRI_BOX_OVERLAP, each cell's facets in ascending id;
// FLAG_FACETS_WHERE_EDGE_COLLISION_IS_IGNORED (:362-480).  Wall velocities
// start at zero.
inline void setup_facets(Setup& s, int nu, const double* vertex, const double* normal,
                         bool default_walls) {
  const mfixdem_params& p = s.p;
  const double xn = p.x_min, xx = s.x_max, yn = p.y_min, yx = s.y_max, zn = p.z_min,
               zx = s.z_max, XL = p.xlength, YL = p.ylength, ZL = p.zlength;
  s.facets.clear();
  s.user_facets = nu;
  s.default_walls = default_walls;
  auto add = [&](const double v[3][3], const double n[3]) {
    Facet f{};
    std::memcpy(f.v, v, sizeof(f.v));
    for (int c = 0; c < 3; c++) f.n[c] = n[c];
    double smallest = 3.4028234663852886e38;  // HUGE(0.0)
    for (int ax = 0; ax < 3; ax++) {
      const double mn = std::min(v[0][ax], std::min(v[1][ax], v[2][ax]));
      const double mx = std::max(v[0][ax], std::max(v[1][ax], v[2][ax]));
      if (std::fabs(mx - mn) < smallest) {
        f.dir = ax;
        f.fmin = mn;
        f.fmax = mx;
        smallest = std::fabs(mx - mn);
      }
    }
    s.facets.push_back(f);
  };
  for (int f = 0; f < nu; f++) {
    double v[3][3], n[3];
    for (int c = 0; c < 3; c++) {
      for (int k = 0; k < 3; k++) v[k][c] = vertex[(f * 3 + c) * 3 + k];
      n[c] = normal[f * 3 + c];
    }
    add(v, n);
  }
  auto box = [&](double a0, double a1, double a2, double b0, double b1, double b2, double c0,
                 double c1, double c2, double n0, double n1, double n2) {
    const double v[3][3] = {{a0, a1, a2}, {b0, b1, b2}, {c0, c1, c2}};
    const double n[3] = {n0, n1, n2};
    add(v, n);
  };
  if (default_walls) {  // one triangle per box face, order W E S N B T
    box(xn, yn, zn, xn, yx + YL, zn, xn, yn, zx + ZL, 1.0, 0.0, 0.0);
    box(xx, yn, zn, xx, yx + YL, zn, xx, yn, zx + ZL, -1.0, 0.0, 0.0);
    box(xn, yn, zn, xx + XL, yn, zn, xn, yn, zx + ZL, 0.0, 1.0, 0.0);
    box(xn, yx, zn, xx + XL, yx, zn, xn, yx, zx + ZL, 0.0, -1.0, 0.0);
    box(xn, yn, zn, xx + XL, yn, zn, xn, yx + YL, zn, 0.0, 0.0, 1.0);
    box(xn, yn, zx, xx + XL, yn, zx, xn, yx + YL, zx, 0.0, 0.0, -1.0);
  }
  s.d.n_facets = (int32_t)s.facets.size();
  const double tol_stl = 1.0e-6;
  const double buffer = 1.1 * p.max_radius;
  const double ldx = 1.0 / s.d.dg_dxinv, ldy = 1.0 / s.d.dg_dyinv, ldz = 1.0 / s.d.dg_dzinv;
  const int iend2 = s.dg_ni, jend2 = s.dg_nj, kend2 = s.dg_nk;
  auto iof = [&](double x) { return floor_sat((x - s.d.dg_xstart) * s.d.dg_dxinv) + 2; };
  auto jof = [&](double y) { return floor_sat((y - s.d.dg_ystart) * s.d.dg_dyinv) + 2; };
  auto kof = [&](double z) { return floor_sat((z - s.d.dg_zstart) * s.d.dg_dzinv) + 2; };
  std::vector<std::vector<int>> lists(s.dg_ncell);
  for (size_t nn = 0; nn < s.facets.size(); nn++) {
    const Facet& f = s.facets[nn];
    double lo[3], hi[3];
    for (int ax = 0; ax < 3; ax++) {
      lo[ax] = std::min(f.v[0][ax], std::min(f.v[1][ax], f.v[2][ax]));
      hi[ax] = std::max(f.v[0][ax], std::max(f.v[1][ax], f.v[2][ax]));
    }
    int I1 = iend2, I2 = 1, J1 = jend2, J2 = 1, K1 = kend2, K2 = 1;
    if (hi[0] >= p.x_min - tol_stl && lo[0] <= s.x_max + tol_stl) {
      I1 = std::max(iof(lo[0]) - 1, 1);
      I2 = std::min(iof(hi[0]) + 1, iend2);
    }
    if (hi[1] >= p.y_min - tol_stl && lo[1] <= s.y_max + tol_stl) {
      J1 = std::max(jof(lo[1]) - 1, 1);
      J2 = std::min(jof(hi[1]) + 1, jend2);
    }
    if (hi[2] >= p.z_min - tol_stl && lo[2] <= s.z_max + tol_stl) {
      K1 = std::max(kof(lo[2]) - 1, 1);
      K2 = std::min(kof(hi[2]) + 1, kend2);
    }
    for (int kk = K1; kk <= K2; kk++)
      for (int jj = J1; jj <= J2; jj++)
        for (int ii = I1; ii <= I2; ii++) {
          const double c[3] = {s.d.dg_xstart + ((double)(ii - 2) + 0.5) * ldx,
                               s.d.dg_ystart + ((double)(jj - 2) + 0.5) * ldy,
                               s.d.dg_zstart + ((double)(kk - 2) + 0.5) * ldz};
          const double h[3] = {0.5 * ldx + buffer, 0.5 * ldy + buffer, 0.5 * ldz + buffer};
          if (tri_box_overlap(c, h, f.v)) lists[dg_ijk(s, ii, jj, kk)].push_back((int)nn);
        }
  }
  s.dg_foff.assign(s.dg_ncell + 1, 0);
  s.dg_fids.clear();
  for (int c = 0; c < s.dg_ncell; c++) {
    s.dg_foff[c + 1] = s.dg_foff[c] + (int)lists[c].size();
    s.dg_fids.insert(s.dg_fids.end(), lists[c].begin(), lists[c].end());
  }
  // facets of one cell sharing an edge at a small angle: edge hits ignored
  const double cos_nb = std::cos(5.0 / 180.0 * kPi) + 1.0e-6;
  auto ec = [&](int f, int e, int ax) {
    const Facet& F = s.facets[f];
    if (e == 0) return 0.5 * (F.v[0][ax] + F.v[1][ax]);
    if (e == 1) return 0.5 * (F.v[0][ax] + F.v[2][ax]);
    return 0.5 * (F.v[1][ax] + F.v[2][ax]);
  };
  for (int cell = 0; cell < s.dg_ncell; cell++) {
    const std::vector<int>& L = lists[cell];
    for (int f1 : L)
      for (int e = 0; e < 3; e++)
        for (int f2 : L) {
          if (f2 == f1) continue;
          for (int ne = 0; ne < 3; ne++) {
            const double d0 = ec(f2, ne, 0) - ec(f1, e, 0), d1 = ec(f2, ne, 1) - ec(f1, e, 1),
                         d2 = ec(f2, ne, 2) - ec(f1, e, 2);
            if (d0 * d0 + d1 * d1 + d2 * d2 < 1.0e-9) {
              const double* a = s.facets[f1].n;
              const double* b = s.facets[f2].n;
              if (a[0] * b[0] + a[1] * b[1] + a[2] * b[2] > cos_nb) {
                s.facets[f1].ignore_edge[e + 1] = 1;
                s.facets[f2].ignore_edge[ne + 1] = 1;
              }
            }
          }
        }
  }
}

// Moving walls of a BC_STL group (get_stl_data.f:1484-1502): facets
// [first, first+count) move with vel + omega x (facet centre - rot_center);
// STL_NORMAL_VEL is that velocity's component along the facet normal,
// STL_TANGENTIAL_VEL the rest.
inline void set_wall_motion(Setup& s, int first, int count, const double vel[3],
                            const double omega[3], const double rot_center[3]) {
  for (int f = first; f < first + count; f++) {
    Facet& F = s.facets[f];
    double center[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) center[c] = center[c] + F.v[v][c];
    for (int c = 0; c < 3; c++) center[c] = center[c] / 3.0;
    const double op[3] = {center[0] - rot_center[0], center[1] - rot_center[1],
                          center[2] - rot_center[2]};
    const double cr[3] = {omega[1] * op[2] - omega[2] * op[1], omega[2] * op[0] - omega[0] * op[2],
                          omega[0] * op[1] - omega[1] * op[0]};
    double avg[3];
    for (int c = 0; c < 3; c++) avg[c] = vel[c] + cr[c];
    const double an = avg[0] * F.n[0] + avg[1] * F.n[1] + avg[2] * F.n[2];
    for (int c = 0; c < 3; c++) F.vn[c] = an * F.n[c];
    for (int c = 0; c < 3; c++) F.vt[c] = avg[c] - F.vn[c];
  }
}

// Full set-up; returns MFIXDEM_OK or an error code.
inline int32_t build_setup(const mfixdem_params& p, Setup& s) {
  if (p.abi_version != MFIXDEM_ABI_VERSION) return MFIXDEM_ERR_ARG;
  if (p.nphases < 1 || p.nphases > MFIXDEM_MAX_PHASES) return MFIXDEM_ERR_ARG;
  if (p.max_pip < 1 || p.imax < 1 || p.jmax < 1 || p.kmax < 1) return MFIXDEM_ERR_ARG;
  if (p.max_radius <= 0.0 || p.neighbor_search_n < 1) return MFIXDEM_ERR_ARG;
  if (p.coll_model != MFIXDEM_LSD && p.coll_model != MFIXDEM_HERTZ) return MFIXDEM_ERR_ARG;
  if (p.integrator != MFIXDEM_EULER && p.integrator != MFIXDEM_ADAMS_BASHFORTH)
    return MFIXDEM_ERR_ARG;
  if (p.precision != MFIXDEM_FP64 && p.precision != MFIXDEM_MIXED) return MFIXDEM_ERR_ARG;
  if (p.coupled && p.model_b) return MFIXDEM_ERR_UNSUPPORTED;
  // explicit coupling: DRAG_GS_DES1's F_GP path, for the filter schemes
  if (p.des_explicitly_coupled &&
      (!p.coupled || p.interp_scheme == MFIXDEM_INTERP_GARG_2012))
    return MFIXDEM_ERR_UNSUPPORTED;
  if (p.des_neighbor_search != MFIXDEM_NSEARCH_GRID &&
      p.des_neighbor_search != MFIXDEM_NSEARCH_NSQUARE)
    return MFIXDEM_ERR_ARG;
  // DES_INTERP_SCHEME checks (check_solids_common_discrete.f:752-860)
  switch (p.interp_scheme) {
    case MFIXDEM_INTERP_GARG_2012:
    case MFIXDEM_INTERP_LINEAR_HAT:
      if (p.des_interp_width != kUndef) return MFIXDEM_ERR_ARG;
      break;
    case MFIXDEM_INTERP_NONE:
      if (p.des_interp_on || p.des_interp_mean_fields || p.des_interp_width != kUndef)
        return MFIXDEM_ERR_ARG;
      break;
    case MFIXDEM_INTERP_SQUARE_DPVM:
    case MFIXDEM_INTERP_GAUSS_DPVM:
      if (p.des_interp_width == kUndef) return MFIXDEM_ERR_ARG;
      break;
    default: return MFIXDEM_ERR_ARG;
  }
  if (p.drag_type != MFIXDEM_DRAG_GIDASPOW && p.drag_type != MFIXDEM_DRAG_WEN_YU)
    return MFIXDEM_ERR_ARG;
  // non-uniform cell sizes: positive, and summing to the domain length
  auto sizes_ok = [](const double* c, int n, double len) {
    if (!c) return true;
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
      if (!(c[i] > 0.0) || !std::isfinite(c[i])) return false;
      sum += c[i];
    }
    return std::fabs(sum - len) <= 1e-12 * std::fabs(len);
  };
  if (!sizes_ok(p.dx, p.imax, p.xlength) || !sizes_ok(p.dy, p.jmax, p.ylength) ||
      !sizes_ok(p.dz, p.kmax, p.zlength))
    return MFIXDEM_ERR_ARG;
  s = Setup{};
  s.p = p;
  if (p.coll_model == MFIXDEM_HERTZ) derive_hertz(s);
  else derive_lsd(s);
  s.d.dtsolid = (p.dtsolid == kUndef) ? s.d.tcoll / p.dtsolid_factor : p.dtsolid;
  setup_grids(s);
  s.p.dx = s.p.dy = s.p.dz = nullptr;  // the sizes now live in cdx, cdy, cdz
  if ((int64_t)s.dg_ni * s.dg_nj * s.dg_nk > (int64_t)1 << 30) return MFIXDEM_ERR_ARG;
  if (s.dg_ni > 1023 || s.dg_nj > 1023 || s.dg_nk > 1023) return MFIXDEM_ERR_UNSUPPORTED;
  // fluid cells are packed in 10 bits per axis on the device
  if (s.ni > 1023 || s.nj > 1023 || s.nk > 1023) return MFIXDEM_ERR_UNSUPPORTED;
  if (p.interp_scheme == MFIXDEM_INTERP_SQUARE_DPVM ||
      p.interp_scheme == MFIXDEM_INTERP_GAUSS_DPVM) {
    // set_filter_des.f:70-87: half-width within a fluid cell and a DES cell
    const double dmin = std::min(std::min(s.dx_min, s.dy_min), s.dz_min);
    const double dgmin = std::min(std::min(1.0 / s.d.dg_dxinv, 1.0 / s.d.dg_dyinv),
                                  1.0 / s.d.dg_dzinv);
    if (0.5 * p.des_interp_width > dmin || 0.5 * p.des_interp_width > dgmin)
      return MFIXDEM_ERR_ARG;
  }
  setup_facets(s, 0, nullptr, nullptr, true);
  return MFIXDEM_OK;
}

}  // namespace mfixdem

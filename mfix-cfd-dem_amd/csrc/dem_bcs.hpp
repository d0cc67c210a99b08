// dem_bcs.hpp -- host side of the DEM mass inlets and outlets.
//
// Set-up: the index box of DES cells each boundary checks
// (SET_DEM_BCMI_IJK, set_bc_dem_mi.f:391-553; SET_BC_DEM_MO,
// set_bc_dem_mo.f:32-177), and for an inlet the window layout of
// LAYOUT_MI_DEM (layout_mi_dem.f:88-444 and its E/W, T/B twins) with the
// seeding rate and schedule of SET_BC_DEM_MI (set_bc_dem_mi.f:18-377).
// Per step: the seeding of MASS_INFLOW_DEM (mass_inflow_dem.f:141-293):
// which inlets seed, how many particles, their phase, position and global
// id (the lowest free one).  The per-particle checks of both routines run on
// the device (k_dem_bcs; step_tail's exit velocity).
//
// RANDOM_NUMBER is stood in for by a seeded splitmix64 stream (Steele, Lea
// and Flood, 2014), 53 bits mapped to [0, 1), drawn in the reference's order:
// the window permutation at set-up, then three numbers per inlet per step
// whether or not it seeds.
#pragma once

#include <cmath>
#include <cstdint>
#include <set>
#include <vector>

#include "../../include/mfixdem.h"
#include "host_setup.hpp"
#include "kernels.hpp"

namespace mfixdem {

constexpr int kNumFracLimit = 10000;  // NUMFRAC_LIMIT (des_bc_mod.f)

struct DemBoundary {
  mfixdem_dem_bc in{};
  int ax = 0, pa = 0, qa = 0;  // normal axis, the two in-plane axes
  bool low = false;            // at the low end of ax (the domain on its high side)
  double face = 0.0;           // XE / YN / ZT of the plane
  int lo[3]{}, hi[3]{};        // DES cell box
  // inlet
  int occupants = 0, vacancy = 1, pi_count = 0, lcell = 0;
  bool poly = false;
  double pi_factor = 0.0, mi_time = 0.0, window = 0.0, offset = 0.0;
  std::vector<double> wp, wq;  // window corners (1-based: entry 0 unused)
  std::vector<int> wi, wj;     // fluid cell of each window along P and Q
  std::vector<int> layout;     // DEM_BC_POLY_LAYOUT: 0-based phase per entry
  double vel[kMaxPh][3]{};     // seeding velocity per phase
};

struct DemBoundaries {
  std::vector<DemBoundary> b;
  uint64_t rng = 0;
  double s_time = 0.0;          // S_TIME of the step being taken
  std::set<int> free_gids;      // NONEXISTENT global ids
  std::vector<SeedDev> seeds;   // this step's seeds
  int max_seeds = 0;            // most seeds one step can make

  double uniform() {
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

namespace bcs_detail {

// CALC_CELL_INTERSECT (calc_cell.f:129-180): the interior cell (2..n+1)
// whose faces bracket loc, both ends inclusive and the lower cell first;
// sizes c[I-1] of cell I.
inline int cell_intersect(double rmin, double loc, const std::vector<double>& c, int n) {
  double lo = rmin, hi = rmin;
  for (int cell = 2; cell <= n + 1; cell++) {
    hi = lo + c[cell - 1];
    if (lo <= loc && loc <= hi) return cell;
    lo = hi;
  }
  if (loc < rmin) return 2;
  if (loc > hi) return n + 1;
  return -1;
}

inline double dg_start(const Setup& S, int a) {
  return a == 0 ? S.d.dg_xstart : (a == 1 ? S.d.dg_ystart : S.d.dg_zstart);
}
inline double dg_inv(const Setup& S, int a) {
  return a == 0 ? S.d.dg_dxinv : (a == 1 ? S.d.dg_dyinv : S.d.dg_dzinv);
}
inline int dg_n(const Setup& S, int a) {
  return a == 0 ? S.dg_ni : (a == 1 ? S.dg_nj : S.dg_nk);
}
// IofPOS (desgrid_mod.f:418-422): DES cell of a coordinate, unclamped
inline int dg_of_pos(const Setup& S, int a, double x) {
  return floor_sat((x - dg_start(S, a)) * dg_inv(S, a)) + 2;
}
// IofBCPOS (desgrid_mod.f:428-432): NINT((x - start) dxinv) + start1 - 1
inline int dg_of_bcpos(const Setup& S, int a, double x) {
  return (int)std::lround((x - dg_start(S, a)) * dg_inv(S, a)) + 1;
}

inline void axes(int plane, int& ax, int& pa, int& qa, bool& low) {
  switch (plane) {
    case MFIXDEM_PLANE_N: case MFIXDEM_PLANE_S: ax = 1; pa = 0; qa = 2; break;
    case MFIXDEM_PLANE_E: case MFIXDEM_PLANE_W: ax = 0; pa = 1; qa = 2; break;
    default: ax = 2; pa = 0; qa = 1; break;
  }
  low = plane == MFIXDEM_PLANE_N || plane == MFIXDEM_PLANE_E || plane == MFIXDEM_PLANE_T;
}

// LAYOUT_MI_DEM and the rate of SET_BC_DEM_MI for one inlet.
inline int32_t layout_inlet(const Setup& S, double dt, DemBoundaries& D, DemBoundary& B) {
  const mfixdem_params& p = S.p;
  const mfixdem_dem_bc& in = B.in;
  const int ax = B.ax, pa = B.pa, qa = B.qa;
  const double lo[3] = {in.x_w, in.y_s, in.z_b}, hi[3] = {in.x_e, in.y_n, in.z_t};
  const std::vector<double>* csz[3] = {&S.cdx, &S.cdy, &S.cdz};
  const int nax[3] = {p.imax, p.jmax, p.kmax};
  const double amin[3] = {p.x_min, p.y_min, p.z_min};
  // phases fed by the inlet and the largest diameter among them
  std::vector<int> ph;
  double max_dia = 0.0;
  for (int m = 0; m < p.nphases; m++)
    if (in.ep_s[m] > 0.0) {
      ph.push_back(m);
      max_dia = std::max(max_dia, p.d_p0[m]);
    }
  if (ph.empty()) return MFIXDEM_ERR_ARG;
  B.poly = ph.size() > 1;
  // windows of the inlet rectangle, one particle each (layout_mi_dem.f:
  // 145-225); the counts go through single precision as REAL does
  const double plen = hi[pa] - lo[pa], qlen = hi[qa] - lo[qa];
  const int nw = (int)std::floor((float)(plen / max_dia));
  const int nh = (int)std::floor((float)(qlen / max_dia));
  if (nw < 1 || nh < 1) return MFIXDEM_ERR_ARG;
  const int jcell = cell_intersect(amin[ax], lo[ax], *csz[ax], nax[ax]);
  const double side = B.low ? -1.0 : 1.0;
  B.offset = lo[ax] + max_dia * side;  // seeds start one diameter outside
  B.lcell = jcell + (int)side;
  B.window = std::min(plen / nw, qlen / nh);
  const double w = B.window;
  std::vector<double> mp(nw), mq(nh);
  std::vector<int> mi(nw), mj(nh);
  const double p0 = lo[pa] + 0.5 * (plen - nw * w), q0 = lo[qa] + 0.5 * (qlen - nh * w);
  for (int k = 0; k < nw; k++) {
    mp[k] = p0 + (double)k * w;
    mi[k] = cell_intersect(amin[pa], mp[k] + 0.5 * w, *csz[pa], nax[pa]);
  }
  for (int k = 0; k < nh; k++) {
    mq[k] = q0 + (double)k * w;
    mj[k] = cell_intersect(amin[qa], mq[k] + 0.5 * w, *csz[qa], nax[qa]);
  }
  // FULL_MAP (:226-270): a window is used when the fluid cells next to the
  // inlet at its four corners are all fluid
  std::vector<uint8_t> full((size_t)nw * nh, 0);
  int occ = 0;
  for (int h = 0; h < nh; h++)
    for (int k = 0; k < nw; k++) {
      bool ok = true;
      for (int c = 0; c < 4 && ok; c++) {
        int ijk[3];
        ijk[ax] = jcell;
        ijk[pa] = cell_intersect(amin[pa], mp[k] + ((c & 1) ? w : 0.0), *csz[pa], nax[pa]);
        ijk[qa] = cell_intersect(amin[qa], mq[h] + ((c & 2) ? w : 0.0), *csz[qa], nax[qa]);
        ok = ijk[0] >= 1 && ijk[1] >= 1 && ijk[2] >= 1 &&
             S.fluid[fluid_ijk(S, ijk[0], ijk[1], ijk[2])];
      }
      full[(size_t)h * nw + k] = ok;
      occ += ok;
    }
  if (occ == 0) return MFIXDEM_ERR_ARG;
  B.occupants = occ;
  // RAND_MAP (:272-300): a random order of the windows, drawn by rejection
  std::vector<int> order((size_t)occ + 1, 0);
  for (int filled = 0; filled < occ;) {
    const int t = (int)std::ceil((float)(D.uniform() * (double)occ));
    bool seen = t < 1;  // a draw of exactly 0 is drawn again
    for (int k = 1; k <= filled && !seen; k++) seen = order[k] == t;
    if (!seen) order[++filled] = t;
  }
  B.vacancy = 1;
  B.wp.assign((size_t)occ + 1, 0.0);
  B.wq.assign((size_t)occ + 1, 0.0);
  B.wi.assign((size_t)occ + 1, 0);
  B.wj.assign((size_t)occ + 1, 0);
  for (int h = 0, used = 0; h < nh; h++)
    for (int k = 0; k < nw; k++) {
      if (!full[(size_t)h * nw + k]) continue;
      const int slot = order[++used];
      B.wp[slot] = mp[k];
      B.wq[slot] = mq[h];
      B.wi[slot] = mi[k];
      B.wj[slot] = mj[h];
    }
  // particles per second of each phase, then per step (set_bc_dem_mi.f:
  // 104-175)
  const double area = plen * qlen;
  const double* bv[3] = {in.u_s, in.v_s, in.w_s};
  double npm[kMaxPh] = {0.0, 0.0, 0.0, 0.0}, max_vel = 0.0;
  for (int m : ph) {
    const double vn = std::fabs(bv[ax][m]);
    max_vel = std::max(max_vel, vn);
    const double d = p.d_p0[m];
    const double pv = (kPi * (d * d * d)) / 6.0;
    if (in.massflow_s[m] == 0.0) npm[m] = 0.0;
    else if (in.massflow_s[m] != kUndef) npm[m] = in.massflow_s[m] / (p.ro_s0[m] * pv);
    else npm[m] = vn * area * in.ep_s[m] / pv;
  }
  double nps = 0.0;
  for (int m = 0; m < kMaxPh; m++) nps += npm[m];
  const double npdt = nps * dt;
  if (npdt == 0.0) {
    B.pi_count = 0;
    B.pi_factor = 987654321.0;  // UNDEFINED_I
  } else if (npdt < 1.0) {
    B.pi_count = 1;
    B.pi_factor = (double)(float)(1.0 / npdt);
  } else {
    B.pi_count = (int)std::ceil((float)npdt);
    B.pi_factor = 1.0;
  }
  // Error 1110: the inlet velocity must clear a seeded particle out of its
  // window before the window is used again
  if (B.pi_count > 0 &&
      max_vel < max_dia / (dt * B.pi_factor * std::floor((float)occ / (float)B.pi_count)))
    return MFIXDEM_ERR_ARG;
  // every phase enters with the largest normal speed (:184-205)
  for (int m = 0; m < kMaxPh; m++)
    for (int c = 0; c < 3; c++) B.vel[m][c] = bv[c][m];
  for (int m : ph)
    B.vel[m][ax] = in.massflow_s[m] == 0.0 ? 0.0 : (B.low ? std::fabs(max_vel) : -std::fabs(max_vel));
  // DEM_BC_POLY_LAYOUT (:240-255): entries per phase in proportion to its rate
  B.layout.assign(kNumFracLimit, ph[0]);
  if (B.poly) {
    int first = 1;
    for (size_t k = 0; k + 1 < ph.size(); k++) {
      const double share = (double)kNumFracLimit * (npm[ph[k]] / nps);
      const int last = (int)std::floor(share) + (first - 1);
      for (int r = first; r <= last; r++) B.layout[r - 1] = ph[k];
      first = last + 1;
    }
    for (int r = first; r <= kNumFracLimit; r++) B.layout[r - 1] = ph.back();
  }
  B.mi_time = B.pi_factor * dt;
  return MFIXDEM_OK;
}

}  // namespace bcs_detail

// Set-up of every boundary; D.rng holds the seed on entry.
inline int32_t setup_dem_bcs(const Setup& S, DemBoundaries& D) {
  using namespace bcs_detail;
  const mfixdem_params& p = S.p;
  const double dt = S.d.dtsolid;
  const std::vector<double>* faces[3] = {&S.xe, &S.yn, &S.zt};
  const int nax[3] = {p.imax, p.jmax, p.kmax};
  double min_radius = 1e300;
  for (int m = 0; m < p.nphases; m++)
    if (p.d_p0[m] > 0.0) min_radius = std::min(min_radius, 0.5 * p.d_p0[m]);
  D.max_seeds = 0;
  for (DemBoundary& B : D.b) {
    const mfixdem_dem_bc& in = B.in;
    axes(in.plane, B.ax, B.pa, B.qa, B.low);
    const int ax = B.ax;
    const double lo[3] = {in.x_w, in.y_s, in.z_b}, hi[3] = {in.x_e, in.y_n, in.z_t};
    B.face = B.low ? (*faces[ax])[1] : (*faces[ax])[nax[ax] + 1];
    if (in.type == MFIXDEM_BC_MASS_INFLOW) {
      // DEM_BCMI_IJK: the rectangle and a cell around it, and along the
      // normal the cells a particle's centre can be in while it enters
      for (int a = 0; a < 3; a++) {
        B.lo[a] = std::max(dg_of_bcpos(S, a, lo[a]) - 1, 2);
        B.hi[a] = std::min(dg_of_bcpos(S, a, hi[a]) + 1, dg_n(S, a) - 1);
      }
      B.lo[ax] = dg_of_pos(S, ax, lo[ax] + (B.low ? min_radius : -p.max_radius));
      B.hi[ax] = dg_of_pos(S, ax, lo[ax] + (B.low ? p.max_radius : -min_radius));
      const int32_t rc = layout_inlet(S, dt, D, B);
      if (rc != MFIXDEM_OK) return rc;
      D.max_seeds += B.pi_count;
    } else {
      for (int a = 0; a < 3; a++) {  // DEM_BCMO_IJK
        B.lo[a] = dg_of_bcpos(S, a, lo[a]);
        B.hi[a] = dg_of_bcpos(S, a, hi[a]) + 1;
      }
    }
    for (int a = 0; a < 3; a++) {  // cells on this rank with one ghost layer
      B.lo[a] = std::max(B.lo[a], 1);
      B.hi[a] = std::min(B.hi[a], dg_n(S, a));
    }
  }
  return MFIXDEM_OK;
}

// The device form of the boundaries.
inline std::vector<BcDev> bcs_dev(const Setup& S, const DemBoundaries& D) {
  std::vector<BcDev> v(D.b.size());
  for (size_t k = 0; k < D.b.size(); k++) {
    const DemBoundary& B = D.b[k];
    BcDev& o = v[k];
    o = BcDev{};
    o.type = B.in.type;
    o.ax = B.ax;
    o.low = B.low ? 1 : 0;
    for (int a = 0; a < 3; a++) {
      o.lo[a] = B.lo[a];
      o.hi[a] = B.hi[a];
    }
    o.dg_ni = S.dg_ni;
    o.dg_nj = S.dg_nj;
    o.face = B.face;
    for (int m = 0; m < kMaxPh; m++) {
      o.vel[m][0] = B.in.u_s[m];
      o.vel[m][1] = B.in.v_s[m];
      o.vel[m][2] = B.in.w_s[m];
    }
  }
  return v;
}

// The seeding part of MASS_INFLOW_DEM for the step at D.s_time: fills
// D.seeds, taking global ids from D.free_gids.
inline int32_t plan_seeds(const Setup& S, DemBoundaries& D) {
  using namespace bcs_detail;
  const mfixdem_params& p = S.p;
  const double dt = S.d.dtsolid;
  D.seeds.clear();
  for (DemBoundary& B : D.b) {
    if (B.in.type != MFIXDEM_BC_MASS_INFLOW) continue;
    double rnd[3];
    for (double& r : rnd) r = D.uniform();  // drawn every step
    if (B.mi_time > D.s_time) continue;
    if (D.s_time < B.in.mi_start_time || D.s_time > B.in.mi_end_time) continue;
    for (int ip = 0; ip < B.pi_count; ip++) {
      const int vac = B.vacancy;
      B.vacancy = vac % B.occupants + 1;
      int m = B.layout[0];
      if (B.poly) {
        const int r = (int)std::ceil((double)kNumFracLimit * rnd[0]);
        m = B.layout[(r < 1 ? 1 : r) - 1];
      }
      const double d = p.d_p0[m];
      SeedDev sd{};
      sd.pos[B.pa] = B.wp[vac] + (0.5 * d + (B.window - d) * rnd[0]);
      sd.pos[B.qa] = B.wq[vac] + (0.5 * d + (B.window - d) * rnd[1]);
      sd.pos[B.ax] = B.offset;
      int ijk[3];
      ijk[B.pa] = B.wi[vac];
      ijk[B.qa] = B.wj[vac];
      ijk[B.ax] = B.lcell;
      if (D.free_gids.empty()) return MFIXDEM_ERR_CAPACITY;  // the reference grows MAX_PIP
      sd.gid = *D.free_gids.begin();
      D.free_gids.erase(D.free_gids.begin());
      for (int c = 0; c < 3; c++) sd.vel[c] = B.vel[m][c];
      const double r = d * 0.5;
      sd.radius = r;
      sd.pvol = (4.0 / 3.0) * kPi * (r * r * r);
      sd.pmass = sd.pvol * p.ro_s0[m];
      sd.omoi = 2.5 / (sd.pmass * (r * r));
      sd.phase = m;
      sd.fcell = pack_ijk(ijk[0], ijk[1], ijk[2]);
      int c3[3];
      for (int a = 0; a < 3; a++)
        c3[a] = std::min(dg_n(S, a), std::max(1, dg_of_pos(S, a, sd.pos[a])));
      sd.dgc = dg_ijk(S, c3[0], c3[1], c3[2]);
      D.seeds.push_back(sd);
    }
    B.mi_time = B.mi_time + B.pi_factor * dt;
  }
  return MFIXDEM_OK;
}

}  // namespace mfixdem

// params.hpp -- device parameter blocks (DevParams, GridDev) from a
// validated Setup; shared by the single-system and the ensemble handles.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host_setup.hpp"
#include "kernels.hpp"

namespace mfixdem {

// Device form of the facets of a set-up.
inline FacetDev facet_dev(const Facet& f) {
  FacetDev F{};
  std::memcpy(F.v, f.v, sizeof(F.v));
  std::memcpy(F.n, f.n, sizeof(F.n));
  F.fmin = f.fmin;
  F.fmax = f.fmax;
  std::memcpy(F.vt, f.vt, sizeof(F.vt));
  std::memcpy(F.vn, f.vn, sizeof(F.vn));
  F.dir = f.dir;
  F.ignore_edges = 0;
  for (int e = 0; e < 4; e++) F.ignore_edges |= f.ignore_edge[e] ? 1 << e : 0;
  return F;
}
inline void facets_dev(const Setup& S, std::vector<FacetDev>& out) {
  out.clear();
  for (const Facet& f : S.facets) out.push_back(facet_dev(f));
}
// Walls of at most kWallSlots facets travel in the kernel arguments, and
// slot k of a particle then holds facet k: with no more facets than slots a
// free slot always exists, so GET_COLLISION never evicts and the history
// each facet sees is the same as with the reference's slot search.
inline bool walls_small(const Setup& S) { return S.facets.size() <= (size_t)kWallSlots; }
inline void small_facets(const Setup& S, DevParams& P) {
  P.nfs = 0;
  for (int f = 0; f < kWallSlots; f++) P.facets[f] = FacetDev{};
  if (!walls_small(S)) return;
  P.nfs = (int)S.facets.size();
  for (int f = 0; f < P.nfs; f++) P.facets[f] = facet_dev(S.facets[f]);
}

inline void fill_params(const Setup& S, int n, int stride, DevParams& P, GridDev& G) {
  const mfixdem_params& p = S.p;
  P = DevParams{};
  P.n = n;
  P.stride = stride;
  P.hertz = p.coll_model == MFIXDEM_HERTZ;
  P.ab2 = p.integrator == MFIXDEM_ADAMS_BASHFORTH;
  P.coupled = p.coupled;
  P.explicit_cpl = p.des_explicitly_coupled;
  P.drag_type = p.drag_type;
  P.nph = p.nphases;
  P.dt = S.d.dtsolid;
  for (int c = 0; c < 3; c++) P.grav[c] = p.grav[c];
  P.kn = p.kn; P.kt = S.d.kt; P.kn_w = p.kn_w; P.kt_w = S.d.kt_w;
  P.mew = p.mew; P.mew_w = p.mew_w;
  for (int a = 0; a < kMaxPh; a++) {
    for (int b = 0; b < kMaxPh; b++) {
      P.etan[a][b] = S.d.des_etan[a][b];
      P.etat[a][b] = S.d.des_etat[a][b];
      P.hkn[a][b] = S.d.hert_kn[a][b];
      P.hkt[a][b] = S.d.hert_kt[a][b];
    }
    P.etan_w[a] = S.d.des_etan_wall[a];
    P.etat_w[a] = S.d.des_etat_wall[a];
    P.hkwn[a] = S.d.hert_kwn[a];
    P.hkwt[a] = S.d.hert_kwt[a];
  }
  P.ni = S.ni; P.nj = S.nj; P.nk = S.nk; P.ncell = S.ncell;
  P.ro_g0 = p.ro_g0; P.mu_g0 = p.mu_g0;
  P.interp = p.interp_scheme;
  P.interp_on = p.des_interp_on;
  if (p.interp_scheme == MFIXDEM_INTERP_SQUARE_DPVM ||
      p.interp_scheme == MFIXDEM_INTERP_GAUSS_DPVM) {
    // set_filter_des.f:85,99-100
    P.fw = 0.500 * p.des_interp_width;
    P.oofv = 0.25 / (P.fw * P.fw * P.fw);
    P.fw3 = P.fw * 3;
  }
  G.dg_ni = S.dg_ni; G.dg_nj = S.dg_nj; G.dg_nk = S.dg_nk;
  for (int d = 0; d < 3; d++) G.coarse[d] = S.coarse[d];
  G.dg_lo[0] = S.d.dg_xstart; G.dg_lo[1] = S.d.dg_ystart; G.dg_lo[2] = S.d.dg_zstart;
  G.dg_inv[0] = S.d.dg_dxinv; G.dg_inv[1] = S.d.dg_dyinv; G.dg_inv[2] = S.d.dg_dzinv;
  G.factor_rlm = p.factor_rlm;
  small_facets(S, P);
}

// bits b per axis of a Morton code over cells 0..count-1 with 2^b > count
inline int morton_axis_bits(int count) {
  int b = 1;
  while ((1 << b) <= count) b++;
  return b;
}

// Neighbour-search grid of one system: cubic cells of cell_scale (>= 1)
// times the largest cut-off FACTOR_RLM*(2 r_max), one cell of margin around
// the box.  The Morton code of a cell takes mbits = 3 b bits of the sort
// key, with 2^b > every axis count, so no cell's code is all ones (the key
// of an empty slot).
inline int32_t search_grid(const mfixdem_params& p, int64_t n, GridDev& G, int& n_fine,
                           double cell_scale = 1.0) {
  const double hcut = p.factor_rlm * (2.0 * p.max_radius) * (1.0 + 1e-9) * cell_scale;
  const double lo[3] = {p.x_min - hcut, p.y_min - hcut, p.z_min - hcut};
  const double len[3] = {p.xlength, p.ylength, p.zlength};
  for (int d = 0; d < 3; d++) {
    G.f_lo[d] = lo[d];
    G.fn[d] = std::max(1, (int)std::ceil((len[d] + 2.0 * hcut) / hcut));
    if (G.fn[d] > 1022) return MFIXDEM_ERR_UNSUPPORTED;
  }
  G.f_inv = 1.0 / hcut;
  n_fine = G.fn[0] * G.fn[1] * G.fn[2];
  // prefilter of the row walk: single-precision coordinates about the
  // grid's centre, good to 2^-24 of the half extent each; a computed
  // distance is then within 2^-20 of it (generously) of the true one
  double half = 0.0;
  for (int d = 0; d < 3; d++) {
    G.pf_c[d] = lo[d] + 0.5 * G.fn[d] * hcut;
    half = std::max(half, 0.5 * G.fn[d] * hcut);
  }
  G.pf_xmax = (float)half;
  G.pf_slack = (float)std::ldexp((double)G.pf_xmax, -20);
  (void)n;
  G.mbits = 3 * morton_axis_bits(std::max(G.fn[0], std::max(G.fn[1], G.fn[2])));
  return MFIXDEM_OK;
}

}  // namespace mfixdem

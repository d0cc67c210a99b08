// device_physics.hpp -- per-pair, per-wall, drag and integrator arithmetic of
// the DEM step for gfx950.
//
// The FP64 functions keep the FP64 expression order of the reference lines
// they cite (the library is compiled with -ffp-contract=off, so no FMA is
// formed and a serial CPU restatement of the same lines agrees bit for bit,
// except where libm pow enters the drag law).  The MIXED variants keep
// positions, velocities, the tangential history and the integration in FP64
// and do the contact-law and drag arithmetic in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mfixdem {

constexpr int kMaxPh = 4;
// wall-collision slots per particle: COLLISION_ARRAY_MAX (discretelement_mod.f:637)
constexpr int kWallSlots = 8;
constexpr double kSmall = 1.0e-15;   // SMALL_NUMBER, param1_mod.f:26
constexpr double kLarge = 1.0e32;    // LARGE_NUMBER, param1_mod.f:25
constexpr double kUndefD = 9.87654321e31;

struct FacetDev {
  double v[3][3];
  double n[3];
  double fmin, fmax;
  double vt[3], vn[3];  // STL_TANGENTIAL_VEL, STL_NORMAL_VEL
  int dir;
  int ignore_edges;  // bit e set: contacts on edge e are ignored
};

// Everything a step kernel needs besides array pointers; passed by value.
struct DevParams {
  int n;            // particle slots
  int stride;       // leading dimension of SoA arrays
  int hertz, ab2, coupled, drag_type, nph;
  int first_pass;   // Adams-Bashforth first step
  int store_forces; // debug: keep FC/TOW
  int fuse_dep;     // fused step: GARG_2012 deposition in the step kernel
  int explicit_cpl; // DES_EXPLICITLY_COUPLED: FC += DRAG_FC, no rebinning
  double dt;
  double grav[3];
  double kn, kt, kn_w, kt_w, mew, mew_w;
  double etan[kMaxPh][kMaxPh], etat[kMaxPh][kMaxPh];
  double hkn[kMaxPh][kMaxPh], hkt[kMaxPh][kMaxPh];
  double etan_w[kMaxPh], etat_w[kMaxPh], hkwn[kMaxPh], hkwt[kMaxPh];
  // fluid grid
  int ni, nj, nk, ncell;
  double ro_g0, mu_g0;
  double dep_scale[kMaxPh];  // fixed-point scale of the deposition sums (power of 2)
  // filter schemes (DES_INTERP_SCHEME other than GARG_2012)
  int interp, interp_on;     // MFIXDEM_INTERP_*, DES_INTERP_ON
  double fw, fw3, oofv;      // FILTER_WIDTH_INTERP, x3, OoFILTER_VOL (DPVM)
  // walls of at most kWallSlots facets: the facets themselves (slot k of a
  // particle holds facet k); nfs = 0 when there are more (facet table and
  // per-cell lists in device memory)
  int nfs;
  FacetDev facets[kWallSlots];
  // DEM mass inlets / outlets (StepPtrs::bcs): 0 when the run has none
  int nbc;
};

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double x[3], const double y[3], double c[3]) {
  c[0] = x[1] * y[2] - x[2] * y[1];
  c[1] = x[2] * y[0] - x[0] * y[2];
  c[2] = x[0] * y[1] - x[1] * y[0];
}
__device__ __forceinline__ float dot3f(const float a[3], const float b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3f(const float x[3], const float y[3], float c[3]) {
  c[0] = x[1] * y[2] - x[2] * y[1];
  c[1] = x[2] * y[0] - x[0] * y[2];
  c[2] = x[0] * y[1] - x[1] * y[0];
}
// FP32 helpers for the mixed path: single-instruction sqrt / reciprocal.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

struct PairIn {
  double xL[3], xI[3], vL[3], vI[3], wL[3], wI[3];
  double rL, rI, mL, mI;
  int phL, phI;
};
struct PairOut {
  double fct[3];    // FN + FT on L (I receives -fct)
  double tl[3];     // torque increment of L (DIST_CL * n x FT)
  double ti[3];     // torque increment of I (DIST_CI * n x FT)
  double pft[3];    // new tangential history
};

// CALC_FORCE_DEM pair body (calc_force_dem.f:210-412) with CFRELVEL
// (cfrelvel.f:54-77).  Returns 0 no contact, 1 contact, 2 zero distance.
__device__ __forceinline__ int pair_force_fp64(const DevParams& P, const PairIn& a,
                                               const double pft_in[3], PairOut& o) {
  const double r_lm = a.rL + a.rI;
  double dist[3] = {a.xI[0] - a.xL[0], a.xI[1] - a.xL[1], a.xI[2] - a.xL[2]};
  double dm = dot3(dist, dist);
  const double rs = r_lm - kSmall;
  if (dm > rs * rs) return 0;
  if (dm == 0.0) return 2;
  dm = sqrt(dm);
  double nrm[3];
  for (int c = 0; c < 3; c++) nrm[c] = dist[c] / dm;
  const double ovl = r_lm - dm;
  double vrel[3];
  for (int c = 0; c < 3; c++) vrel[c] = a.vL[c] - a.vI[c];
  const double dcl = (dm * dm + a.rL * a.rL - a.rI * a.rI) / (2.0 * dm);
  const double dci = dm - dcl;
  double osum[3], vrot[3];
  for (int c = 0; c < 3; c++) osum[c] = a.wL[c] * dcl + a.wI[c] * dci;
  cross3(osum, nrm, vrot);
  for (int c = 0; c < 3; c++) vrel[c] = vrel[c] + vrot[c];
  const double vrn = dot3(vrel, nrm);
  double vs[3];
  for (int c = 0; c < 3; c++) vs[c] = vrel[c] - vrn * nrm[c];
  const double reff = (a.rL * a.rI) / (a.rL + a.rI);
  const double meff = (a.mL * a.mI) / (a.mL + a.mI);
  double kn, kt, etan, etat;
  if (P.hertz) {
    double sqo = sqrt(ovl);
    kn = P.hkn[a.phL][a.phI] * sqo * sqrt(reff);
    kt = P.hkt[a.phL][a.phI] * sqo * sqrt(reff);
    sqo = sqrt(sqo);
    etan = P.etan[a.phL][a.phI] * sqo * sqrt(P.hkn[a.phL][a.phI] * sqrt(reff) * meff);
    etat = P.etat[a.phL][a.phI] * sqo * sqrt(P.hkt[a.phL][a.phI] * sqrt(reff) * meff);
  } else {
    kn = P.kn;
    kt = P.kt;
    etan = P.etan[a.phL][a.phI] * sqrt(meff);
    etat = P.etat[a.phL][a.phI] * sqrt(meff);
  }
  double fn[3], ot[3], ft[3];
  for (int c = 0; c < 3; c++) fn[c] = -(kn * ovl * nrm[c] + etan * vrn * nrm[c]);
  for (int c = 0; c < 3; c++) ot[c] = P.dt * vs[c] + pft_in[c];
  const double mag = sqrt(dot3(ot, ot));
  if (mag > 0.0) {
    for (int c = 0; c < 3; c++) ft[c] = -kt * ot[c] - etat * vs[c];
    const double ftmd = sqrt(dot3(ft, ft));
    const double fnmd = P.mew * sqrt(dot3(fn, fn));
    if (ftmd > fnmd) {
      double tg[3];
      for (int c = 0; c < 3; c++) tg[c] = ot[c] / mag;
      for (int c = 0; c < 3; c++) ft[c] = -fnmd * tg[c];
      for (int c = 0; c < 3; c++) ot[c] = (fnmd / kt) * tg[c];
    }
  } else {
    ft[0] = ft[1] = ft[2] = 0.0;
  }
  for (int c = 0; c < 3; c++) o.pft[c] = ot[c];
  const double dist_cl = dm / 2.0 + (a.rL * a.rL - a.rI * a.rI) / (2.0 * dm);
  const double dist_ci = dm - dist_cl;
  double tf[3];
  cross3(nrm, ft, tf);
  for (int c = 0; c < 3; c++) o.fct[c] = 0.0 + fn[c] + ft[c];
  for (int c = 0; c < 3; c++) o.tl[c] = dist_cl * tf[c];
  for (int c = 0; c < 3; c++) o.ti[c] = dist_ci * tf[c];
  return 1;
}

// Mixed precision: contact geometry test in FP64 (same contact set as the
// FP64 path), everything after it in FP32 except the history update.
__device__ __forceinline__ int pair_force_mixed(const DevParams& P, const PairIn& a,
                                                const double pft_in[3], PairOut& o) {
  const double r_lm = a.rL + a.rI;
  const double dist[3] = {a.xI[0] - a.xL[0], a.xI[1] - a.xL[1], a.xI[2] - a.xL[2]};
  const double dm2 = dot3(dist, dist);
  const double rs = r_lm - kSmall;
  if (dm2 > rs * rs) return 0;
  if (dm2 == 0.0) return 2;
  const float dmf = fsqrt((float)dm2);
  const float inv = __builtin_amdgcn_rcpf(dmf);
  float nrm[3];
  for (int c = 0; c < 3; c++) nrm[c] = (float)dist[c] * inv;
  // overlap from the FP64 squared distance: r_lm - |d| = (r_lm^2 - |d|^2) / (r_lm + |d|)
  const float ovl = (float)((r_lm * r_lm - dm2)) * __builtin_amdgcn_rcpf((float)r_lm + dmf);
  const float rL = (float)a.rL, rI = (float)a.rI;
  float vrel[3];
  for (int c = 0; c < 3; c++) vrel[c] = (float)(a.vL[c] - a.vI[c]);
  const float dcl = (dmf * dmf + rL * rL - rI * rI) * (0.5f * inv);
  const float dci = dmf - dcl;
  float osum[3], vrot[3];
  for (int c = 0; c < 3; c++) osum[c] = (float)a.wL[c] * dcl + (float)a.wI[c] * dci;
  cross3f(osum, nrm, vrot);
  for (int c = 0; c < 3; c++) vrel[c] += vrot[c];
  const float vrn = dot3f(vrel, nrm);
  float vs[3];
  for (int c = 0; c < 3; c++) vs[c] = vrel[c] - vrn * nrm[c];
  const float reff = fdiv(rL * rI, rL + rI);
  const float mL = (float)a.mL, mI = (float)a.mI;
  const float meff = fdiv(mL * mI, mL + mI);
  float kn, kt, etan, etat;
  if (P.hertz) {
    const float sr = fsqrt(reff);
    float sqo = fsqrt(ovl);
    const float hkn = (float)P.hkn[a.phL][a.phI], hkt = (float)P.hkt[a.phL][a.phI];
    kn = hkn * sqo * sr;
    kt = hkt * sqo * sr;
    sqo = fsqrt(sqo);
    etan = (float)P.etan[a.phL][a.phI] * sqo * fsqrt(hkn * sr * meff);
    etat = (float)P.etat[a.phL][a.phI] * sqo * fsqrt(hkt * sr * meff);
  } else {
    const float sm = fsqrt(meff);
    kn = (float)P.kn;
    kt = (float)P.kt;
    etan = (float)P.etan[a.phL][a.phI] * sm;
    etat = (float)P.etat[a.phL][a.phI] * sm;
  }
  float fn[3], ft[3];
  for (int c = 0; c < 3; c++) fn[c] = -(kn * ovl + etan * vrn) * nrm[c];
  double ot[3];
  for (int c = 0; c < 3; c++) ot[c] = P.dt * (double)vs[c] + pft_in[c];
  const double mag2 = dot3(ot, ot);
  if (mag2 > 0.0) {
    float otf[3] = {(float)ot[0], (float)ot[1], (float)ot[2]};
    for (int c = 0; c < 3; c++) ft[c] = -kt * otf[c] - etat * vs[c];
    const float ftmd = fsqrt(dot3f(ft, ft));
    const float fnmd = (float)P.mew * fsqrt(dot3f(fn, fn));
    if (ftmd > fnmd) {
      const double mag = sqrt(mag2);
      double tg[3];
      for (int c = 0; c < 3; c++) tg[c] = ot[c] / mag;
      for (int c = 0; c < 3; c++) ft[c] = -fnmd * (float)tg[c];
      const double sc = (double)fnmd / (double)kt;
      for (int c = 0; c < 3; c++) ot[c] = sc * tg[c];
    }
  } else {
    ft[0] = ft[1] = ft[2] = 0.0f;
  }
  for (int c = 0; c < 3; c++) o.pft[c] = ot[c];
  const float dist_cl = 0.5f * dmf + (rL * rL - rI * rI) * (0.5f * inv);
  const float dist_ci = dmf - dist_cl;
  float tf[3];
  cross3f(nrm, ft, tf);
  for (int c = 0; c < 3; c++) o.fct[c] = (double)(fn[c] + ft[c]);
  for (int c = 0; c < 3; c++) o.tl[c] = (double)(dist_cl * tf[c]);
  for (int c = 0; c < 3; c++) o.ti[c] = (double)(dist_ci * tf[c]);
  return 1;
}

// ClosestPtPointTriangle, stl_functions_des_mod.f:25-160.
__device__ __forceinline__ void closest_pt_tri(const FacetDev& F, const double pp[3],
                                               double cp[3], int* pe, int* edge) {
  double ab[3], ac[3], ap[3];
  *pe = 0;
  *edge = 0;
  for (int i = 0; i < 3; i++) {
    ab[i] = F.v[1][i] - F.v[0][i];
    ac[i] = F.v[2][i] - F.v[0][i];
    ap[i] = pp[i] - F.v[0][i];
  }
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 < 0.0 && d2 < 0.0) {
    for (int i = 0; i < 3; i++) cp[i] = F.v[0][i];
    *pe = 1; *edge = 1;
    if ((F.ignore_edges >> 1) & 1) cp[0] = kUndefD;
    return;
  }
  double bp[3];
  for (int i = 0; i < 3; i++) bp[i] = pp[i] - F.v[1][i];
  const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 > 0.0 && d4 < d3) {
    for (int i = 0; i < 3; i++) cp[i] = F.v[1][i];
    *pe = 1; *edge = 1;
    if ((F.ignore_edges >> 1) & 1) cp[0] = kUndefD;
    return;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc < 0.0 && d1 > 0.0 && d3 < 0.0) {
    const double v = d1 / (d1 - d3);
    for (int i = 0; i < 3; i++) cp[i] = F.v[0][i] + v * ab[i];
    *pe = 1; *edge = 1;
    if ((F.ignore_edges >> 1) & 1) cp[0] = kUndefD;
    return;
  }
  double cpv[3];
  for (int i = 0; i < 3; i++) cpv[i] = pp[i] - F.v[2][i];
  const double d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
  if (d6 > 0.0 && d5 < d6) {
    for (int i = 0; i < 3; i++) cp[i] = F.v[2][i];
    *pe = 1; *edge = 2;
    if ((F.ignore_edges >> 2) & 1) cp[0] = kUndefD;
    return;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb < 0.0 && d2 > 0.0 && d6 < 0.0) {
    const double w = d2 / (d2 - d6);
    for (int i = 0; i < 3; i++) cp[i] = F.v[0][i] + w * ac[i];
    *pe = 1; *edge = 2;
    if ((F.ignore_edges >> 2) & 1) cp[0] = kUndefD;
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va < 0.0 && (d4 - d3) > 0.0 && (d5 - d6) > 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; i++) cp[i] = F.v[1][i] + w * (F.v[2][i] - F.v[1][i]);
    *pe = 1; *edge = 3;
    if ((F.ignore_edges >> 3) & 1) cp[0] = kUndefD;
    return;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  for (int i = 0; i < 3; i++) cp[i] = F.v[0][i] + ab[i] * v + ac[i] * w;
}

// One particle-facet contact of CALC_DEM_FORCE_WITH_WALL_STL
// (calc_collision_wall_mod.f:300-606).  hist/hist_set: the facet's stored
// tangential history.  Returns: 0 = removed (history cleared), 1 = contact
// (fc/tow updated, hist written), 2 = skipped (edge contact ignored).
template <bool MIXED>
__device__ __forceinline__ int wall_contact(const DevParams& P, const FacetDev& F,
                                            const double pos[3], const double vel[3],
                                            const double om[3], double r, double pmass,
                                            int ph, double hist[3], bool hist_set,
                                            double fc[3], double tow[3]) {
  // pos[F.dir] by selects: no dynamic index into an array in registers
  const double pd = F.dir == 0 ? pos[0] : (F.dir == 1 ? pos[1] : pos[2]);
  const double pmax = pd + r, pmin = pd - r;
  if (F.fmin > pmax) return 0;
  if (F.fmax < pmin) return 0;
  const double vv[3] = {F.v[0][0] - pos[0], F.v[0][1] - pos[1], F.v[0][2] - pos[2]};
  const double line_t = dot3(vv, F.n);
  if (line_t <= -1.0001 * r) return 0;
  double cp[3];
  int pe, edge;
  closest_pt_tri(F, pos, cp, &pe, &edge);
  const double dist[3] = {cp[0] - pos[0], cp[1] - pos[1], cp[2] - pos[2]};
  const double distsq = dot3(dist, dist);
  if (distsq >= r * r - kSmall) return 0;
  if (pe && ((F.ignore_edges >> edge) & 1)) return 2;
  const double distmod = sqrt(distsq);
  double nrm[3] = {dist[0], dist[1], dist[2]};
  if (distsq > 0.0) {
    const double sq = sqrt(distsq);
    for (int c = 0; c < 3; c++) nrm[c] = dist[c] / sq;
  }
  const double ovl = r - distmod;
  double h0[3] = {0.0, 0.0, 0.0};
  if (hist_set) { h0[0] = hist[0]; h0[1] = hist[1]; h0[2] = hist[2]; }
  if (!MIXED) {
    double cr[3], vrot[3], vrel[3], vrt[3];
    cross3(om, nrm, cr);
    for (int c = 0; c < 3; c++) vrot[c] = distmod * cr[c];
    for (int c = 0; c < 3; c++) {  // CFRELVEL_WALL, calc_collision_wall_mod.f:836-837
      vrel[c] = vel[c] + vrot[c];
      vrel[c] = vrel[c] - (F.vt[c] + F.vn[c]);
    }
    const double vrn = dot3(vrel, nrm);
    for (int c = 0; c < 3; c++) vrt[c] = vrel[c] - vrn * nrm[c];
    double kn, kt, etan, etat;
    if (P.hertz) {
      double sqo = 0.0;
      if (ovl > 0.0) sqo = sqrt(ovl);
      kn = P.hkwn[ph] * sqo * sqrt(r);
      kt = P.hkwt[ph] * sqo * sqrt(r);
      etan = P.etan_w[ph] * sqrt(sqo) * sqrt(P.hkwn[ph] * sqrt(r) * pmass);
      etat = P.etat_w[ph] * sqrt(sqo) * sqrt(P.hkwt[ph] * sqrt(r) * pmass);
    } else {
      kn = P.kn_w;
      kt = P.kt_w;
      etan = P.etan_w[ph] * sqrt(pmass);
      etat = P.etat_w[ph] * sqrt(pmass);
    }
    double fnorm[3], ot[3], ftan[3];
    for (int c = 0; c < 3; c++) fnorm[c] = -(kn * ovl * nrm[c] + etan * vrn * nrm[c]);
    for (int c = 0; c < 3; c++) ot[c] = P.dt * vrt[c] + h0[c];
    const double mag = sqrt(dot3(ot, ot));
    if (mag > 0.0) {
      for (int c = 0; c < 3; c++) ftan[c] = -kt * ot[c] - etat * vrt[c];
      const double ftmd = sqrt(dot3(ftan, ftan));
      const double fnmd = P.mew_w * sqrt(dot3(fnorm, fnorm));
      if (ftmd > fnmd) {
        double tg[3];
        for (int c = 0; c < 3; c++) tg[c] = ot[c] / mag;
        for (int c = 0; c < 3; c++) ftan[c] = -fnmd * tg[c];
        for (int c = 0; c < 3; c++) ot[c] = (fnmd / kt) * tg[c];
      }
    } else {
      ftan[0] = ftan[1] = ftan[2] = 0.0;
    }
    for (int c = 0; c < 3; c++) hist[c] = ot[c];
    for (int c = 0; c < 3; c++) fc[c] = fc[c] + fnorm[c] + ftan[c];
    double tw[3];
    cross3(nrm, ftan, tw);
    for (int c = 0; c < 3; c++) tow[c] = tow[c] + distmod * tw[c];
  } else {
    const float nf[3] = {(float)nrm[0], (float)nrm[1], (float)nrm[2]};
    const float dmf = (float)distmod, ovf = (float)ovl;
    const float omf[3] = {(float)om[0], (float)om[1], (float)om[2]};
    float cr[3], vrel[3], vrt[3];
    cross3f(omf, nf, cr);
    for (int c = 0; c < 3; c++)
      vrel[c] = (float)vel[c] + dmf * cr[c] - (float)(F.vt[c] + F.vn[c]);
    const float vrn = dot3f(vrel, nf);
    for (int c = 0; c < 3; c++) vrt[c] = vrel[c] - vrn * nf[c];
    float kn, kt, etan, etat;
    const float rf = (float)r, mf = (float)pmass;
    if (P.hertz) {
      const float sqo = ovf > 0.0f ? fsqrt(ovf) : 0.0f;
      const float sr = fsqrt(rf), s4 = fsqrt(sqo);
      kn = (float)P.hkwn[ph] * sqo * sr;
      kt = (float)P.hkwt[ph] * sqo * sr;
      etan = (float)P.etan_w[ph] * s4 * fsqrt((float)P.hkwn[ph] * sr * mf);
      etat = (float)P.etat_w[ph] * s4 * fsqrt((float)P.hkwt[ph] * sr * mf);
    } else {
      const float sm = fsqrt(mf);
      kn = (float)P.kn_w;
      kt = (float)P.kt_w;
      etan = (float)P.etan_w[ph] * sm;
      etat = (float)P.etat_w[ph] * sm;
    }
    float fnorm[3], ftan[3];
    for (int c = 0; c < 3; c++) fnorm[c] = -(kn * ovf + etan * vrn) * nf[c];
    double ot[3];
    for (int c = 0; c < 3; c++) ot[c] = P.dt * (double)vrt[c] + h0[c];
    const double mag2 = dot3(ot, ot);
    if (mag2 > 0.0) {
      for (int c = 0; c < 3; c++) ftan[c] = -kt * (float)ot[c] - etat * vrt[c];
      const float ftmd = fsqrt(dot3f(ftan, ftan));
      const float fnmd = (float)P.mew_w * fsqrt(dot3f(fnorm, fnorm));
      if (ftmd > fnmd) {
        const double mag = sqrt(mag2);
        const double sc = (double)fnmd / (double)kt;
        for (int c = 0; c < 3; c++) {
          const double tg = ot[c] / mag;
          ftan[c] = -fnmd * (float)tg;
          ot[c] = sc * tg;
        }
      }
    } else {
      ftan[0] = ftan[1] = ftan[2] = 0.0f;
    }
    for (int c = 0; c < 3; c++) hist[c] = ot[c];
    for (int c = 0; c < 3; c++) fc[c] = fc[c] + (double)(fnorm[c] + ftan[c]);
    float tw[3];
    cross3f(nf, ftan, tw);
    for (int c = 0; c < 3; c++) tow[c] = tow[c] + (double)(dmf * tw[c]);
  }
  return 1;
}

// DES_DRAG_GP (drag_gp_des.f:29-260) with DRAG_GIDASPOW (drag_gs.f:536-598),
// DRAG_WEN_YU (drag_gs.f:703-758), C_DS_SN (drag_mod.f:48-56); Model A,
// UR_F_gs = 1.  Returns F_GP = beta * PVOL.
__device__ __forceinline__ double drag_fgp_fp64(int drag_type, double epg, double mu, double rog,
                                                double vrel, double dpm, double pvol) {
  const double ropg = rog * epg;
  double dga;
  const double re = (mu > 0.0) ? dpm * vrel * ropg / mu : kLarge;
  if (re == 0.0) {
    dga = 0.0;
  } else if (drag_type == 1) {
    double cd;
    if (re <= 1000.0) cd = (24.0 / (re + kSmall)) * (1.0 + 0.15 * pow(re, 0.687));
    else cd = 0.44;
    dga = 0.75 * cd * vrel * ropg * pow(epg, -2.65) / dpm;
  } else if (epg <= 0.8) {
    dga = 150.0 * (1.0 - epg) * mu / (epg * (dpm * dpm)) + 1.75 * rog * vrel / dpm;
  } else {
    double cd;
    if (re <= 1000.0) cd = 24.0 * (1.0 + 0.15 * pow(re, 0.687)) / (re + kSmall);
    else cd = 0.44;
    dga = 0.75 * cd * vrel * ropg * pow(epg, -2.65) / dpm;
  }
  return dga * pvol;
}

__device__ __forceinline__ float drag_fgp_fp32(int drag_type, float epg, float mu, float rog,
                                               float vrel, float dpm, float pvol) {
  const float ropg = rog * epg;
  float dga;
  const float re = (mu > 0.0f) ? fdiv(dpm * vrel * ropg, mu) : 1.0e30f;
  if (re == 0.0f) {
    dga = 0.0f;
  } else if (drag_type != 1 && epg <= 0.8f) {
    dga = fdiv(150.0f * (1.0f - epg) * mu, epg * (dpm * dpm)) + fdiv(1.75f * rog * vrel, dpm);
  } else {
    const float cd = (re <= 1000.0f)
                         ? fdiv(24.0f * (1.0f + 0.15f * __powf(re, 0.687f)), re + 1.0e-15f)
                         : 0.44f;
    dga = fdiv(0.75f * cd * vrel * ropg * __powf(epg, -2.65f), dpm);
  }
  return dga * pvol;
}

// Unique cell I with faces[I-1] <= x < faces[I], 1 <= I <= nmax; 0 if none
// (PARTICLES_IN_CELL + PIC_SEARCH, particles_in_cell.f:87-154,347-370).
__device__ __forceinline__ int locate_cell(const double* faces, int nmax, double inv_d, double x) {
  if (!(x >= faces[0] && x < faces[nmax])) return 0;
  int i = (int)((x - faces[0]) * inv_d) + 1;
  if (i < 1) i = 1;
  if (i > nmax) i = nmax;
  while (i > 1 && x < faces[i - 1]) i--;
  while (i < nmax && x >= faces[i]) i++;
  return i;
}

}  // namespace mfixdem

// kernels.hpp -- gfx950 kernels of the DEM time step.
//
// Layout in HBM (all per-particle arrays in the current Morton order of the
// neighbour-search grid; component c of particle i at [c*stride + i]):
//   pos/vel/omega      double, double-buffered (read one, write the other)
//   radius/pmass/omoi/pvol, gid, state, phase, dgc (DES cell), fcell (fluid
//   cell), wall mask + per-facet wall history, AB2 accelerations
//   neighbour rows     full rows: [incoming (owner has lower id) | own]
//                      nbr_start[i] .. nbr_split[i] .. nbr_start[i+1]
//                      nbr = neighbour slot, nbr_pid = pair id (history slot)
//   pair history       pft[pid*3 + d], double-buffered per step
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "../../include/mfixdem.h"
#include "device_physics.hpp"

namespace mfixdem {

constexpr int kBS = 256;

// PARTICLE_STATE values (discretelement_mod.f:68-78).  A particle seeded by
// an inlet since the last neighbour search is ENTERING but in no DES cell
// list yet: kSeeded marks it until the next binning turns it into kEntering,
// so that the boundary checks, which visit the particles of the lists, pass
// it over.
enum : uint8_t { kNonexistent = 0, kNormal = 1, kEntering = 2, kExiting = 3, kSeeded = 4 };

// A DEM boundary as the particle kernels see it (des_bc_mod.f): the DES
// cells whose particles it checks (DEM_BCMI_IJK / DEM_BCMO_IJK as an index
// box, 1-based, ghost layer included), its normal axis and side, the
// position of its plane and, per phase, the velocity BC_U_s/V_s/W_s.
struct BcDev {
  int type;              // MFIXDEM_BC_*
  int ax, low;           // normal axis; 1 when the boundary is at the low end
  int lo[3], hi[3];      // DES cell box
  int dg_ni, dg_nj;      // DES grid extents (to decode a cell index)
  double face;
  double vel[kMaxPh][3];
};
__device__ __forceinline__ bool bc_holds(const BcDev& b, int dc) {
  const int j = dc % b.dg_nj + 1, t = dc / b.dg_nj;
  const int i = t % b.dg_ni + 1, k = t / b.dg_ni + 1;
  return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] &&
         k <= b.hi[2];
}
// distance of a centre from the boundary plane, into the domain
__device__ __forceinline__ double bc_dist(const BcDev& b, const double x[3]) {
  return b.low ? x[b.ax] - b.face : b.face - x[b.ax];
}

// One particle an inlet seeds (SET_NEW_PARTICLE_PROPS, mass_inflow_dem.f:
// 419-642), formed on the host; slot_of_gid places it.
struct SeedDev {
  double pos[3], vel[3];
  double radius, pvol, pmass, omoi;
  int gid, dgc, fcell, phase;
};

struct StepPtrs {
  const double *pos, *vel, *omega;
  double *pos_o, *vel_o, *omega_o;
  const double *radius, *pmass, *omoi, *pvol;
  const int* gid;
  uint8_t* state;
  const uint8_t* phase;
  const int* dgc;
  // walls: per particle kWallSlots wall-collision slots, bit k of wmask =
  // slot k in use, wpft[(k*3 + d)*stride + i] its history.  At most
  // kWallSlots facets (DevParams::nfs > 0): slot k holds facet k and
  // cell_fmask[c] has bit f set for each facet f of DES cell c.  More
  // facets: the table fac, the facet ids of each DES cell
  // (dg_fids[dg_foff[c] .. dg_foff[c+1]), ascending) and wfid[k*stride + i]
  // the facet of slot k.
  const uint8_t* cell_fmask;
  const FacetDev* fac;
  const int *dg_foff, *dg_fids;
  uint8_t* wmask;
  int* wfid;
  double* wpft;
  const int *nbr_start, *nbr_split, *nbr, *nbr_pid;
  const double* pft_in;         // tangential history [3*pair + d], valid where pfl is 1
  double* pft_out;
  const uint8_t* pfl_in;        // 1: the pair's history is set; 0: it is zero (pft unread)
  uint8_t* pfl_out;
  double *fc, *tow;             // FC/TOW in slot order (unfused path)
  double *fc_id, *tow_id;       // FC/TOW in global-id order (inspection)
  double *acc_old, *rot_acc_old;
  int* fcell;
  const double *xe, *yn, *zt;
  const double* csz;  // cell sizes DX(1..ni), DY(1..nj), DZ(1..nk), concatenated
  const double *vnode, *p_force, *ep_g;
  const double* ugc;            // cell-centre gas velocity [3*ijk + d] (filter schemes)
  const uint8_t* fluid;
  int* flags;
  // fused deposition (DevParams::fuse_dep): fixed-point node sums of this
  // step, and the queue of particles that left the fluid grid (slot ids;
  // count at rm_cnt[0]), whose removal the next kernel completes
  unsigned long long* dep_fix;
  int* rm_cnt;
  int* rm_list;
  // explicit coupling, in global-id order: F_GP of the last gas-side drag
  // and DRAG_FC [3*gid + d] of the last DES_TIME_INIT (null otherwise)
  double* f_gp;
  double* drag_fc;
  const BcDev* bcs;             // DEM boundaries (DevParams::nbc of them)
};

enum { kFlagRogueGid = 0, kFlagRogueCount, kFlagZeroDist, kFlagOwnOverflow, kFlagEpg,
       kFlagOutOfDomain, kFlagContacts, kFlagWallSlots, kFlagPairs, kFlagCount };

enum { kModeFused = 0, kModeForce = 1, kModeDrag = 2, kModeIntegrate = 3 };

// XCD-aware tile order.  Workgroups are dealt to the 8 XCDs of the chip
// round-robin (workgroup b runs on XCD b % 8), and each XCD has its own L2.
// Particles are stored in spatial (Morton) order, so neighbouring tiles
// share neighbour and gas-node data; mapping the workgroups of one XCD to a
// contiguous range of tiles keeps that data in one L2 instead of eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb / kXcds, r = nb % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Ensembles: many independent systems in one set of particle arrays.  The
// slots of system s are a fixed range; each slot carries its system id, and
// parameters, grids, search-grid cell offsets, DES-cell offsets and rogue
// flags are per-system tables.  The BATCH kernel variants read them; with
// BATCH = false the kernels take the single system's blocks by value.
struct GridDev;
struct BatchDev {
  const uint16_t* sys;      // system of each slot
  const DevParams* ptab;    // per-system DevParams
  const GridDev* gtab;      // per-system grids
  const int* cell_off;      // search cells of system s: [cell_off[s], cell_off[s+1])
  const int* dg_off;        // first DES cell of system s in the facet lists
  int* rogue;               // [2s]: smallest rogue id, [2s+1]: rogue count
  int nsys;
  int sys_shift;            // bit position of the system id in the sort key
};
template <bool BATCH>
__device__ __forceinline__ const DevParams& sys_params(const DevParams& P0, const BatchDev& B,
                                                       int i) {
  if constexpr (BATCH) return B.ptab[B.sys[i]];
  else return P0;
}

__device__ __forceinline__ void load3(const double* a, int S, int i, double v[3]) {
  v[0] = a[i];
  v[1] = a[S + i];
  v[2] = a[2 * S + i];
}
__device__ __forceinline__ void store3(double* a, int S, int i, const double v[3]) {
  a[i] = v[0];
  a[S + i] = v[1];
  a[2 * S + i] = v[2];
}

// Fluid cell of a particle, kept packed as I | J << 10 | K << 20 (1-based,
// ghost layer included; negative = outside the grid) so that kernels get
// I, J, K with shifts instead of divisions by the grid size.
__host__ __device__ __forceinline__ int pack_ijk(int I, int J, int K) {
  return I | (J << 10) | (K << 20);
}
__device__ __forceinline__ void unpack_ijk(int c, int& I, int& J, int& K) {
  I = c & 1023;
  J = (c >> 10) & 1023;
  K = c >> 20;
}
// linear IJK (FUNIJK, J fastest), 0-based
__device__ __forceinline__ int lin_ijk(int I, int J, int K, int ni, int nj) {
  return (J - 1) + (I - 1) * nj + (K - 1) * nj * ni;
}

// 10-bit value spread to every third bit (Morton codes)
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

// ---------------------------------------------------- filter schemes (DES1)
__device__ __forceinline__ bool filter_scheme(int interp) {
  return interp == MFIXDEM_INTERP_SQUARE_DPVM || interp == MFIXDEM_INTERP_GAUSS_DPVM ||
         interp == MFIXDEM_INTERP_LINEAR_HAT;
}
// CALC_FILTER_WEIGHTS (calc_interp_weights.f:141-156): mass of the cubic
// kernel of half-width FILTER_WIDTH_INTERP beyond a face
__device__ __forceinline__ double dpvm_w(const DevParams& P, double pos1, double pos2) {
  const double overlap = pos1 - pos2;
  if (overlap < P.fw) {
    const double height = P.fw - overlap;
    return height * height * (P.fw3 - height) * P.oofv;
  }
  return 0.0;
}
// IS_ON_MYPE_PLUS2LAYERS (functions.inc:169-181, serial) and FLUID_AT
__device__ __forceinline__ bool on_grid(const DevParams& P, int I, int J, int K) {
  return I >= 1 && I <= P.ni && J >= 1 && J <= P.nj && K >= 1 && K <= P.nk;
}
__device__ __forceinline__ bool fluid_at(const DevParams& P, const uint8_t* fluid, int I, int J,
                                         int K) {
  return on_grid(P, I, J, K) && fluid[lin_ijk(I, J, K, P.ni, P.nj)];
}
// one axis of CALC_INTERP_WEIGHTS2 (calc_interp_weights.f:365-460) in cell
// I: face_lo = XE(I-1), c = &DX(I) (its neighbours at c[-1], c[1]);
// oDX_E(L) = ONE/(HALF*(DX(L+1)+DX(L))) (set_geometry.f:341-343)
__device__ __forceinline__ void lhat_axis(double xpos, double face_lo, const double* c, bool fm,
                                          bool fp, double sw[3]) {
  const double xc = face_lo + 0.5 * c[0];
  sw[0] = 0.0;
  sw[1] = 1.0;
  sw[2] = 0.0;
  if (xpos < xc) {
    if (fm) {
      sw[0] = (xc - xpos) * (1.0 / (0.5 * (c[0] + c[-1])));
      sw[1] = 1.0 - sw[0];
    }
  } else if (fp) {
    sw[2] = (xpos - xc) * (1.0 / (0.5 * (c[1] + c[0])));
    sw[1] = 1.0 - sw[2];
  }
}
// The 27 (FILTER_CELL, FILTER_WEIGHT) pairs of a particle in fluid cell
// (ci, cj, ck), in the reference's order: SQUARE/GAUSS_DPVM
// (CALC_INTERP_WEIGHTS1 :47-139, K then I then J fastest; a non-fluid
// target gives its weight to the own cell) and LINEAR_HAT
// (CALC_INTERP_WEIGHTS2 :351-528, K then J then I fastest; a non-fluid
// target gets weight 0).  fn(cell, weight) is called for each, cell 0-based.
template <class Fn>
__device__ __forceinline__ void for_filter(const DevParams& P, const double* xe, const double* yn,
                                           const double* zt, const double* csz,
                                           const uint8_t* fluid, int ci, int cj, int ck,
                                           const double pos[3], Fn fn) {
  const int own = lin_ijk(ci, cj, ck, P.ni, P.nj);
  if (P.interp == MFIXDEM_INTERP_LINEAR_HAT) {
    double sx[3], sy[3], sz[3];
    // c[-1] / c[1] are read only where that neighbour is a fluid cell
    lhat_axis(pos[0], xe[ci - 1], csz + (ci - 1), fluid_at(P, fluid, ci - 1, cj, ck),
              fluid_at(P, fluid, ci + 1, cj, ck), sx);
    lhat_axis(pos[1], yn[cj - 1], csz + P.ni + (cj - 1), fluid_at(P, fluid, ci, cj - 1, ck),
              fluid_at(P, fluid, ci, cj + 1, ck), sy);
    lhat_axis(pos[2], zt[ck - 1], csz + P.ni + P.nj + (ck - 1),
              fluid_at(P, fluid, ci, cj, ck - 1), fluid_at(P, fluid, ci, cj, ck + 1), sz);
#pragma unroll 1
    for (int kc = 0; kc < 3; kc++)
#pragma unroll 1
      for (int jc = 0; jc < 3; jc++)
#pragma unroll 1
        for (int ic = 0; ic < 3; ic++) {
          const int I = ci - 1 + ic, J = cj - 1 + jc, K = ck - 1 + kc;
          const double w = sx[ic] * sy[jc] * sz[kc];
          if (fluid_at(P, fluid, I, J, K)) fn(lin_ijk(I, J, K, P.ni, P.nj), w);
          else fn(own, 0.0);
        }
  } else {
    double wi[3], wj[3], wk[3];
    wi[0] = dpvm_w(P, pos[0], xe[ci - 1]);
    wi[2] = dpvm_w(P, xe[ci], pos[0]);
    wi[1] = 1.0 - wi[0] - wi[2];
    wj[0] = dpvm_w(P, pos[1], yn[cj - 1]);
    wj[2] = dpvm_w(P, yn[cj], pos[1]);
    wj[1] = 1.0 - wj[0] - wj[2];
    wk[0] = dpvm_w(P, pos[2], zt[ck - 1]);
    wk[2] = dpvm_w(P, zt[ck], pos[2]);
    wk[1] = 1.0 - wk[0] - wk[2];
#pragma unroll 1
    for (int kc = 0; kc < 3; kc++)
#pragma unroll 1
      for (int ic = 0; ic < 3; ic++)
#pragma unroll 1
        for (int jc = 0; jc < 3; jc++) {
          const int I = ci - 1 + ic, J = cj - 1 + jc, K = ck - 1 + kc;
          const double w = wi[ic] * wj[jc] * wk[kc];
          if (!on_grid(P, I, J, K)) fn(own, 0.0);
          else if (fluid[lin_ijk(I, J, K, P.ni, P.nj)]) fn(lin_ijk(I, J, K, P.ni, P.nj), w);
          else fn(own, w);
        }
  }
}

// DRAG_GS_DES1 on one particle (drag_gs_des1.f:50-160): gas velocity at
// the cell centres, voidage and P_FORCE interpolated with the filter weights
// when DES_INTERP_ON (else the particle's cell), voidage below 1e-6 falls
// back to the cell value, no drag outside fluid cells.
template <bool MIXED>
__device__ __forceinline__ void add_drag_des1(const DevParams& P, const StepPtrs& a, int i,
                                              const double pos[3], const double vel[3],
                                              double fc[3]) {
  if (P.explicit_cpl) {  // CALC_DRAG_DES, calc_drag_des.f:34-43
    const double* d = a.drag_fc + 3 * (size_t)a.gid[i];
    for (int c = 0; c < 3; c++) fc[c] = fc[c] + d[c];
    return;
  }
  const int pc = a.fcell[i];
  if (pc < 0) return;
  int ci, cj, ck;
  unpack_ijk(pc, ci, cj, ck);
  const int c0 = lin_ijk(ci, cj, ck, P.ni, P.nj);
  double lepg = 0.0, velfp[3] = {0.0, 0.0, 0.0}, lpf[3] = {0.0, 0.0, 0.0};
  if (P.interp_on && filter_scheme(P.interp)) {
    for_filter(P, a.xe, a.yn, a.zt, a.csz, a.fluid, ci, cj, ck, pos, [&](int c, double w) {
      lepg = lepg + a.ep_g[c] * w;
      for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + a.ugc[3 * c + d] * w;
      for (int d = 0; d < 3; d++) lpf[d] = lpf[d] + a.p_force[3 * c + d] * w;
    });
  } else {
    lepg = a.ep_g[c0];
    for (int d = 0; d < 3; d++) velfp[d] = a.ugc[3 * c0 + d];
    for (int d = 0; d < 3; d++) lpf[d] = a.p_force[3 * c0 + d];
  }
  if (lepg < 1.0e-6) lepg = a.ep_g[c0];
  if (!a.fluid[c0]) return;
  const double pvol = a.pvol[i];
  if (!MIXED) {
    const double vs[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
    const double vrel = sqrt(dot3(vs, vs));
    const double fgp = drag_fgp_fp64(P.drag_type, lepg, P.mu_g0, P.ro_g0, vrel,
                                     2.0 * a.radius[i], pvol);
    for (int d = 0; d < 3; d++) fc[d] = fc[d] + fgp * (velfp[d] - vel[d]);
  } else {
    const float vs[3] = {(float)(velfp[0] - vel[0]), (float)(velfp[1] - vel[1]),
                         (float)(velfp[2] - vel[2])};
    const float vrel = fsqrt(dot3f(vs, vs));
    const float fgp = drag_fgp_fp32(P.drag_type, (float)lepg, (float)P.mu_g0, (float)P.ro_g0,
                                    vrel, (float)(2.0 * a.radius[i]), (float)pvol);
    for (int d = 0; d < 3; d++) fc[d] = fc[d] + (double)(fgp * vs[d]);
  }
  for (int d = 0; d < 3; d++) fc[d] = fc[d] + lpf[d] * pvol;
}

// GARG_2012 drag on one particle (DRAG_GS_DES0 :140-212 with
// DRAG_INTERPOLATION :507-566); adds D_FORCE and P_FORCE*PVOL to fc.
template <bool MIXED, bool FILTER>
__device__ __forceinline__ void add_drag(const DevParams& P, const StepPtrs& a, int i,
                                         const double pos[3], const double vel[3], double fc[3]) {
  if (FILTER) {  // any DES_INTERP_SCHEME other than GARG_2012
    add_drag_des1<MIXED>(P, a, i, pos, vel, fc);
    return;
  }
  const int pc = a.fcell[i];
  if (pc < 0) return;
  int ci, cj, ck;
  unpack_ijk(pc, ci, cj, ck);
  const int c = lin_ijk(ci, cj, ck, P.ni, P.nj);
  if (!a.fluid[c]) return;
  const double x0 = a.xe[ci - 1], y0 = a.yn[cj - 1], z0 = a.zt[ck - 1];
  const double dxx = a.xe[ci] - x0, dyy = a.yn[cj] - y0, dzz = a.zt[ck] - z0;
  double zx = pos[0] - x0, zy = pos[1] - y0;
  zx = zx / dxx;
  zy = zy / dyy;
  double zz = pos[2] - z0;
  zz = zz / dzz;
  const double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
  double velfp[3] = {0.0, 0.0, 0.0};
  const int base = (cj - 2) + (ci - 2) * P.nj + (ck - 2) * P.nj * P.ni;  // node (ci-1,cj-1,ck-1)
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const double w = xv[ii] * yv[jj] * zv[kk];
        const double* vn = a.vnode + 3 * (base + jj + ii * P.nj + kk * P.nj * P.ni);
        for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + vn[d] * w;
      }
  const double pvol = a.pvol[i];
  if (!MIXED) {
    const double vs[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
    const double vrel = sqrt(dot3(vs, vs));
    const double fgp = drag_fgp_fp64(P.drag_type, a.ep_g[c], P.mu_g0, P.ro_g0, vrel,
                                     2.0 * a.radius[i], pvol);
    for (int d = 0; d < 3; d++) fc[d] = fc[d] + fgp * (velfp[d] - vel[d]);
  } else {
    const float vs[3] = {(float)(velfp[0] - vel[0]), (float)(velfp[1] - vel[1]),
                         (float)(velfp[2] - vel[2])};
    const float vrel = fsqrt(dot3f(vs, vs));
    const float fgp = drag_fgp_fp32(P.drag_type, (float)a.ep_g[c], (float)P.mu_g0,
                                    (float)P.ro_g0, vrel, (float)(2.0 * a.radius[i]), (float)pvol);
    for (int d = 0; d < 3; d++) fc[d] = fc[d] + (double)(fgp * vs[d]);
  }
  for (int d = 0; d < 3; d++) fc[d] = fc[d] + a.p_force[3 * c + d] * pvol;
}

// The same GARG_2012 drag, evaluated ahead of the pair pass: returns the
// two increments add_drag adds, D_FORCE in t and P_FORCE*PVOL in q (same
// operations, so adding t and then q later rounds exactly as add_drag),
// or false when the particle adds none.  Every gas load is issued in one
// batch before the fluid test: indices are clamped into the grid, so the
// loads stay in bounds for a particle outside the domain or in a ghost
// cell, whose values are then dropped.  pc is the particle's packed fluid
// cell (FCELL), pvol its volume, dpm its diameter.
template <bool MIXED>
__device__ __forceinline__ bool garg_drag_terms(const DevParams& P, const StepPtrs& a, int pc,
                                                const double pos[3], const double vel[3],
                                                double pvol, double dpm, double t[3],
                                                double q[3]) {
  int ci = 2, cj = 2, ck = 2;
  if (pc >= 0) unpack_ijk(pc, ci, cj, ck);
  const int c = lin_ijk(ci, cj, ck, P.ni, P.nj);
  const int qi = ci > 2 ? ci : 2, qj = cj > 2 ? cj : 2, qk = ck > 2 ? ck : 2;
  const uint8_t fl = a.fluid[c];
  const double epg = a.ep_g[c];
  double pf[3];
  for (int d = 0; d < 3; d++) pf[d] = a.p_force[3 * c + d];
  const double x0 = a.xe[qi - 1], y0 = a.yn[qj - 1], z0 = a.zt[qk - 1];
  const double x1 = a.xe[qi], y1 = a.yn[qj], z1 = a.zt[qk];
  const int base = (qj - 2) + (qi - 2) * P.nj + (qk - 2) * P.nj * P.ni;
  double vn[8][3];
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const double* v = a.vnode + 3 * (base + jj + ii * P.nj + kk * P.nj * P.ni);
        for (int d = 0; d < 3; d++) vn[(kk * 2 + jj) * 2 + ii][d] = v[d];
      }
  if (pc < 0 || !fl) return false;
  const double dxx = x1 - x0, dyy = y1 - y0, dzz = z1 - z0;
  double zx = pos[0] - x0, zy = pos[1] - y0;
  zx = zx / dxx;
  zy = zy / dyy;
  double zz = pos[2] - z0;
  zz = zz / dzz;
  const double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
  double velfp[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const double w = xv[ii] * yv[jj] * zv[kk];
        for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + vn[(kk * 2 + jj) * 2 + ii][d] * w;
      }
  if (!MIXED) {
    const double vs[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
    const double vrel = sqrt(dot3(vs, vs));
    const double fgp = drag_fgp_fp64(P.drag_type, epg, P.mu_g0, P.ro_g0, vrel, dpm, pvol);
    for (int d = 0; d < 3; d++) t[d] = fgp * (velfp[d] - vel[d]);
  } else {
    const float vs[3] = {(float)(velfp[0] - vel[0]), (float)(velfp[1] - vel[1]),
                         (float)(velfp[2] - vel[2])};
    const float vrel = fsqrt(dot3f(vs, vs));
    const float fgp = drag_fgp_fp32(P.drag_type, (float)epg, (float)P.mu_g0, (float)P.ro_g0,
                                    vrel, (float)dpm, (float)pvol);
    for (int d = 0; d < 3; d++) t[d] = (double)(fgp * vs[d]);
  }
  for (int d = 0; d < 3; d++) q[d] = pf[d] * pvol;
  return true;
}

// GET_COLLISION's slot for a new wall contact (calc_collision_wall_mod.f:
// 651-702): the last free slot, else the first slot whose facet is not
// among the cell's facets [fl0, fl1); -1 when neither exists (the
// reference then doubles its slot table).
__device__ __forceinline__ int wall_slot(const StepPtrs& a, int S, int i, uint8_t wm, int fl0,
                                         int fl1) {
  for (int k = kWallSlots - 1; k >= 0; k--)
    if (!((wm >> k) & 1)) return k;
  for (int k = 0; k < kWallSlots; k++) {
    const int id = a.wfid[k * S + i];
    bool listed = false;
    for (int q = fl0; q < fl1; q++) listed = listed || a.dg_fids[q] == id;
    if (!listed) return k;
  }
  return -1;
}

// Facet f of the table read through the scalar cache: f must be the same
// in every lane of the wavefront, and the table is not written while a
// kernel runs (constant address space).
__device__ __forceinline__ FacetDev facet_uniform(const FacetDev* fac, int f) {
  typedef const __attribute__((address_space(4))) FacetDev* ConstFacets;
  const ConstFacets c = (ConstFacets)fac + f;
  FacetDev F;
  for (int v = 0; v < 3; v++)
    for (int d = 0; d < 3; d++) F.v[v][d] = c->v[v][d];
  for (int d = 0; d < 3; d++) {
    F.n[d] = c->n[d];
    F.vt[d] = c->vt[d];
    F.vn[d] = c->vn[d];
  }
  F.fmin = c->fmin;
  F.fmax = c->fmax;
  F.dir = c->dir;
  F.ignore_edges = c->ignore_edges;
  return F;
}

// The pair's history carried unchanged into the next step.
__device__ __forceinline__ void keep_history(const StepPtrs& a, int pid) {
  const uint8_t f = a.pfl_in[pid];
  a.pfl_out[pid] = f;
  if (f)
    for (int c = 0; c < 3; c++) a.pft_out[3 * pid + c] = a.pft_in[3 * pid + c];
}

// One pair entry of a row (calc_force_dem.f:210-412): particle ip (position
// xp, radius rp) is L of the pair (owner, own row) or I (incoming row);
// partner j, pair id pid.
// Writes the pair's new history when ip owns it and returns 1 with the
// increments of ip's force and torque in d[0..5] when the pair is in
// contact; returns 0 when it adds nothing.
template <bool MIXED, bool BATCH>
__device__ __forceinline__ int pair_entry(const DevParams& P0, const StepPtrs& a,
                                          const BatchDev& B, int ip, const double xp[3],
                                          double rp, int j, int pid, bool owner, double d[6]) {
  const int S = P0.stride;
  const DevParams& P = sys_params<BATCH>(P0, B, ip);
  const uint8_t stj = a.state[j];
  double xj[3];
  load3(a.pos, S, j, xj);
  const double rj = a.radius[j];
  if (stj == 0) {  // IF(IS_NONEXISTENT(I)) CYCLE: history kept
    if (owner) keep_history(a, pid);
    return 0;
  }
  // contact test in the reference's orientation: |x_I - x_L|^2, r_L + r_I
  const double r_lm = owner ? rp + rj : rj + rp;
  const double d0 = owner ? xj[0] - xp[0] : xp[0] - xj[0];
  const double d1 = owner ? xj[1] - xp[1] : xp[1] - xj[1];
  const double d2 = owner ? xj[2] - xp[2] : xp[2] - xj[2];
  const double dm = d0 * d0 + d1 * d1 + d2 * d2;
  const double rs = r_lm - kSmall;
  if (dm > rs * rs) {  // PFT_NEIGHBOR(:,CC) = 0: the flag says so, no 24-byte write
    if (owner) a.pfl_out[pid] = 0;
    return 0;
  }
  PairIn in;
  double vj[3], wj[3], vp[3], wp[3];
  load3(a.vel, S, j, vj);
  load3(a.omega, S, j, wj);
  load3(a.vel, S, ip, vp);
  load3(a.omega, S, ip, wp);
  const double mj = a.pmass[j], mp = a.pmass[ip];
  const int phj = a.phase[j], php = a.phase[ip];
  for (int c = 0; c < 3; c++) {
    in.xL[c] = owner ? xp[c] : xj[c];
    in.xI[c] = owner ? xj[c] : xp[c];
    in.vL[c] = owner ? vp[c] : vj[c];
    in.vI[c] = owner ? vj[c] : vp[c];
    in.wL[c] = owner ? wp[c] : wj[c];
    in.wI[c] = owner ? wj[c] : wp[c];
  }
  in.rL = owner ? rp : rj;
  in.rI = owner ? rj : rp;
  in.mL = owner ? mp : mj;
  in.mI = owner ? mj : mp;
  in.phL = owner ? php : phj;
  in.phI = owner ? phj : php;
  double hin[3] = {0.0, 0.0, 0.0};
  if (a.pfl_in[pid])
    for (int c = 0; c < 3; c++) hin[c] = a.pft_in[3 * pid + c];
  PairOut o;
  const int r = MIXED ? pair_force_mixed(P, in, hin, o) : pair_force_fp64(P, in, hin, o);
  if (r == 2) {  // counted once per pair, on the owner's row
    if (owner) {
      atomicAdd(&a.flags[kFlagZeroDist], 1);
      keep_history(a, pid);
    }
    return 0;
  }
  if (owner) {
    for (int c = 0; c < 3; c++) a.pft_out[3 * pid + c] = o.pft[c];
    a.pfl_out[pid] = 1;
    if (P0.store_forces) atomicAdd(&a.flags[kFlagContacts], 1);
    for (int c = 0; c < 3; c++) d[c] = o.fct[c];
    for (int c = 0; c < 3; c++) d[3 + c] = o.tl[c];
  } else {  // FC(I) = FC(I) - FC_pair: adding the negation rounds the same
    for (int c = 0; c < 3; c++) d[c] = -o.fct[c];
    for (int c = 0; c < 3; c++) d[3 + c] = o.ti[c];
  }
  return 1;
}

// The pair part of CALC_FORCE_DEM for the 64 particles of a wavefront.
// Their rows ([incoming | own], consecutive particles) are one contiguous
// run of entries; the lanes take the entries 64 at a time (one pair entry
// per lane, so a long row does not hold the other lanes back), put each
// entry's force and torque increment in LDS, and then every lane adds the
// increments of its own row in row order -- the reference's summation
// order, so the sums are bit-identical to a per-particle loop.  Rows of a
// nonexistent particle only carry their history over.  Every lane of the
// wavefront must call this (lanes past n with live = false).  The lane's
// row is [r0, r1) with its own entries from sp on; xr is the lane's own
// position and radius, which the lanes working on its row read from LDS.
template <bool MIXED, bool BATCH>
__device__ __forceinline__ void wave_pairs(const DevParams& P0, const StepPtrs& a,
                                           const BatchDev& B, int i, int st, int r0, int sp,
                                           int r1, const double xr[4], double fc[3],
                                           double tow[3]) {
  __shared__ double s_xr[kBS / 64][4][64];
  __shared__ int s_row[kBS / 64][65];
  __shared__ int s_split[kBS / 64][64];
  __shared__ int s_st[kBS / 64][64];
  __shared__ int s_hit[kBS / 64][64];
  __shared__ double s_d[kBS / 64][6][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  s_row[w][lane] = r0;
  s_split[w][lane] = sp;
  s_st[w][lane] = st;
  for (int c = 0; c < 4; c++) s_xr[w][c][lane] = xr[c];
  if (lane == 63) s_row[w][64] = r1;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int e_first = s_row[w][0], e_last = s_row[w][64];
  const int my1 = s_row[w][lane + 1];
  const int wbase = i - lane;
  for (int base = e_first; base < e_last; base += 64) {
    const int f = base + lane;
    int hit = 0;
    double d[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (f < e_last) {
      // the lane whose row holds entry f: the last p with s_row[p] <= f
      int p = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (s_row[w][p + step] <= f) p += step;
      const bool owner = f >= s_split[w][p];
      const int pid = a.nbr_pid[f];
      if (s_st[w][p] == 0) {
        if (owner) keep_history(a, pid);  // rows of a removed particle keep their history
      } else {
        const double xp[3] = {s_xr[w][0][p], s_xr[w][1][p], s_xr[w][2][p]};
        hit = pair_entry<MIXED, BATCH>(P0, a, B, wbase + p, xp, s_xr[w][3][p], a.nbr[f], pid,
                                       owner, d);
      }
    }
    s_hit[w][lane] = hit;
#pragma unroll
    for (int k = 0; k < 6; k++) s_d[w][k][lane] = d[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int lo = r0 > base ? r0 : base, hi = my1 < base + 64 ? my1 : base + 64;
    for (int g = lo; g < hi; g++) {
      const int t = g - base;
      if (s_hit[w][t]) {
        for (int c = 0; c < 3; c++) fc[c] = fc[c] + s_d[w][c][t];
        for (int c = 0; c < 3; c++) tow[c] = tow[c] + s_d[w][3 + c][t];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Inputs of the fused deposition from one particle's step: its new
// position, mass and phase and the fluid cell it is deposited into (-1:
// none), or a removal (the particle left the fluid grid).
struct DepIn {
  int pc = -1;
  int ph = 0;
  bool remove = false;
  double x[3] = {0.0, 0.0, 0.0};
  double m = 0.0;
};

// COMP_MEAN_FIELDS0 deposition (comp_mean_fields0.f:186-210) fused into the
// step kernel, one particle per thread: trilinear node weights x PMASS are
// rounded to int64 fixed point (as in k_deposit_list, so the integer sums
// and hence ROP_S / EP_G are identical) and summed in an LDS tile over the
// bounding box of the workgroup's nodes (tile, cap entries, which aliases
// LDS that the step no longer uses), then one global atomic per non-zero
// node; a larger box or a multi-phase system adds straight to fix.  A
// particle that left the grid is queued (slot id) for the next kernel,
// which sets its state and radius to 0: other lanes of this launch may
// still read them as a partner's.  Every thread of the workgroup calls
// this (it holds barriers).
__device__ __forceinline__ void fused_deposit(const DevParams& P, const StepPtrs& a, int i,
                                              const DepIn& d, unsigned long long* tile,
                                              int cap) {
  __shared__ int box[6];
  if (d.remove) {
    const int q = atomicAdd(a.rm_cnt, 1);
    a.rm_list[q] = i;
    atomicAdd(&a.flags[kFlagOutOfDomain], 1);
  }
  int ci = 0, cj = 0, ck = 0;
  if (d.pc >= 0) unpack_ijk(d.pc, ci, cj, ck);
  int lo0 = d.pc >= 0 ? ci : INT_MAX, lo1 = d.pc >= 0 ? cj : INT_MAX;
  int lo2 = d.pc >= 0 ? ck : INT_MAX;
  int hi0 = d.pc >= 0 ? ci : INT_MIN, hi1 = d.pc >= 0 ? cj : INT_MIN;
  int hi2 = d.pc >= 0 ? ck : INT_MIN;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo0 = min(lo0, __shfl_xor(lo0, o, 64));
    lo1 = min(lo1, __shfl_xor(lo1, o, 64));
    lo2 = min(lo2, __shfl_xor(lo2, o, 64));
    hi0 = max(hi0, __shfl_xor(hi0, o, 64));
    hi1 = max(hi1, __shfl_xor(hi1, o, 64));
    hi2 = max(hi2, __shfl_xor(hi2, o, 64));
  }
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
  __syncthreads();  // also: every lane is done with the LDS that tile aliases
  if ((threadIdx.x & 63) == 0 && lo0 != INT_MAX) {
    atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
    atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
  }
  __syncthreads();
  // tile nodes I in [bi-1, box_hi]: a cell I touches nodes I-1 and I
  const int bi = box[0], bj = box[1], bk = box[2];
  const int ti_n = box[3] - bi + 2, tj_n = box[4] - bj + 2, tk_n = box[5] - bk + 2;
  const bool any = bi != INT_MAX;
  const int vol = any ? ti_n * tj_n * tk_n : 0;
  const bool tiled = any && P.nph == 1 && vol <= cap;
  if (tiled)
    for (int q = threadIdx.x; q < vol; q += kBS) tile[q] = 0ull;
  __syncthreads();
  if (d.pc >= 0) {
    const double x0 = a.xe[ci - 1], y0 = a.yn[cj - 1], z0 = a.zt[ck - 1];
    const double dxx = a.xe[ci] - x0, dyy = a.yn[cj] - y0, dzz = a.zt[ck] - z0;
    double zx = d.x[0] - x0, zy = d.x[1] - y0;
    zx = zx / dxx;
    zy = zy / dyy;
    double zz = d.x[2] - z0;
    zz = zz / dzz;
    const double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
    const double sc = P.dep_scale[d.ph];
    if (tiled) {
      const int tb = (ci - bi) + ti_n * ((cj - bj) + tj_n * (ck - bk));
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const double w = xv[q & 1] * yv[(q >> 1) & 1] * zv[q >> 2];
        const double v = w * d.m;
        const unsigned long long qv = (unsigned long long)__double2ll_rn(v * sc);
        atomicAdd(&tile[tb + (q & 1) + ti_n * (((q >> 1) & 1) + tj_n * (q >> 2))], qv);
      }
    } else {
      unsigned long long* f = a.dep_fix + (size_t)d.ph * P.ncell;
      const int base = lin_ijk(ci - 1, cj - 1, ck - 1, P.ni, P.nj);
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const double w = xv[q & 1] * yv[(q >> 1) & 1] * zv[q >> 2];
        const double v = w * d.m;
        const unsigned long long qv = (unsigned long long)__double2ll_rn(v * sc);
        atomicAdd(&f[base + ((q >> 1) & 1) + (q & 1) * P.nj + (q >> 2) * P.nj * P.ni], qv);
      }
    }
  }
  if (!tiled) return;  // uniform over the workgroup
  __syncthreads();
  int ti = threadIdx.x % ti_n, tj = threadIdx.x / ti_n, tk = tj / tj_n;
  tj -= tk * tj_n;
  const int step_i = kBS % ti_n, step_r = kBS / ti_n;
  for (int q = threadIdx.x; q < vol; q += kBS) {
    const unsigned long long v = tile[q];
    if (v != 0ull) {
      const int I = bi - 1 + ti, J = bj - 1 + tj, K = bk - 1 + tk;  // node (I,J,K), 1-based
      atomicAdd(&a.dep_fix[lin_ijk(I, J, K, P.ni, P.nj)], v);
    }
    ti += step_i;
    int rows = step_r;
    if (ti >= ti_n) {
      ti -= ti_n;
      rows++;
    }
    tj += rows;
    while (tj >= tj_n) {
      tj -= tj_n;
      tk++;
    }
  }
}

// The per-particle part of a step after the pair pass (drag, integrator,
// fluid-cell locate) for a live particle i; fills dep for the fused
// deposition when P0.fuse_dep.
// UPDATE_EXIT_DES_VEL (mass_outflow_dem.f:444-514), Euler: an exiting
// particle in the cells of an outlet other than MASS_OUTFLOW takes this
// step's acceleration and keeps only the component normal to the outlet,
// once per such outlet.
__device__ __forceinline__ void exit_velocity(const DevParams& P, const BcDev* bcs, int dc,
                                           const double fc[3], double mi, double vel[3]) {
  for (int k = 0; k < P.nbc; k++) {
    const BcDev& b = bcs[k];
    if (b.type != MFIXDEM_BC_OUTFLOW || !bc_holds(b, dc)) continue;
    for (int c = 0; c < 3; c++) {
      vel[c] = vel[c] + P.dt * (fc[c] / mi + P.grav[c]);
      vel[c] = vel[c] * (c == b.ax ? 1.0 : 0.0);
    }
  }
}

template <bool MIXED, int MODE, bool FILTER, bool BATCH, bool kEarlyDrag>
__device__ __forceinline__ void step_tail(const DevParams& P0, const DevParams& P,
                                          const StepPtrs& a, int i, int st, double fc[3],
                                          double tow[3], bool drag, const double* s_drag,
                                          DepIn& dep) {
  const int S = P0.stride;
  if (st == 0) {
    if (MODE == kModeFused || MODE == kModeIntegrate) {
      for (int c = 0; c < 3; c++) {
        a.pos_o[c * S + i] = a.pos[c * S + i];
        a.vel_o[c * S + i] = a.vel[c * S + i];
        a.omega_o[c * S + i] = a.omega[c * S + i];
      }
    }
    return;
  }
  if (MODE == kModeForce) {
    store3(a.fc, S, i, fc);
    store3(a.tow, S, i, tow);
    return;
  }
  double pos[3], vel[3], om[3];
  load3(a.pos, S, i, pos);
  load3(a.vel, S, i, vel);
  load3(a.omega, S, i, om);
  const double mi = a.pmass[i];
  if (MODE == kModeDrag) {
    load3(a.fc, S, i, fc);
    if (P.coupled) add_drag<MIXED, FILTER>(P, a, i, pos, vel, fc);
    store3(a.fc, S, i, fc);
    return;
  }
  if (kEarlyDrag) {
    if (drag) {
      for (int d = 0; d < 3; d++) fc[d] = fc[d] + s_drag[d * kBS + threadIdx.x];
      for (int d = 0; d < 3; d++) fc[d] = fc[d] + s_drag[(3 + d) * kBS + threadIdx.x];
    }
  } else if (MODE == kModeFused && P.coupled) {
    add_drag<MIXED, FILTER>(P, a, i, pos, vel, fc);
  }
  if (MODE == kModeIntegrate) {
    load3(a.fc, S, i, fc);
    load3(a.tow, S, i, tow);
  }
  if (P0.store_forces) {
    const int g = a.gid[i];
    store3(a.fc_id, S, g, fc);
    store3(a.tow_id, S, g, tow);
  }

  // ---- CFNEWVALUES (cfnewvalues.f:700-811)
  const double dt = P.dt;
  const double omoi = a.omoi[i];
  if (!P.ab2) {
    // entering and exiting particles move at their velocity (:492, :593)
    const bool normal = st == kNormal;
    if (st == kExiting && P.nbc) exit_velocity(P, a.bcs, a.dgc[i], fc, mi, vel);
    for (int c = 0; c < 3; c++) {
      if (normal) vel[c] = vel[c] + dt * (fc[c] / mi + P.grav[c]);
      pos[c] = pos[c] + vel[c] * dt;
    }
    if (normal)
      for (int c = 0; c < 3; c++) om[c] = om[c] + tow[c] * omoi * dt;
  } else {
    double acc[3], racc[3];
    if (P0.first_pass) {
      for (int c = 0; c < 3; c++) acc[c] = fc[c] / mi + P.grav[c];
      for (int c = 0; c < 3; c++) racc[c] = tow[c];
    } else {
      load3(a.acc_old, S, i, acc);
      load3(a.rot_acc_old, S, i, racc);
    }
    for (int c = 0; c < 3; c++) {
      const double f = fc[c] / mi + P.grav[c];
      const double vn = vel[c] + 0.5 * (3.0 * f - acc[c]) * dt;
      acc[c] = f;
      pos[c] = pos[c] + 0.5 * (vel[c] + vn) * dt;
      vel[c] = vn;
    }
    for (int c = 0; c < 3; c++) {
      om[c] = om[c] + 0.5 * (3.0 * tow[c] * omoi - racc[c]) * dt;
      racc[c] = tow[c] * omoi;
    }
    store3(a.acc_old, S, i, acc);
    store3(a.rot_acc_old, S, i, racc);
  }

  // ---- PARTICLES_IN_CELL locate, fused (particles_in_cell.f:87-154)
  int pc = 0;
  if (MODE == kModeFused && P.coupled && !P.explicit_cpl) {
    const int I = locate_cell(a.xe, P.ni, 1.0 / (a.xe[1] - a.xe[0]), pos[0]);
    const int J = locate_cell(a.yn, P.nj, 1.0 / (a.yn[1] - a.yn[0]), pos[1]);
    const int K = locate_cell(a.zt, P.nk, 1.0 / (a.zt[1] - a.zt[0]), pos[2]);
    pc = (I && J && K) ? pack_ijk(I, J, K) : -1;
    a.fcell[i] = pc;
  }
  if (kEarlyDrag && P0.fuse_dep && pc < 0) {
    // PARTICLES_IN_CELL deletion (particles_in_cell.f:175-184): state and
    // radius follow in the next kernel
    const double z[3] = {0.0, 0.0, 0.0};
    store3(a.pos_o, S, i, z);
    store3(a.vel_o, S, i, z);
    store3(a.omega_o, S, i, z);
    a.wmask[i] = 0;
    dep.remove = true;
    return;
  }
  store3(a.pos_o, S, i, pos);
  store3(a.vel_o, S, i, vel);
  store3(a.omega_o, S, i, om);
  if (kEarlyDrag && P0.fuse_dep) {
    int I, J, K;
    unpack_ijk(pc, I, J, K);
    if (a.fluid[lin_ijk(I, J, K, P.ni, P.nj)]) {
      dep.pc = pc;
      dep.ph = a.phase[i];
      dep.m = mi;
      for (int c = 0; c < 3; c++) dep.x[c] = pos[c];
    }
  }
}

// One DES_TIME_STEP particle body.  MODE selects the fused step or one of the
// unfused reference routines (CALC_FORCE_DEM / CALC_DRAG_DES / CFNEWVALUES);
// FILTER selects the DRAG_GS_DES1 drag (any interpolation scheme other than
// GARG_2012), so that each kernel carries only one drag path.
//
// The fused GARG_2012 step evaluates the drag before the walls and the
// pair pass (garg_drag_terms), so that its gas loads travel with the
// particle's own loads instead of forming a second chain of dependent
// loads after the pairs; the two increments wait in LDS and are added
// after the pair forces, where CALC_DRAG_DES adds them.  With
// P0.fuse_dep it also deposits the particles' new positions
// (COMP_MEAN_FIELDS0 node sums) before it ends.
template <bool MIXED, int MODE, bool FILTER, bool BATCH = false, bool STL = false>
__global__ __launch_bounds__(kBS) void k_step(DevParams P0, StepPtrs a, BatchDev B) {
  constexpr bool kPairs = MODE == kModeFused || MODE == kModeForce;
  constexpr bool kEarlyDrag = MODE == kModeFused && !FILTER && !BATCH;
  __shared__ double s_drag[kEarlyDrag ? 6 : 1][kBS];
  const int i = xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
  const bool live = i < P0.n;
  if (!kPairs && !live) return;  // the pair modes keep every lane for the wave passes
  const int S = P0.stride;
  const DevParams& P = sys_params<BATCH>(P0, B, live ? i : 0);
  const int st = live ? a.state[i] : 0;
  double fc[3] = {0.0, 0.0, 0.0}, tow[3] = {0.0, 0.0, 0.0};
  bool drag = false;

  if (kPairs) {
    // the row bounds go out with the particle's own loads
    const int n = P0.n;
    const int r0 = a.nbr_start[live ? i : n];
    const int sp = live ? a.nbr_split[i] : r0;
    const int r1 = a.nbr_start[live ? i + 1 : n];
    double xr[4] = {0.0, 0.0, 0.0, 0.0};
    if (live && st != 0) {
      double pos[3], vel[3], om[3];
      load3(a.pos, S, i, pos);
      load3(a.vel, S, i, vel);
      load3(a.omega, S, i, om);
      const double ri = a.radius[i], mi = a.pmass[i];
      for (int c = 0; c < 3; c++) xr[c] = pos[c];
      xr[3] = ri;
      const int phi = a.phase[i];
      int pc = -1;
      double pv = 0.0;
      if (kEarlyDrag && P.coupled) {
        pc = a.fcell[i];
        pv = a.pvol[i];
      }
      // ---- walls (calc_collision_wall_mod.f:184-606), the cell's facets in
      // ascending id
      // only normal particles meet the walls (calc_collision_wall_mod.f:
      // 105); entering and exiting ones keep their wall slots as they are
      const bool walls = st == kNormal;
      const int dc = (BATCH ? B.dg_off[B.sys[i]] : 0) + a.dgc[i];
      const int fm = STL || !walls ? 0 : a.cell_fmask[dc];
      const int fl0 = STL && walls ? a.dg_foff[dc] : 0;
      const int fl1 = STL && walls ? a.dg_foff[dc + 1] : 0;
      uint8_t wm = a.wmask[i];
      if (kEarlyDrag && P.coupled) {
        double t[3], q[3];
        drag = garg_drag_terms<MIXED>(P, a, pc, pos, vel, pv, 2.0 * ri, t, q);
        if (drag)
          for (int d = 0; d < 3; d++) {
            s_drag[d][threadIdx.x] = t[d];
            s_drag[3 + d][threadIdx.x] = q[d];
          }
      }
      const uint8_t wm0 = wm;
      if (!STL && walls) {
        if (fm == 0) wm = 0;  // no facets in the cell: every slot cleared (:189-193)
        for (int f = 0; f < P.nfs; f++) {
          if (!((fm >> f) & 1)) continue;
          double hist[3];
          const bool set = (wm >> f) & 1;
          if (set)
            for (int d = 0; d < 3; d++) hist[d] = a.wpft[(f * 3 + d) * S + i];
          const int r = wall_contact<MIXED>(P, P.facets[f], pos, vel, om, ri, mi, phi, hist, set,
                                            fc, tow);
          if (r == 0) {  // REMOVE_COLLISION
            wm &= (uint8_t)~(1u << f);
          } else if (r == 1) {  // GET_COLLISION / UPDATE_COLLISION
            wm |= (uint8_t)(1u << f);
            for (int d = 0; d < 3; d++) a.wpft[(f * 3 + d) * S + i] = hist[d];
          }
        }
      }
      if (STL && walls && fl0 == fl1) wm = 0;
      // The facets are walked in wave-uniform passes: a pass takes the next
      // facet of the first lane that has one left, reads it once for the
      // wavefront through the scalar cache, and the lanes whose next facet
      // it is evaluate it; every lane still meets its own facets in
      // ascending id.
      for (int q = fl0; STL;) {
        const int mine = q < fl1 ? a.dg_fids[q] : -1;
        const unsigned long long left = __ballot(mine >= 0);
        if (left == 0ull) break;
        const int f = __builtin_amdgcn_readlane(mine, __builtin_ctzll(left));
        const FacetDev F = facet_uniform(a.fac, f);
        // lanes taking this pass (a ballot, so that the facet index stays
        // the wave-uniform f rather than the lane's own copy of it)
        if (!((__ballot(mine == f) >> __lane_id()) & 1ull)) continue;
        q++;
        int k = -1;  // the slot holding facet f
        for (int s2 = 0; s2 < kWallSlots; s2++)
          if (k < 0 && ((wm >> s2) & 1) && a.wfid[s2 * S + i] == f) k = s2;
        double hist[3];
        if (k >= 0)
          for (int d = 0; d < 3; d++) hist[d] = a.wpft[(k * 3 + d) * S + i];
        const int r = wall_contact<MIXED>(P, F, pos, vel, om, ri, mi, phi, hist, k >= 0, fc, tow);
        if (r == 0) {  // REMOVE_COLLISION
          if (k >= 0) wm &= (uint8_t)~(1u << k);
        } else if (r == 1) {  // GET_COLLISION / UPDATE_COLLISION
          if (k < 0) k = wall_slot(a, S, i, wm, fl0, fl1);
          if (k >= 0) {
            a.wfid[k * S + i] = f;
            for (int d = 0; d < 3; d++) a.wpft[(k * 3 + d) * S + i] = hist[d];
            wm |= (uint8_t)(1u << k);
          } else {
            atomicOr(&a.flags[kFlagWallSlots], 1);
          }
        }
      }
      if (wm != wm0) a.wmask[i] = wm;
    }
    // ---- pairs (calc_force_dem.f:180-412), summed after the walls
    wave_pairs<MIXED, BATCH>(P0, a, B, i, st, r0, sp, r1, xr, fc, tow);
  }
  DepIn dep;
  if (live)
    step_tail<MIXED, MODE, FILTER, BATCH, kEarlyDrag>(P0, P, a, i, st, fc, tow, drag,
                                                      &s_drag[0][0], dep);
  if constexpr (kEarlyDrag) {
    // s_drag is read above and free once the first barrier inside has passed
    if (P0.fuse_dep)
      fused_deposit(P0, a, i, dep, reinterpret_cast<unsigned long long*>(&s_drag[0][0]), 6 * kBS);
  }
}

// PARTICLES_IN_CELL deletion of a particle outside the fluid grid
// (particles_in_cell.f:175-184); the deposition kernels, which visit every
// particle once, apply it.
struct Removal {
  uint8_t* state;
  uint8_t* wmask;
  double *pos, *vel, *omega, *radius;
  int* flags;
};
__device__ __forceinline__ void remove_outside(const Removal& r, int S, int i) {
  r.state[i] = 0;
  r.wmask[i] = 0;
  for (int d = 0; d < 3; d++) {
    r.pos[d * S + i] = 0.0;
    r.vel[d * S + i] = 0.0;
    r.omega[d * S + i] = 0.0;
  }
  r.radius[i] = 0.0;
  atomicAdd(&r.flags[kFlagOutOfDomain], 1);
}

// COMP_MEAN_FIELDS0 deposition (comp_mean_fields0.f:186-210) of the current
// positions (at set-up and for the unfused routines; the fused step kernel
// deposits itself).  The particles are visited in slot order (list null,
// the Morton order of their search cells) or in list order, kDepPPT
// consecutive ones per thread (2 measured fastest at 1M spheres, against 1
// and 4): a run of particles in the same fluid cell
// is summed in registers, each run adds its 8 corner values into an int64
// tile in LDS covering the bounding box of the workgroup's nodes (up to
// kDepCap nodes), and each touched tile node then costs one global atomic.
// A workgroup whose box is larger, or a multi-phase system, adds straight
// to global memory.  Particles that changed fluid cell since the rebuild
// only widen the box.  Weights are rounded to fixed point per particle and
// summed as integers, so the result does not depend on the order.
constexpr int kDepPPT = 2;
constexpr int kDepCap = 4096;
__global__ __launch_bounds__(kBS) void k_deposit_list(DevParams P, const int* list,
                                                      const double* pos, const double* pmass,
                                                      const uint8_t* phase, const uint8_t* state,
                                                      const int* fcell, const uint8_t* fluid,
                                                      const double* xe, const double* yn,
                                                      const double* zt, unsigned long long* fix,
                                                      Removal rm) {
  __shared__ unsigned long long tile[kDepCap];
  __shared__ int box[6];
  const int S = P.stride;
  const int k0 = (xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x) * kDepPPT;
  int pi[kDepPPT], pcell[kDepPPT], pph[kDepPPT];
  double px[kDepPPT][3], pm[kDepPPT];
  int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
#pragma unroll
  for (int u = 0; u < kDepPPT; u++) {
    const int k = k0 + u;
    pi[u] = k < P.n ? (list ? list[k] : k) : -1;
    pcell[u] = -1;
  }
  // every per-particle load is issued here, ahead of the workgroup barriers
#pragma unroll
  for (int u = 0; u < kDepPPT; u++) {
    const int i = pi[u] < 0 ? 0 : pi[u];
    pph[u] = phase[i];
    px[u][0] = pos[i];
    px[u][1] = pos[S + i];
    px[u][2] = pos[2 * S + i];
    pm[u] = pmass[i];
  }
#pragma unroll
  for (int u = 0; u < kDepPPT; u++) {
    const int i = pi[u];
    if (i < 0) continue;
    const int pc = fcell[i];
    if (state[i] == 0) continue;
    if (pc < 0) {
      remove_outside(rm, S, i);
      continue;
    }
    int I, J, K;
    unpack_ijk(pc, I, J, K);
    if (!fluid[lin_ijk(I, J, K, P.ni, P.nj)]) continue;
    pcell[u] = pc;
    lo0 = min(lo0, I); lo1 = min(lo1, J); lo2 = min(lo2, K);
    hi0 = max(hi0, I); hi1 = max(hi1, J); hi2 = max(hi2, K);
  }
  // bounding box of the workgroup's cells: reduce within the wavefront,
  // then one LDS atomic per wavefront and bound
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo0 = min(lo0, __shfl_xor(lo0, o, 64));
    lo1 = min(lo1, __shfl_xor(lo1, o, 64));
    lo2 = min(lo2, __shfl_xor(lo2, o, 64));
    hi0 = max(hi0, __shfl_xor(hi0, o, 64));
    hi1 = max(hi1, __shfl_xor(hi1, o, 64));
    hi2 = max(hi2, __shfl_xor(hi2, o, 64));
  }
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && lo0 != INT_MAX) {
    atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
    atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
  }
  __syncthreads();
  // tile nodes I in [bi-1, box_hi]: a cell I touches nodes I-1 and I
  const int bi = box[0], bj = box[1], bk = box[2];
  const int ti_n = box[3] - bi + 2, tj_n = box[4] - bj + 2, tk_n = box[5] - bk + 2;
  const bool any = bi != INT_MAX;
  const int vol = any ? ti_n * tj_n * tk_n : 0;
  const bool tiled = any && P.nph == 1 && vol <= kDepCap;
  if (tiled)
    for (int q = threadIdx.x; q < vol; q += kBS) tile[q] = 0ull;
  __syncthreads();
  unsigned long long acc[8];
  int run_cell = -1, run_ph = 0;
  auto flush_run = [&]() {
    if (run_cell < 0) return;
    int ci, cj, ck;
    unpack_ijk(run_cell, ci, cj, ck);
    if (tiled) {
      const int tb = (ci - bi) + ti_n * ((cj - bj) + tj_n * (ck - bk));
#pragma unroll
      for (int q = 0; q < 8; q++)
        atomicAdd(&tile[tb + (q & 1) + ti_n * (((q >> 1) & 1) + tj_n * (q >> 2))], acc[q]);
    } else {
      unsigned long long* f = fix + (size_t)run_ph * P.ncell;
      const int base = lin_ijk(ci - 1, cj - 1, ck - 1, P.ni, P.nj);
#pragma unroll
      for (int q = 0; q < 8; q++)
        atomicAdd(&f[base + ((q >> 1) & 1) + (q & 1) * P.nj + (q >> 2) * P.nj * P.ni], acc[q]);
    }
  };
#pragma unroll
  for (int u = 0; u < kDepPPT; u++) {
    if (pcell[u] < 0) continue;
    const int ph = pph[u];
    const double x[3] = {px[u][0], px[u][1], px[u][2]};
    const double m = pm[u];
    int ci, cj, ck;
    unpack_ijk(pcell[u], ci, cj, ck);
    const double x0 = xe[ci - 1], y0 = yn[cj - 1], z0 = zt[ck - 1];
    const double dxx = xe[ci] - x0, dyy = yn[cj] - y0, dzz = zt[ck] - z0;
    double zx = x[0] - x0, zy = x[1] - y0;
    zx = zx / dxx;
    zy = zy / dyy;
    double zz = x[2] - z0;
    zz = zz / dzz;
    const double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
    const double sc = P.dep_scale[ph];
    const bool same = pcell[u] == run_cell && ph == run_ph;
    if (!same) {
      flush_run();
      run_cell = pcell[u];
      run_ph = ph;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const double w = xv[q & 1] * yv[(q >> 1) & 1] * zv[q >> 2];
      const double v = w * m;
      const unsigned long long qv = (unsigned long long)__double2ll_rn(v * sc);
      acc[q] = same ? acc[q] + qv : qv;
    }
  }
  flush_run();
  if (!tiled) return;
  __syncthreads();
  // walk this thread's tile entries q = tid, tid + kBS, ... keeping (ti, tj,
  // tk) by carries: one division at the start, none in the loop
  int ti = threadIdx.x % ti_n, tj = threadIdx.x / ti_n, tk = tj / tj_n;
  tj -= tk * tj_n;
  const int step_i = kBS % ti_n, step_r = kBS / ti_n;
  for (int q = threadIdx.x; q < vol; q += kBS) {
    const unsigned long long v = tile[q];
    if (v != 0ull) {
      const int I = bi - 1 + ti, J = bj - 1 + tj, K = bk - 1 + tk;  // node (I,J,K), 1-based
      atomicAdd(&fix[lin_ijk(I, J, K, P.ni, P.nj)], v);
    }
    ti += step_i;
    int rows = step_r;
    if (ti >= ti_n) {
      ti -= ti_n;
      rows++;
    }
    tj += rows;
    while (tj >= tj_n) {
      tj -= tj_n;
      tk++;
    }
  }
}

// COMP_MEAN_FIELDS1 deposition (comp_mean_fields1.f:42-64): PVOL times the
// filter weight into the 27 filter cells (or PVOL into the own cell for the
// NONE scheme), fixed-point int64 sums per cell.  Visits particles in slot
// order (list null) or list order; the workgroup's cells are summed in an LDS tile over
// the bounding box of its targets (up to kDepCap cells), else directly.
__global__ __launch_bounds__(kBS) void k_deposit_filter(DevParams P, const int* list,
                                                        const double* pos, const double* pvol,
                                                        const uint8_t* phase, const uint8_t* state,
                                                        const int* fcell, const double* xe,
                                                        const double* yn, const double* zt,
                                                        const double* csz, const uint8_t* fluid,
                                                        unsigned long long* fix, Removal rm) {
  __shared__ unsigned long long tile[kDepCap];
  __shared__ int box[6];
  const int S = P.stride;
  const int k = blockIdx.x * kBS + threadIdx.x;
  int i = -1, pc = -1, ci = 0, cj = 0, ck = 0;
  if (k < P.n) {
    i = list ? list[k] : k;
    pc = fcell[i];
    if (state[i] == 0) {
      pc = -1;
    } else if (pc < 0) {
      remove_outside(rm, S, i);
    } else {
      unpack_ijk(pc, ci, cj, ck);
    }
  }
  // bounding box: reduce within the wavefront, one LDS atomic per bound
  int lo0 = pc >= 0 ? ci : INT_MAX, lo1 = pc >= 0 ? cj : INT_MAX, lo2 = pc >= 0 ? ck : INT_MAX;
  int hi0 = pc >= 0 ? ci : INT_MIN, hi1 = pc >= 0 ? cj : INT_MIN, hi2 = pc >= 0 ? ck : INT_MIN;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo0 = min(lo0, __shfl_xor(lo0, o, 64));
    lo1 = min(lo1, __shfl_xor(lo1, o, 64));
    lo2 = min(lo2, __shfl_xor(lo2, o, 64));
    hi0 = max(hi0, __shfl_xor(hi0, o, 64));
    hi1 = max(hi1, __shfl_xor(hi1, o, 64));
    hi2 = max(hi2, __shfl_xor(hi2, o, 64));
  }
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && lo0 != INT_MAX) {
    atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
    atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
  }
  __syncthreads();
  // target cells lie in [min-1, max+1] per axis
  const int bi = box[0] - 1, bj = box[1] - 1, bk = box[2] - 1;
  const int ti_n = box[3] - box[0] + 3, tj_n = box[4] - box[1] + 3, tk_n = box[5] - box[2] + 3;
  const bool any = box[0] != INT_MAX;
  const int vol = any ? ti_n * tj_n * tk_n : 0;
  const bool tiled = any && P.nph == 1 && vol <= kDepCap;
  if (tiled)
    for (int q = threadIdx.x; q < vol; q += kBS) tile[q] = 0ull;
  __syncthreads();
  if (pc >= 0) {
    const int ph = phase[i];
    const double vol_wt = pvol[i];
    const double sc = P.dep_scale[ph];
    unsigned long long* f = fix + (size_t)ph * P.ncell;
    const int nj = P.nj, ninj = P.nj * P.ni;
    auto add = [&](int c, double w) {
      const double v = vol_wt * w;
      const unsigned long long qv = (unsigned long long)__double2ll_rn(v * sc);
      if (qv == 0ull) return;
      if (tiled) {
        const int K = c / ninj, r = c - K * ninj, I = r / nj, J = r - I * nj;  // 0-based
        atomicAdd(&tile[(I + 1 - bi) + ti_n * ((J + 1 - bj) + tj_n * (K + 1 - bk))], qv);
      } else {
        atomicAdd(&f[c], qv);
      }
    };
    if (filter_scheme(P.interp)) {
      const double x[3] = {pos[i], pos[S + i], pos[2 * S + i]};
      for_filter(P, xe, yn, zt, csz, fluid, ci, cj, ck, x, add);
    } else {
      add(lin_ijk(ci, cj, ck, P.ni, P.nj), 1.0);
    }
  }
  if (!tiled) return;
  __syncthreads();
  for (int q = threadIdx.x; q < vol; q += kBS) {
    const unsigned long long v = tile[q];
    if (v == 0ull) continue;
    const int ti = q % ti_n, tj = (q / ti_n) % tj_n, tk = q / (ti_n * tj_n);
    atomicAdd(&fix[lin_ijk(bi + ti, bj + tj, bk + tk, P.ni, P.nj)], v);
  }
}

// ---------------------------------------------------------------- scans
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* total, T* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = 0;
    for (int w = 0; w < kBS / 64; w++) {
      const T t = lds[w];
      lds[w] = s;
      s += t;
    }
    lds[kBS / 64] = s;
  }
  __syncthreads();
  const T res = x - v + lds[wid];
  *total = lds[kBS / 64];
  __syncthreads();
  return res;
}

constexpr int kScanPer = 4;
constexpr int kScanTile = kBS * kScanPer;

// exclusive scan of in[0..n) into out[0..n], out[n] = total (3 launches)
__global__ __launch_bounds__(kBS) void k_scan_tiles(const int* in, int* out, int n, int* bsum) {
  __shared__ int lds[kBS / 64 + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  int v[kScanPer], s = 0;
  for (int q = 0; q < kScanPer; q++) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    s += v[q];
  }
  int tot;
  int pre = block_excl_scan(s, &tot, lds);
  for (int q = 0; q < kScanPer; q++) {
    if (base + q < n) out[base + q] = pre;
    pre += v[q];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_sums(int* bsum, int nb) {
  __shared__ int lds[kBS / 64 + 1];
  const int per = (nb + kBS - 1) / kBS;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) s += bsum[b0 + q];
  int tot;
  int pre = block_excl_scan(s, &tot, lds);
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) {
      const int t = bsum[b0 + q];
      bsum[b0 + q] = pre;
      pre += t;
    }
  if (threadIdx.x == 0) bsum[nb] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_add(int* out, int n, const int* bsum, int nb) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) out[i] += bsum[i / kScanTile];
  if (i == 0) out[n] = bsum[nb];
}

// The offsets of the neighbour rows in one scan: each particle's own and
// incoming counts packed as (own << 32) + in, so one exclusive scan gives
// the pair ids (high half) and the row starts (the sum of the halves).
// The low halves never carry: their total is the pair count, below 2^31.
__global__ __launch_bounds__(kBS) void k_scan_rows_tiles(const int* own, const int* in,
                                                         int* pair_base, int* row_start, int n,
                                                         long long* bsum) {
  __shared__ long long lds[kBS / 64 + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  long long v[kScanPer], s = 0;
  for (int q = 0; q < kScanPer; q++) {
    v[q] = (base + q < n) ? ((long long)own[base + q] << 32) + in[base + q] : 0;
    s += v[q];
  }
  long long tot;
  long long pre = block_excl_scan(s, &tot, lds);
  for (int q = 0; q < kScanPer; q++) {
    if (base + q < n) {
      pair_base[base + q] = (int)(pre >> 32);
      row_start[base + q] = (int)(pre >> 32) + (int)(pre & 0xffffffffll);
    }
    pre += v[q];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_sums64(long long* bsum, int nb) {
  __shared__ long long lds[kBS / 64 + 1];
  const int per = (nb + kBS - 1) / kBS;
  const int b0 = threadIdx.x * per;
  long long s = 0;
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) s += bsum[b0 + q];
  long long tot;
  long long pre = block_excl_scan(s, &tot, lds);
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) {
      const long long t = bsum[b0 + q];
      bsum[b0 + q] = pre;
      pre += t;
    }
  if (threadIdx.x == 0) bsum[nb] = tot;
}

__global__ __launch_bounds__(kBS) void k_scan_rows_add(int* pair_base, int* row_start, int n,
                                                       const long long* bsum, int nb,
                                                       int* flags) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) {
    const long long o = bsum[i / kScanTile];
    pair_base[i] += (int)(o >> 32);
    row_start[i] += (int)(o >> 32) + (int)(o & 0xffffffffll);
  }
  if (i == 0) {
    const long long t = bsum[nb];
    pair_base[n] = (int)(t >> 32);
    row_start[n] = (int)(t >> 32) + (int)(t & 0xffffffffll);
    flags[kFlagPairs] = (int)(t >> 32);  // read back with the other flags
  }
}

__global__ void k_fill_i32(int* a, int n, int v) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_f64(double* a, int n, double v) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------------------------- rebuild
struct GridDev {
  // reference DES grid (ghost layer included): cells 1..dg_n*
  int dg_ni, dg_nj, dg_nk;
  int coarse[3];
  double dg_lo[3], dg_inv[3];
  double factor_rlm;
  // neighbour-search grid
  int fn[3];
  double f_lo[3], f_inv;
  int mbits;  // bits of the Morton code in the sort key
  // single-precision prefilter of the row walk: coordinates relative to
  // pf_c; |x - pf_c| <= pf_xmax on every axis for particles it covers;
  // pf_slack bounds the rounding of a distance so computed
  double pf_c[3];
  float pf_xmax, pf_slack;
};

__device__ __forceinline__ int floor_sat_d(double v) {
  if (!(v > -1.0e9)) return -1000000000;
  if (v > 1.0e9) return 1000000000;
  return (int)floor(v);
}
__device__ __forceinline__ void dg_decode(const GridDev& G, int c, int& i, int& j, int& k) {
  j = c % G.dg_nj + 1;
  const int t = c / G.dg_nj;
  i = t % G.dg_ni + 1;
  k = t / G.dg_ni + 1;
}
__device__ __forceinline__ bool dg_ghost(const GridDev& G, int i, int j, int k) {
  return i < 2 || i > G.dg_ni - 1 || j < 2 || j > G.dg_nj - 1 || k < 2 || k > G.dg_nk - 1;
}
// coarse half-stencil offsets (desgrid_mod.f:1111-1152): lo/hi per axis
__device__ __forceinline__ void stencil_off(const GridDev& G, const double x[3], int lo[3],
                                            int hi[3]) {
  for (int d = 0; d < 3; d++) {
    lo[d] = 1;
    hi[d] = 1;
    if (G.coarse[d]) {
      const double t = (x[d] - G.dg_lo[d]) * G.dg_inv[d];
      if (t - floor(t) >= 0.5) lo[d] = 0;
      else hi[d] = 0;
    }
  }
}

template <bool BATCH>
__device__ __forceinline__ const GridDev& sys_grid(const GridDev& G0, const BatchDev& B, int s) {
  if constexpr (BATCH) return B.gtab[s];
  else return G0;
}
// system owning global search cell c (ensembles): last s with cell_off[s] <= c
__device__ __forceinline__ int cell_system(const BatchDev& B, int c) {
  int lo = 0, hi = B.nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (B.cell_off[mid] <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// DESGRID_PIC locate (desgrid_mod.f:900-915) + sort key on the search grid.
// Ensembles put the system id above the Morton code, so every system keeps
// its slot range (its empty slots sort to the end of that range).
template <bool BATCH = false>
__global__ __launch_bounds__(kBS) void k_bin(int n, int S, GridDev G0, const double* pos,
                                             const uint8_t* state, const int* gid, int* dgc,
                                             uint64_t* keys, int* vals, int* flags, BatchDev B) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  vals[i] = i;
  const int s = BATCH ? B.sys[i] : 0;
  const uint64_t sys_key = BATCH ? (uint64_t)s << B.sys_shift : 0ull;
  if (state[i] == 0) {
    keys[i] = BATCH ? sys_key | ((1ull << B.sys_shift) - 1) : ~0ull;
    dgc[i] = 0;
    return;
  }
  const GridDev& G = sys_grid<BATCH>(G0, B, s);
  const double x[3] = {pos[i], pos[S + i], pos[2 * S + i]};
  const int nn[3] = {G.dg_ni, G.dg_nj, G.dg_nk};
  int c3[3], f3[3];
  for (int d = 0; d < 3; d++) {
    int v = floor_sat_d((x[d] - G.dg_lo[d]) * G.dg_inv[d]) + 2;
    c3[d] = v < 1 ? 1 : (v > nn[d] ? nn[d] : v);
    int f = floor_sat_d((x[d] - G.f_lo[d]) * G.f_inv);
    f3[d] = f < 0 ? 0 : (f >= G.fn[d] ? G.fn[d] - 1 : f);
  }
  dgc[i] = (c3[1] - 1) + (c3[0] - 1) * G.dg_nj + (c3[2] - 1) * G.dg_nj * G.dg_ni;
  // rogue: a normal particle in the ghost layer (only normal particles have
  // rows, desgrid_mod.f:1096-1104); entering ones start outside the box
  if (state[i] == kNormal && dg_ghost(G, c3[0], c3[1], c3[2])) {
    if (BATCH) {
      atomicMin(&B.rogue[2 * s], gid[i]);
      atomicAdd(&B.rogue[2 * s + 1], 1);
    }
    atomicMin(&flags[kFlagRogueGid], gid[i]);
    atomicAdd(&flags[kFlagRogueCount], 1);
  }
  // the sort is stable, so particles keep their previous relative order
  // within a search cell (no result depends on that order)
  const uint64_t m = (uint64_t)(spread10(f3[0]) | (spread10(f3[1]) << 1) | (spread10(f3[2]) << 2));
  keys[i] = sys_key | m;
}

// The single-precision copy is relative to G.pf_c; a particle beyond
// pf_xmax on some axis gets radius -1 there (every test with it is then
// made in double precision), a nonexistent one a position no cut-off
// reaches.
constexpr double kFarAway = 1.0e300;
__device__ __forceinline__ void pack_candidate(const GridDev& G, const double p[3], double r,
                                               bool exists, double* xr, float4* xf, int k) {
  double2* o = reinterpret_cast<double2*>(xr) + 2 * k;
  if (!exists) {
    o[0] = make_double2(kFarAway, kFarAway);
    o[1] = make_double2(kFarAway, r);
    xf[k] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
    return;
  }
  o[0] = make_double2(p[0], p[1]);
  o[1] = make_double2(p[2], r);
  float f[3];
  bool in = true;
  for (int d = 0; d < 3; d++) {
    const double q = p[d] - G.pf_c[d];
    in = in && fabs(q) <= (double)G.pf_xmax;
    f[d] = (float)q;
  }
  xf[k] = make_float4(f[0], f[1], f[2], in ? (float)r : -1.0f);
}

// DESGRID_PIC locate as k_bin, for one system, with a counting sort in
// place of the key sort: each particle counts itself into its search cell
// (cells numbered in Morton order by cell_rank; nonexistent particles into
// a last cell after them) and keeps the place it drew.  A cell's particles
// end up in no particular order among themselves (no result depends on it).
__global__ __launch_bounds__(kBS) void k_bin_count(int n, int S, GridDev G, const double* pos,
                                                   uint8_t* state, const int* gid,
                                                   int* dgc, const int* cell_rank,
                                                   int* cell_cnt, int nfine, uint64_t* place,
                                                   int* flags) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  int r = nfine;
  if (state[i] == kNonexistent) {
    dgc[i] = 0;
  } else {
    if (state[i] == kSeeded) state[i] = kEntering;  // binned now: in a DES cell list
    const double x[3] = {pos[i], pos[S + i], pos[2 * S + i]};
    const int nn[3] = {G.dg_ni, G.dg_nj, G.dg_nk};
    int c3[3], f3[3];
    for (int d = 0; d < 3; d++) {
      const int v = floor_sat_d((x[d] - G.dg_lo[d]) * G.dg_inv[d]) + 2;
      c3[d] = v < 1 ? 1 : (v > nn[d] ? nn[d] : v);
      const int f = floor_sat_d((x[d] - G.f_lo[d]) * G.f_inv);
      f3[d] = f < 0 ? 0 : (f >= G.fn[d] ? G.fn[d] - 1 : f);
    }
    dgc[i] = (c3[1] - 1) + (c3[0] - 1) * G.dg_nj + (c3[2] - 1) * G.dg_nj * G.dg_ni;
    if (state[i] == kNormal && dg_ghost(G, c3[0], c3[1], c3[2])) {  // rogue, as in k_bin
      atomicMin(&flags[kFlagRogueGid], gid[i]);
      atomicAdd(&flags[kFlagRogueCount], 1);
    }
    r = cell_rank[f3[0] + G.fn[0] * (f3[1] + G.fn[1] * f3[2])];
  }
  // The particles are still in the last binning's order, so neighbouring
  // lanes mostly share a cell: each run of equal cells takes its places
  // with one atomic, by its first lane.  The active lanes are a prefix of
  // the wavefront (lanes past n returned above), so a run ends at the next
  // run's first lane or at the last active lane.
  const int lane = __lane_id();
  const int prev = __shfl_up(r, 1, 64);
  const unsigned long long act = __ballot(1);
  const unsigned long long head = __ballot(lane == 0 || prev != r);
  const unsigned long long below = head & ((2ull << lane) - 1ull);   // heads at or below lane
  const int first = 63 - __clzll(below);
  const unsigned long long above = head & ~((2ull << lane) - 1ull);   // heads after lane
  const int last = above ? __builtin_ctzll(above) - 1 : 63 - __clzll(act);
  int w0 = 0;
  if (lane == first) w0 = atomicAdd(&cell_cnt[r], last - first + 1);
  w0 = __shfl(w0, first, 64);
  place[i] = ((uint64_t)r << 32) | (uint32_t)(w0 + lane - first);
}

// The new order (perm[new slot] = old slot) from the counted places and the
// cells' first slots (cell_off: exclusive scan of the counts, in rank
// order); and each search cell's slot range, indexed by its grid position.
// The counts go back to zero for the next binning.
struct BinPack {               // k_bin_place's packing of the row walk's candidates
  GridDev G;
  int S;
  const double *pos, *radius;
  const uint8_t* state;
  double* xr;
  float4* xf;
};
__global__ __launch_bounds__(kBS) void k_bin_place(int n, int nfine, const uint64_t* place,
                                                   const int* cell_rank, const int* cell_off,
                                                   int* cell_cnt, int* perm, int* inv,
                                                   int* cstart, int* cend, int* zero0,
                                                   int* zero1, int* zero2, BinPack bp) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) {
    const uint64_t p = place[i];
    const int np = cell_off[(int)(p >> 32)] + (int)(uint32_t)p;
    perm[np] = i;
    inv[i] = np;
    // the candidate data of the row walk, at the particle's new place
    const double x[3] = {bp.pos[i], bp.pos[bp.S + i], bp.pos[2 * bp.S + i]};
    pack_candidate(bp.G, x, bp.radius[i], bp.state[i] != 0, bp.xr, bp.xf, np);
    zero0[i] = 0;  // the row counters of the neighbour build that follows
    zero1[i] = 0;
    zero2[i] = 0;
  }
  if (i < nfine) {
    const int r = cell_rank[i];
    cstart[i] = cell_off[r];
    cend[i] = cell_off[r + 1];
    cell_cnt[r] = 0;
  }
  if (i == 0) cell_cnt[nfine] = 0;
}

struct PermPtrs {
  const int* perm;
  const double *pos, *vel, *omega;
  double *pos_o, *vel_o, *omega_o;
  const double *radius, *pmass, *omoi, *pvol, *acc, *racc;
  double *radius_o, *pmass_o, *omoi_o, *pvol_o, *acc_o, *racc_o;
  const int *gid, *fcell, *dgc;
  int *gid_o, *fcell_o, *dgc_o;
  const uint8_t *state, *phase, *wmask;
  uint8_t *state_o, *phase_o, *wmask_o;
  const int* wfid;
  int* wfid_o;
  const double* wpft;
  double* wpft_o;
  int ab2;
};

__global__ __launch_bounds__(kBS) void k_permute(int n, int S, PermPtrs p) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const int o = p.perm[i];
  for (int c = 0; c < 3; c++) {
    p.pos_o[c * S + i] = p.pos[c * S + o];
    p.vel_o[c * S + i] = p.vel[c * S + o];
    p.omega_o[c * S + i] = p.omega[c * S + o];
  }
  p.radius_o[i] = p.radius[o];
  p.pmass_o[i] = p.pmass[o];
  p.omoi_o[i] = p.omoi[o];
  p.pvol_o[i] = p.pvol[o];
  p.gid_o[i] = p.gid[o];
  p.fcell_o[i] = p.fcell[o];
  p.dgc_o[i] = p.dgc[o];
  const uint8_t st = p.state[o];
  p.state_o[i] = st == kSeeded ? kEntering : st;  // binned now: in a DES cell list
  p.phase_o[i] = p.phase[o];
  const uint8_t wm = p.wmask[o];
  p.wmask_o[i] = wm;
  for (int k = 0; k < kWallSlots; k++)
    if ((wm >> k) & 1) {
      p.wfid_o[k * S + i] = p.wfid[k * S + o];
      for (int d = 0; d < 3; d++) p.wpft_o[(k * 3 + d) * S + i] = p.wpft[(k * 3 + d) * S + o];
    }
  if (p.ab2)
    for (int c = 0; c < 3; c++) {
      p.acc_o[c * S + i] = p.acc[c * S + o];
      p.racc_o[c * S + i] = p.racc[c * S + o];
    }
}

__device__ __forceinline__ uint32_t compact10(uint64_t v) {
  v &= 0x09249249;
  v = (v | (v >> 2)) & 0x030C30C3;
  v = (v | (v >> 4)) & 0x0300F00F;
  v = (v | (v >> 8)) & 0x030000FF;
  v = (v | (v >> 16)) & 0x3ff;
  return (uint32_t)v;
}

// search-grid cell ranges of the sorted particles (from the Morton part of
// the sort key)
template <bool BATCH = false>
__global__ __launch_bounds__(kBS) void k_cell_ranges(int n, GridDev G0, const uint64_t* keys,
                                                     const uint8_t* state, int* cstart,
                                                     int* cend, BatchDev B) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || state[i] == 0) return;
  const int s = BATCH ? B.sys[i] : 0;
  const GridDev& G = sys_grid<BATCH>(G0, B, s);
  const uint64_t k = keys[i];  // system id (ensembles) and Morton code
  const bool first = (i == 0) || keys[i - 1] != k;
  const bool last = (i + 1 == n) || keys[i + 1] != k;
  const uint64_t m = k & ((1ull << G.mbits) - 1);
  const int c = (BATCH ? B.cell_off[s] : 0) + (int)compact10(m) +
                G.fn[0] * ((int)compact10(m >> 1) + G.fn[1] * (int)compact10(m >> 2));
  if (first) cstart[c] = i;
  if (last) cend[c] = i + 1;
}

struct BuildPtrs {
  const double *pos, *radius;
  const double* xr;             // [4*i + c]: x, y, z, radius (k_pack_xr), read per candidate
  const float4* xf;             // the same in single precision about GridDev::pf_c (radius
                                // < 0: outside the prefilter's range)
  const int *gid, *dgc;
  const uint8_t* state;
  const int *fstart, *fend;     // search-grid cell ranges
  int* own_cnt;
  int* own_tmp;                 // own row entry k of slot i at [k*ld + i] (slot-major:
  uint64_t* own_key;            // consecutive slots are adjacent, so row walks coalesce)
  int ld;
  int maxrow;
  int* flags;
  int* in_cnt;                  // incoming entries per particle (thread-row kernel)
  const int* idx;               // slot at each position of the search-cell order, when the
                                // binning left the slots as they were (null: the same)
};

// Own rows of DESGRID_NEIGH_BUILD (desgrid_mod.f:1057-1214): j in the
// reference stencil of i, j after i (ids), within FACTOR_RLM*(ri+rj).
// Half-stencil search: the thread of particle h tests each unordered pair
// once -- partners in its own search cell after it (storage order) and in
// the 13 neighbouring cells with a positive offset (z, then y, then x).  A
// pair within the cut-off belongs to the row of its owner (own_row_hit),
// which must be outside the ghost DES cells and before the first rogue id;
// the partner must lie in the owner's reference stencil.  Entries are
// appended to the owner's row through an atomic slot (own_cnt and in_cnt
// zeroed beforehand); k_place then puts each row in the reference's order
// (stencil lk, lj, li, then ascending id).
//
// Search cells hold about one particle, so a loop per cell would cost a
// wavefront the slowest lane's count in every one of the 14 cells.  The
// thread instead looks up all 14 cell ranges at once, keeps the non-empty
// ones in LDS and walks its candidates as one sequence, loading the next
// candidate before it tests the current one.  Candidates within the
// cut-off (about one in ten) wait in a short LDS list and go through the
// owner and stencil checks afterwards, so the wavefront runs that long path
// as often as its busiest lane has hits rather than in almost every step
// of the walk.
constexpr int kHalfCells = 14;
constexpr int kHitCap = 8;

// The owner and stencil checks of a pair within the cut-off (h, t) and the
// append to the owner's row.  Only normal particles have rows; a row takes
// every existing partner except normal ones of lower id (desgrid_mod.f:
// 1160-1175), so a pair of two normal particles belongs to the lower id and
// a pair of a normal particle with an entering or exiting one to the normal
// particle.
template <bool BATCH>
__device__ __forceinline__ void own_row_hit(const GridDev& G, const BuildPtrs& b, int S, int g_r,
                                            int h, int gh, int t) {
  const uint8_t sh = b.state[h], stt = b.state[t];
  if (stt == kNonexistent) return;
  const int gt = b.gid[t];
  if (sh != kNormal && stt != kNormal) return;
  const bool h_owns = sh == kNormal && (stt != kNormal || gh < gt);
  const int own = h_owns ? h : t, oth = h_owns ? t : h;
  const int go = h_owns ? gh : gt, gp = h_owns ? gt : gh;
  if (go >= g_r) return;
  int ci, cj, ck, li, lj, lk;
  dg_decode(G, b.dgc[own], ci, cj, ck);
  if (dg_ghost(G, ci, cj, ck)) return;
  const double xo[3] = {b.pos[own], b.pos[S + own], b.pos[2 * S + own]};
  int slo[3], shi[3];
  stencil_off(G, xo, slo, shi);
  dg_decode(G, b.dgc[oth], li, lj, lk);
  if (li < ci - slo[0] || li > ci + shi[0] || lj < cj - slo[1] || lj > cj + shi[1] ||
      lk < ck - slo[2] || lk > ck + shi[2])
    return;
  const int slot = atomicAdd(&b.own_cnt[own], 1);
  if (slot < b.maxrow) {
    b.own_tmp[(size_t)slot * b.ld + own] = oth;
    b.own_key[(size_t)slot * b.ld + own] =
        ((uint64_t)((lk * 1024 + lj) * 1024 + li) << 32) | (uint32_t)gp;
    atomicAdd(&b.in_cnt[oth], 1);
  } else {
    atomicMax(&b.flags[kFlagOwnOverflow], slot + 1);
  }
}

// A candidate that passed the walk's prefilter (t: its position in the
// search-cell order): the cut-off test of DESGRID_NEIGH_BUILD in double
// precision (desgrid_mod.f:1177-1183), then the owner and stencil checks.
// A nonexistent candidate sits far away and fails the test.
template <bool BATCH>
__device__ __forceinline__ void own_row_exact(const GridDev& G, const BuildPtrs& b, int S,
                                              int g_r, int h, int gh, const double x[3],
                                              double r, int t) {
  const double2* xr2 = reinterpret_cast<const double2*>(b.xr);
  const double2 u = xr2[2 * t], v = xr2[2 * t + 1];  // x, y | z, radius: one 32-byte line
  const double sr = G.factor_rlm * (r + v.y);
  const double d0 = x[0] - u.x, d1 = x[1] - u.y, d2 = x[2] - v.x;
  const double ds = d0 * d0 + d1 * d1 + d2 * d2;
  if (ds > sr * sr) return;
  own_row_hit<BATCH>(G, b, S, g_r, h, gh, b.idx ? b.idx[t] : t);
}

// kRowLanes lanes share one particle's half stencil (lane p takes the cells
// p, p + kRowLanes, ...): shorter candidate walks per lane, and the range
// table each lane keeps in LDS shrinks with them, so more wavefronts fit on a
// CU to hide the walk's dependent loads.
#ifndef MFIXDEM_ROW_LANES
#define MFIXDEM_ROW_LANES 2
#endif
constexpr int kRowLanes = MFIXDEM_ROW_LANES;

template <bool BATCH = false>
__global__ __launch_bounds__(kBS) void k_own_rows(int n, int S, GridDev G0, BuildPtrs b,
                                                  BatchDev B) {
  constexpr int kCells = (kHalfCells + kRowLanes - 1) / kRowLanes;
  constexpr int kHits = kRowLanes == 1 ? kHitCap : kHitCap / 2;
  __shared__ int s_lo[kCells][kBS], s_hi[kCells][kBS];
  __shared__ int s_hit[kHits][kBS];
  // XCD-aware order: the cells a row walks in the next planes of search
  // cells are read by workgroups of the same XCD, from its L2
  const int id = xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
  const int hp = id / kRowLanes, part = id % kRowLanes;  // position in the search-cell order
  if (hp >= n) return;
  const int h = b.idx ? b.idx[hp] : hp;                   // and slot
  if (b.state[h] == 0) return;
  const int s = BATCH ? B.sys[h] : 0;
  const GridDev& G = sys_grid<BATCH>(G0, B, s);
  const int cbase = BATCH ? B.cell_off[s] : 0;
  const int g_r = BATCH ? B.rogue[2 * s] : b.flags[kFlagRogueGid];
  const double x[3] = {b.pos[h], b.pos[S + h], b.pos[2 * S + h]};
  const double r = b.radius[h];
  const int gh = b.gid[h];
  int f3[3];
  for (int d = 0; d < 3; d++) {
    const int f = floor_sat_d((x[d] - G.f_lo[d]) * G.f_inv);
    f3[d] = f < 0 ? 0 : (f >= G.fn[d] ? G.fn[d] - 1 : f);
  }
  int lo[kCells], hi[kCells];
#pragma unroll
  for (int c = 0; c < kCells; c++) {
    // q = 0: own cell; q = 1..13: offsets (dx, dy, dz) after (0, 0, 0)
    const int q = part + c * kRowLanes;
    const int o = q == 0 ? 13 : 13 + q;  // index in the 3x3x3 block, x fastest
    const int fx = f3[0] + o % 3 - 1, fy = f3[1] + (o / 3) % 3 - 1, fz = f3[2] + o / 9 - 1;
    const bool in = q < kHalfCells && fx >= 0 && fx < G.fn[0] && fy >= 0 && fy < G.fn[1] &&
                    fz >= 0 && fz < G.fn[2];
    const int fc = in ? cbase + fx + G.fn[0] * (fy + G.fn[1] * fz) : cbase;
    lo[c] = q == 0 ? hp + 1 : b.fstart[fc];
    hi[c] = in ? b.fend[fc] : 0;
  }
  int m = 0;
#pragma unroll
  for (int c = 0; c < kCells; c++)
    if (hi[c] > lo[c]) {
      s_lo[m][threadIdx.x] = lo[c];
      s_hi[m][threadIdx.x] = hi[c];
      m++;
    }
  if (m == 0) return;
  // The walk tests each candidate in single precision, 16 bytes a candidate,
  // with a margin no rounding can cross; the few that pass are tested again
  // in double precision (the reference's criterion) before they count.
  const float4 xh = b.xf[hp];
  const bool coarse = xh.w >= 0.0f;  // h inside the prefilter's range
  const float frlm = (float)G.factor_rlm * (1.0f + 1.0f / 65536.0f);
  int k = 0, t = s_lo[0][threadIdx.x], t1 = s_hi[0][threadIdx.x], nh = 0;
  float4 u = b.xf[t];  // candidate t, loaded one step ahead
#pragma unroll 1
  while (k < m) {
    const int tc = t;
    const float4 uc = u;
    if (++t == t1 && ++k < m) {
      t = s_lo[k][threadIdx.x];
      t1 = s_hi[k][threadIdx.x];
    }
    const int tn = k < m ? t : tc;  // in range either way
    u = b.xf[tn];
    const float lim = frlm * (xh.w + uc.w) + G0.pf_slack;  // the grid k_pack_xr used
    const float e0 = xh.x - uc.x, e1 = xh.y - uc.y, e2 = xh.z - uc.z;
    if (coarse && uc.w >= 0.0f && e0 * e0 + e1 * e1 + e2 * e2 > lim * lim) continue;
    if (nh < kHits) s_hit[nh++][threadIdx.x] = tc;
    else own_row_exact<BATCH>(G, b, S, g_r, h, gh, x, r, tc);  // a crowded stencil: at once
  }
  if (nh == 0) return;
  // The waiting candidates, all at once: their loads issued together, the
  // owner and stencil checks with h's own cell and stencil formed once, the
  // row slots taken with atomics issued side by side.  Same rules as
  // own_row_exact and own_row_hit.
  const uint8_t sh = b.state[h];
  int hc[3], hlo[3], hhi[3];
  dg_decode(G, b.dgc[h], hc[0], hc[1], hc[2]);
  stencil_off(G, x, hlo, hhi);
  const double2* xr2 = reinterpret_cast<const double2*>(b.xr);
  int tv[kHits], own[kHits], oth[kHits], slot[kHits];
  uint64_t key[kHits];
  bool ok[kHits];
#pragma unroll
  for (int e = 0; e < kHits; e++) {
    ok[e] = false;
    if (e < nh) tv[e] = s_hit[e][threadIdx.x];
  }
#pragma unroll
  for (int e = 0; e < kHits; e++) {
    if (e >= nh) continue;
    const int tp = tv[e], t = b.idx ? b.idx[tp] : tp;  // position, slot
    const double2 u = xr2[2 * tp], v = xr2[2 * tp + 1];
    const uint8_t stt = b.state[t];
    const int gt = b.gid[t], dt = b.dgc[t];
    const double sr = G.factor_rlm * (r + v.y);
    const double d0 = x[0] - u.x, d1 = x[1] - u.y, d2 = x[2] - v.x;
    if (d0 * d0 + d1 * d1 + d2 * d2 > sr * sr) continue;  // DESGRID_NEIGH_BUILD cut-off
    if (stt == kNonexistent || (sh != kNormal && stt != kNormal)) continue;
    const bool h_owns = sh == kNormal && (stt != kNormal || gh < gt);
    if ((h_owns ? gh : gt) >= g_r) continue;
    int oc[3], lo[3], hi[3], pc[3];
    if (h_owns) {
      for (int d = 0; d < 3; d++) {
        oc[d] = hc[d];
        lo[d] = hlo[d];
        hi[d] = hhi[d];
      }
      dg_decode(G, dt, pc[0], pc[1], pc[2]);
    } else {
      dg_decode(G, dt, oc[0], oc[1], oc[2]);
      const double xt[3] = {u.x, u.y, v.x};
      stencil_off(G, xt, lo, hi);
      for (int d = 0; d < 3; d++) pc[d] = hc[d];
    }
    if (dg_ghost(G, oc[0], oc[1], oc[2])) continue;
    bool in = true;
    for (int d = 0; d < 3; d++) in = in && pc[d] >= oc[d] - lo[d] && pc[d] <= oc[d] + hi[d];
    if (!in) continue;
    ok[e] = true;
    own[e] = h_owns ? h : t;
    oth[e] = h_owns ? t : h;
    key[e] = ((uint64_t)((pc[2] * 1024 + pc[1]) * 1024 + pc[0]) << 32) |
             (uint32_t)(h_owns ? gt : gh);
  }
#pragma unroll
  for (int e = 0; e < kHits; e++)
    if (ok[e]) slot[e] = atomicAdd(&b.own_cnt[own[e]], 1);
#pragma unroll
  for (int e = 0; e < kHits; e++) {
    if (!ok[e]) continue;
    if (slot[e] < b.maxrow) {
      b.own_tmp[(size_t)slot[e] * b.ld + own[e]] = oth[e];
      b.own_key[(size_t)slot[e] * b.ld + own[e]] = key[e];
      atomicAdd(&b.in_cnt[oth[e]], 1);
    } else {
      atomicMax(&b.flags[kFlagOwnOverflow], slot[e] + 1);
    }
  }
}

// Position and radius of the particles packed in the search-cell order for
// the candidate walk of k_own_rows (one cache line per candidate instead of
// five), after a binning that did not pack them (k_bin_place does).  A
// nonexistent particle gets a position no cut-off reaches, so the walk
// needs no state load; the hit path checks the state itself.
__global__ __launch_bounds__(kBS) void k_pack_xr(int n, int S, GridDev G, const int* idx,
                                                 const double* pos, const double* radius,
                                                 const uint8_t* state, double* xr, float4* xf) {
  const int k = blockIdx.x * kBS + threadIdx.x;  // position in the search-cell order
  if (k >= n) return;
  const int i = idx ? idx[k] : k;
  const double p[3] = {pos[i], pos[S + i], pos[2 * S + i]};
  pack_candidate(G, p, radius[i], state[i] != 0, xr, xf, k);
}

// A row of k_own_rows put in the reference's order (keys are distinct: the
// partner id is part of the key), in place.  Rows are short: up to kRegRow
// entries are sorted in registers, longer ones by insertion in place.
constexpr int kRegRow = 8;
__device__ __forceinline__ void sort_own_row(int* row, uint64_t* key, size_t ld, int cnt) {
  if (cnt <= kRegRow) {
    uint64_t k[kRegRow];
    int v[kRegRow];
#pragma unroll
    for (int a = 0; a < kRegRow; a++)
      if (a < cnt) {
        k[a] = key[a * ld];
        v[a] = row[a * ld];
      }
    // insertion sort, unrolled: entry a sinks past larger keys before it
#pragma unroll
    for (int a = 1; a < kRegRow; a++)
#pragma unroll
      for (int q = a; q > 0; q--)
        if (a < cnt && k[q - 1] > k[q]) {
          const uint64_t tk = k[q - 1];
          k[q - 1] = k[q];
          k[q] = tk;
          const int tv = v[q - 1];
          v[q - 1] = v[q];
          v[q] = tv;
        }
#pragma unroll
    for (int a = 0; a < kRegRow; a++)
      if (a < cnt) row[a * ld] = v[a];
    return;
  }
  for (int a = 1; a < cnt; a++) {
    const uint64_t k = key[a * ld];
    const int v = row[a * ld];
    int q = a;
    while (q > 0 && key[(q - 1) * ld] > k) {
      key[q * ld] = key[(q - 1) * ld];
      row[q * ld] = row[(q - 1) * ld];
      q--;
    }
    key[q * ld] = k;
    row[q * ld] = v;
  }
}

// Own rows, one wavefront per search-grid cell.  The candidates of the 27
// surrounding cells are staged in LDS once per cell (DES cells already
// decoded); the home particles of the cell are walked in a uniform loop and
// the lanes take the candidates 64 at a time.  The lanes that pass the
// reference's criteria are appended to the home row in lane order with a
// ballot, the row is kept in registers (lane k holds entry k, at most 64
// per row) and finally put in the reference's order by a rank sort.  A row
// longer than 64 raises the overflow flag (the host then reruns the
// thread-per-particle kernel).
// kRowCap = candidates staged in LDS per wavefront (a tuning choice)
template <bool BATCH = false, int kRowCap = 192>
__global__ __launch_bounds__(kBS) void k_own_rows_wave(int S, GridDev G0, BuildPtrs b, int ncells,
                                                       BatchDev B) {
  __shared__ double s_x[kBS / 64][kRowCap], s_y[kBS / 64][kRowCap], s_z[kBS / 64][kRowCap],
      s_r[kBS / 64][kRowCap];
  __shared__ int s_j[kBS / 64][kRowCap], s_g[kBS / 64][kRowCap], s_d[kBS / 64][kRowCap];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int c = (int)(((unsigned)xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x) >> 6);
  if (c >= ncells) return;                       // whole wavefront leaves together
  const int h0 = b.fstart[c], h1 = b.fend[c];
  if (h1 <= h0) return;
  const int s = BATCH ? cell_system(B, c) : 0;
  const GridDev& G = sys_grid<BATCH>(G0, B, s);
  const int cbase = BATCH ? B.cell_off[s] : 0;
  const int lc = c - cbase;
  const int fx = lc % G.fn[0], fy = (lc / G.fn[0]) % G.fn[1], fz = lc / (G.fn[0] * G.fn[1]);
  // lane k < 27 holds neighbour cell k (x fastest): its first slot and the
  // inclusive prefix of the candidate counts
  int ns = 0, nc = 0;
  if (lane < 27) {
    const int x = fx + lane % 3 - 1, y = fy + (lane / 3) % 3 - 1, z = fz + lane / 9 - 1;
    if (x >= 0 && x < G.fn[0] && y >= 0 && y < G.fn[1] && z >= 0 && z < G.fn[2]) {
      const int cc = cbase + x + G.fn[0] * (y + G.fn[1] * z);
      ns = b.fstart[cc];
      nc = b.fend[cc] - ns;
    }
  }
  int pre = nc;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    const int y = __shfl_up(pre, o, 64);
    if (lane >= o) pre += y;
  }
  const int total = __shfl(pre, 26, 64);
  // slot of candidate t: first neighbour cell k with prefix > t (every lane
  // takes part in the shuffles)
  auto slot_of = [&](int t) {
    const int tq = t < total ? t : total - 1;
    int k = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1) {
      const int q = k + step - 1 < 26 ? k + step - 1 : 26;
      const int pq = __shfl(pre, q, 64);
      if (k + step <= 26 && pq <= tq) k += step;
    }
    const int before = __shfl(pre - nc, k, 64);
    const int cs = __shfl(ns, k, 64);
    return cs + (tq - before);
  };
  // stage the first kRowCap candidates; a nonexistent one gets id INT_MIN
  // so that it never passes the id test
  const int staged = total < kRowCap ? total : kRowCap;
  for (int base = 0; base < staged; base += 64) {
    const int t = base + lane;
    const int j = slot_of(t);
    if (t < staged) {
      const uint8_t st = b.state[j];
      const int g = b.gid[j];
      s_x[wv][t] = b.pos[j];
      s_y[wv][t] = b.pos[S + j];
      s_z[wv][t] = b.pos[2 * S + j];
      s_r[wv][t] = b.radius[j];
      int li, lj, lk;
      dg_decode(G, b.dgc[j], li, lj, lk);
      s_d[wv][t] = (lk * 1024 + lj) * 1024 + li;
      s_j[wv][t] = j;
      s_g[wv][t] = st != 0 ? g : INT_MIN;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int g_r = BATCH ? B.rogue[2 * s] : b.flags[kFlagRogueGid];
  for (int h = h0; h < h1; h++) {
    if (b.state[h] != 1) {
      if (lane == 0) b.own_cnt[h] = 0;
      continue;
    }
    const int gh = b.gid[h];
    int ci, cj, ck;
    dg_decode(G, b.dgc[h], ci, cj, ck);
    if (dg_ghost(G, ci, cj, ck) || gh >= g_r) {
      if (lane == 0) b.own_cnt[h] = 0;
      continue;
    }
    const double x[3] = {b.pos[h], b.pos[S + h], b.pos[2 * S + h]};
    const double r = b.radius[h];
    int lo[3], hi[3];
    stencil_off(G, x, lo, hi);
    int cnt = 0;
    int row_j = -1;
    uint64_t row_key = ~0ull;
    for (int base = 0; base < total; base += 64) {
      const int t = base + lane;
      int j = -1, gj = INT_MIN, dj = 0;
      double xj = 0.0, yj = 0.0, zj = 0.0, rj = 0.0;
      if (base < kRowCap) {
        if (t < staged) {
          j = s_j[wv][t];
          gj = s_g[wv][t];
          xj = s_x[wv][t];
          yj = s_y[wv][t];
          zj = s_z[wv][t];
          rj = s_r[wv][t];
          dj = s_d[wv][t];
        }
      } else {
        const int js = slot_of(t);
        if (t < total) {
          j = js;
          gj = b.state[j] != 0 ? b.gid[j] : INT_MIN;
          xj = b.pos[j];
          yj = b.pos[S + j];
          zj = b.pos[2 * S + j];
          rj = b.radius[j];
          int li, lj, lk;
          dg_decode(G, b.dgc[j], li, lj, lk);
          dj = (lk * 1024 + lj) * 1024 + li;
        }
      }
      bool pass = false;
      uint64_t key = 0;
      if (j != h && gj > gh) {
        const double sr = G.factor_rlm * (r + rj);
        const double d0 = x[0] - xj, d1 = x[1] - yj, d2 = x[2] - zj;
        const double ds = d0 * d0 + d1 * d1 + d2 * d2;
        if (!(ds > sr * sr)) {
          const int li = dj & 1023, lj = (dj >> 10) & 1023, lk = dj >> 20;
          if (li >= ci - lo[0] && li <= ci + hi[0] && lj >= cj - lo[1] && lj <= cj + hi[1] &&
              lk >= ck - lo[2] && lk <= ck + hi[2]) {
            pass = true;
            key = ((uint64_t)(uint32_t)dj << 32) | (uint32_t)gj;
          }
        }
      }
      // append the passing lanes in lane order: lane cnt+k takes the k-th
      uint64_t mask = __ballot(pass);
      while (mask) {
        const int src = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int jj = __builtin_amdgcn_readlane(j, src);
        const uint32_t klo = __builtin_amdgcn_readlane((int)(uint32_t)key, src);
        const uint32_t khi = __builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), src);
        if (lane == cnt) {
          row_j = jj;
          row_key = ((uint64_t)khi << 32) | klo;
        }
        cnt++;
      }
    }
    if (cnt > 64) {
      if (lane == 0) atomicMax(&b.flags[kFlagOwnOverflow], 1 << 20);
      continue;
    }
    if (cnt > b.maxrow) {
      if (lane == 0) atomicMax(&b.flags[kFlagOwnOverflow], cnt);
      continue;
    }
    // rank sort by key (keys are distinct: the id is part of the key)
    int rank = 0;
    for (int m = 0; m < cnt; m++) {
      const uint32_t klo = __builtin_amdgcn_readlane((int)(uint32_t)row_key, m);
      const uint32_t khi = __builtin_amdgcn_readlane((int)(uint32_t)(row_key >> 32), m);
      const uint64_t km = ((uint64_t)khi << 32) | klo;
      rank += (km < row_key) ? 1 : 0;
    }
    if (lane < cnt) {
      b.own_tmp[(size_t)rank * b.ld + h] = row_j;
      b.own_key[(size_t)rank * b.ld + h] = row_key;
    }
    if (lane == 0) b.own_cnt[h] = cnt;
  }
}

// Own rows of NSQUARE (nsquare.f:15-142, no periodic walls): every existing
// particle against every existing particle with a larger id, DIST < R_LM**2
// (strict), rows in ascending id.  Brute force over LDS tiles of candidates.
__global__ __launch_bounds__(kBS) void k_own_rows_nsquare(int n, int S, double factor_rlm,
                                                          BuildPtrs b) {
  __shared__ double s_x[kBS], s_y[kBS], s_z[kBS], s_r[kBS];
  __shared__ int s_g[kBS];
  const int i = blockIdx.x * kBS + threadIdx.x;
  const bool home = i < n && b.state[i] != 0;
  double xi[3] = {0.0, 0.0, 0.0}, ri = 0.0;
  int gi = INT_MAX;
  if (home) {
    xi[0] = b.pos[i];
    xi[1] = b.pos[S + i];
    xi[2] = b.pos[2 * S + i];
    ri = b.radius[i];
    gi = b.gid[i];
  }
  int cnt = 0;
  int* row = b.own_tmp + (home ? i : 0);
  uint64_t* key = b.own_key + (home ? i : 0);
  const size_t ld = b.ld;
  for (int base = 0; base < n; base += kBS) {
    const int j = base + threadIdx.x;
    if (j < n) {
      s_x[threadIdx.x] = b.pos[j];
      s_y[threadIdx.x] = b.pos[S + j];
      s_z[threadIdx.x] = b.pos[2 * S + j];
      s_r[threadIdx.x] = b.radius[j];
      s_g[threadIdx.x] = b.state[j] != 0 ? b.gid[j] : INT_MIN;
    } else {
      s_g[threadIdx.x] = INT_MIN;
    }
    __syncthreads();
    if (home) {
      const int m = min(kBS, n - base);
      for (int t = 0; t < m; t++) {
        const int gj = s_g[t];
        if (gj <= gi) continue;
        double r_lm = ri + s_r[t];
        r_lm = factor_rlm * r_lm;
        const double d0 = s_x[t] - xi[0], d1 = s_y[t] - xi[1], d2 = s_z[t] - xi[2];
        const double dist = d0 * d0 + d1 * d1 + d2 * d2;
        if (!(dist < r_lm * r_lm)) continue;
        if (cnt < b.maxrow) {
          const uint64_t sk = (uint64_t)(uint32_t)gj;
          int q = cnt;
          while (q > 0 && key[(q - 1) * ld] > sk) {
            key[q * ld] = key[(q - 1) * ld];
            row[q * ld] = row[(q - 1) * ld];
            q--;
          }
          key[q * ld] = sk;
          row[q * ld] = base + t;
        }
        cnt++;
      }
    }
    __syncthreads();
  }
  if (i < n) {
    if (cnt > b.maxrow) atomicMax(&b.flags[kFlagOwnOverflow], cnt);
    b.own_cnt[i] = home ? cnt : 0;
  }
}

__global__ __launch_bounds__(kBS) void k_count_in(int n, const int* own_cnt, const int* own_tmp,
                                                  int maxrow, int* in_cnt) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const int c = own_cnt[i];
  for (int k = 0; k < c; k++) atomicAdd(&in_cnt[own_tmp[(size_t)k * n + i]], 1);
}

__global__ __launch_bounds__(kBS) void k_add_counts(int n, const int* a, const int* b, int* o) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) o[i] = a[i] + b[i];
}

// history carry-over (neighbour.f:77-117), matched by (owner id, neighbour id)
struct CarryPtrs {
  int* flags;
  const int *old_start, *old_split, *old_nbr, *old_pair_base;
  const double* pft_old;
  double* pft_new;
  const uint8_t* pfl_old;       // history-set flags of pft_old / pft_new
  uint8_t* pfl_new;
};

// Placement launched before the host has read the build's flags: it runs
// only if the rows did not overflow and the lists and histories fit the
// buffers (otherwise the host sees the same flags and places again).
struct PlaceGuard {
  const int* flags;  // null: no guard
  int maxrow;
  int64_t cap_entries, cap_pairs;
};
__device__ __forceinline__ bool place_ok(const PlaceGuard& g) {
  if (!g.flags) return true;
  const int64_t pairs = g.flags[kFlagPairs];
  return g.flags[kFlagOwnOverflow] <= g.maxrow && 2 * pairs <= g.cap_entries &&
         pairs <= g.cap_pairs;
}

// Places the own entries of each row (put in the reference's order first
// when they were appended through atomic slots: own_key given) and
// scatters the transposed (incoming) entries through atomic slots.  Every
// new pair starts with no history; k_carry then brings over the histories
// there are.
__global__ __launch_bounds__(kBS) void k_place(int n, const int* own_cnt, int* own_tmp,
                                               uint64_t* own_key, const int* pair_base,
                                               const int* nbr_start, const int* in_cnt,
                                               int* nbr_split, int* in_cur, int* nbr,
                                               int* nbr_pid, uint8_t* pfl_new,
                                               PlaceGuard pg) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || !place_ok(pg)) return;
  const int split = nbr_start[i] + in_cnt[i];
  nbr_split[i] = split;
  const int cnt = own_cnt[i];
  if (cnt == 0) return;
  if (own_key && cnt > 1) sort_own_row(own_tmp + i, own_key + i, (size_t)n, cnt);
  const int pb = pair_base[i];
  if (cnt <= kRegRow) {
    // a short row: every load, atomic and store of the row issued side by
    // side rather than one entry after another
    int jv[kRegRow], nj[kRegRow], sl[kRegRow];
#pragma unroll
    for (int k = 0; k < kRegRow; k++)
      if (k < cnt) jv[k] = own_tmp[(size_t)k * n + i];
#pragma unroll
    for (int k = 0; k < kRegRow; k++)
      if (k < cnt) {
        nj[k] = nbr_start[jv[k]];
        sl[k] = atomicAdd(&in_cur[jv[k]], 1);
      }
#pragma unroll
    for (int k = 0; k < kRegRow; k++)
      if (k < cnt) {
        nbr[split + k] = jv[k];
        nbr_pid[split + k] = pb + k;
        nbr[nj[k] + sl[k]] = i;
        nbr_pid[nj[k] + sl[k]] = pb + k;
        pfl_new[pb + k] = 0;
      }
    return;
  }
  for (int k = 0; k < cnt; k++) {
    const int j = own_tmp[(size_t)k * n + i];
    nbr[split + k] = j;
    nbr_pid[split + k] = pb + k;
    const int slot = atomicAdd(&in_cur[j], 1);
    nbr[nbr_start[j] + slot] = i;
    nbr_pid[nbr_start[j] + slot] = pb + k;
    pfl_new[pb + k] = 0;
  }
}

// NEIGHBOUR's history carry-over (neighbour.f:67-126): a new pair takes the
// history of the same pair (same owner, same partner) in the owner's old
// row, or starts at zero.  Only old pairs that hold a history (flag set:
// the pairs in contact, a small share of the list) have anything to give,
// so the carry starts from them: the thread of each old row finds each of
// its flagged pairs in the owner's new row and copies the history there.
// Slots map through inv (old slot -> new slot; null: the slots stayed).
// A pair whose owner changed is not in the owner's new row and, as in the
// reference, starts at zero.
__global__ __launch_bounds__(kBS) void k_carry(int n, const int* inv, CarryPtrs c,
                                               const int* pair_base, const int* nbr_split,
                                               const int* nbr_start, const int* nbr,
                                               PlaceGuard pg) {
  const int o = blockIdx.x * kBS + threadIdx.x;  // the owner's old slot
  if (!place_ok(pg)) return;
  if (o == 0) {  // the host has read this build's flags: the next build's start values
    c.flags[kFlagRogueGid] = INT_MAX;
    c.flags[kFlagRogueCount] = 0;
    c.flags[kFlagOwnOverflow] = 0;
  }
  if (o >= n) return;
  const int os = c.old_split[o], oe = c.old_start[o + 1];
  if (os == oe) return;
  const int opb = c.old_pair_base[o];
  int i = -1, s0 = 0, s1 = 0, pb = 0;
  for (int r = os; r < oe; r++) {
    const int op = opb + (r - os);
    if (!c.pfl_old[op]) continue;
    if (i < 0) {
      i = inv ? inv[o] : o;
      s0 = nbr_split[i];
      s1 = nbr_start[i + 1];
      pb = pair_base[i];
    }
    const int jo = c.old_nbr[r];
    const int j = inv ? inv[jo] : jo;
    for (int q = s0; q < s1; q++)
      if (nbr[q] == j) {
        const int p = pb + (q - s0);
        c.pfl_new[p] = 1;
        for (int d = 0; d < 3; d++) c.pft_new[3 * p + d] = c.pft_old[3 * op + d];
        break;
      }
  }
}

// incoming entries in ascending owner id (short rows in registers)
__global__ __launch_bounds__(kBS) void k_sort_in(int n, const int* gid, const int* nbr_start,
                                                 const int* nbr_split, int* nbr, int* nbr_pid,
                                                 PlaceGuard pg) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || !place_ok(pg)) return;
  const int s = nbr_start[i], e = nbr_split[i], cnt = e - s;
  if (cnt < 2) return;
  if (cnt <= kRegRow) {
    int g[kRegRow], v[kRegRow], pv[kRegRow];
#pragma unroll
    for (int a = 0; a < kRegRow; a++)
      if (a < cnt) {
        v[a] = nbr[s + a];
        pv[a] = nbr_pid[s + a];
      }
#pragma unroll
    for (int a = 0; a < kRegRow; a++)
      if (a < cnt) g[a] = gid[v[a]];
#pragma unroll
    for (int a = 1; a < kRegRow; a++)
#pragma unroll
      for (int q = a; q > 0; q--)
        if (a < cnt && g[q - 1] > g[q]) {
          int t = g[q - 1]; g[q - 1] = g[q]; g[q] = t;
          t = v[q - 1]; v[q - 1] = v[q]; v[q] = t;
          t = pv[q - 1]; pv[q - 1] = pv[q]; pv[q] = t;
        }
#pragma unroll
    for (int a = 0; a < kRegRow; a++)
      if (a < cnt) {
        nbr[s + a] = v[a];
        nbr_pid[s + a] = pv[a];
      }
    return;
  }
  for (int a = s + 1; a < e; a++) {
    const int j = nbr[a], pj = nbr_pid[a], gj = gid[j];
    int q = a;
    while (q > s && gid[nbr[q - 1]] > gj) {
      nbr[q] = nbr[q - 1];
      nbr_pid[q] = nbr_pid[q - 1];
      q--;
    }
    nbr[q] = j;
    nbr_pid[q] = pj;
  }
}

// DELETE_PARTICLE of rogue particles (mass_outflow_dem.f:201-280)
template <bool BATCH = false>
__global__ __launch_bounds__(kBS) void k_delete_rogue(int n, int S, GridDev G0, const int* dgc,
                                                      uint8_t* state, uint8_t* wmask, double* pos,
                                                      double* vel, double* omega, double* radius,
                                                      BatchDev B) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || state[i] != 1) return;
  const GridDev& G = sys_grid<BATCH>(G0, B, BATCH ? B.sys[i] : 0);
  int ci, cj, ck;
  dg_decode(G, dgc[i], ci, cj, ck);
  if (!dg_ghost(G, ci, cj, ck)) return;
  state[i] = 0;
  wmask[i] = 0;
  for (int c = 0; c < 3; c++) {
    pos[c * S + i] = 0.0;
    vel[c * S + i] = 0.0;
    omega[c * S + i] = 0.0;
  }
  radius[i] = 0.0;
}

// ------------------------------------------------- DEM inlets / outlets
struct BcPtrs {
  double *pos, *vel, *omega, *radius, *pmass, *omoi, *pvol;
  const int *gid, *dgc;
  uint8_t *state, *wmask;
  const uint8_t* phase;
  const BcDev* bcs;
  int nbc, n, S;
  int* del_cnt;   // deleted particles: count and global ids
  int* del_list;
};

// The checks MASS_INFLOW_DEM (mass_inflow_dem.f:68-139) and MASS_OUTFLOW_DEM
// (mass_outflow_dem.f:13-188) make of the particles in each boundary's DES
// cells, one particle per thread.  The reference walks the boundaries in
// turn and the cell lists inside each; every check reads and writes only
// its own particle, so a thread taking its particle through the boundaries
// in the same order (inlets first, as MASS_INFLOW_DEM runs first) ends in
// the same state.  A particle is in a cell list once binned: seeded ones
// (kSeeded) are not.
//   inlet:  ENTERING -> NORMAL once the centre is a radius inside;
//   outlet: NORMAL beyond a radius inside; EXITING between (MASS_OUTFLOW
//           gives it the outlet's velocity, other outlets keep only its
//           outward normal velocity, and only while it moves outwards);
//           deleted (DELETE_PARTICLE, :201-280) once the centre is out.
__global__ __launch_bounds__(kBS) void k_dem_bcs(BcPtrs b) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= b.n) return;
  const int st0 = b.state[i];
  if (st0 == kNonexistent || st0 == kSeeded) return;
  const int S = b.S, dc = b.dgc[i];
  int st = st0;
  double x[3], v[3];
  load3(b.pos, S, i, x);
  load3(b.vel, S, i, v);
  const double r = b.radius[i];
  bool vset = false;
  for (int k = 0; k < b.nbc; k++) {
    const BcDev& B = b.bcs[k];
    if (B.type != MFIXDEM_BC_MASS_INFLOW || !bc_holds(B, dc)) continue;
    if (st == kExiting) continue;
    if (bc_dist(B, x) > r && st == kEntering) st = kNormal;
  }
  for (int k = 0; k < b.nbc; k++) {
    const BcDev& B = b.bcs[k];
    if (B.type == MFIXDEM_BC_MASS_INFLOW || !bc_holds(B, dc)) continue;
    if (st == kNonexistent || st == kEntering) continue;
    const double dist = bc_dist(B, x);
    if (dist > r) {
      st = kNormal;
    } else if (dist > 0.0) {
      if (B.type == MFIXDEM_BC_MASS_OUTFLOW) {
        const int m = b.phase[i];
        for (int c = 0; c < 3; c++) v[c] = B.vel[m][c];
        st = kExiting;
        vset = true;
      } else if (v[B.ax] * (B.low ? -1.0 : 1.0) > 0.0) {
        for (int c = 0; c < 3; c++) v[c] = v[c] * (c == B.ax ? 1.0 : 0.0);
        st = kExiting;
        vset = true;
      }
    } else {
      st = kNonexistent;
    }
  }
  if (st == kNonexistent) {
    const double z[3] = {0.0, 0.0, 0.0};
    store3(b.pos, S, i, z);
    store3(b.vel, S, i, z);
    store3(b.omega, S, i, z);
    b.radius[i] = 0.0;
    b.pmass[i] = 3.4028234663852886e38;  // HUGE(0.0) as DELETE_PARTICLE sets it
    b.pvol[i] = 3.4028234663852886e38;
    b.omoi[i] = 0.0;
    b.wmask[i] = 0;
    b.state[i] = kNonexistent;
    b.del_list[atomicAdd(b.del_cnt, 1)] = b.gid[i];
    return;
  }
  if (st != st0) b.state[i] = (uint8_t)st;
  if (vset) store3(b.vel, S, i, v);
}

struct SeedPtrs {
  double *pos, *vel, *omega, *radius, *pmass, *omoi, *pvol;
  int *dgc, *fcell;
  uint8_t *state, *phase, *wmask;
  const int* slot_of_gid;
  int S;
};
// SET_NEW_PARTICLE_PROPS (mass_inflow_dem.f:419-642): the seeded particles
// written into the slots of their global ids, ENTERING and not yet binned.
__global__ __launch_bounds__(kBS) void k_seed(int ns, const SeedDev* sd, SeedPtrs p) {
  const int k = blockIdx.x * kBS + threadIdx.x;
  if (k >= ns) return;
  const SeedDev d = sd[k];
  const int i = p.slot_of_gid[d.gid], S = p.S;
  const double z[3] = {0.0, 0.0, 0.0};
  store3(p.pos, S, i, d.pos);
  store3(p.vel, S, i, d.vel);
  store3(p.omega, S, i, z);
  p.radius[i] = d.radius;
  p.pvol[i] = d.pvol;
  p.pmass[i] = d.pmass;
  p.omoi[i] = d.omoi;
  p.dgc[i] = d.dgc;
  p.fcell[i] = d.fcell;
  p.phase[i] = (uint8_t)d.phase;
  p.wmask[i] = 0;
  p.state[i] = kSeeded;
}

// the inverse of a binning's order (ensembles, whose key sort gives the
// order only)
__global__ __launch_bounds__(kBS) void k_invert(int n, const int* perm, int* inv) {
  const int k = blockIdx.x * kBS + threadIdx.x;
  if (k < n) inv[perm[k]] = k;
}

// slot of each global id, after a binning has reordered the slots
__global__ __launch_bounds__(kBS) void k_slot_of_gid(int n, const int* gid, int* slot_of_gid) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) slot_of_gid[gid[i]] = i;
}

// ---------------------------------------------------------- gas coupling
__global__ __launch_bounds__(kBS) void k_locate(int n, int S, int ni, int nj, int nk,
                                                const double* xe, const double* yn,
                                                const double* zt, const double* pos,
                                                const uint8_t* state, int* fcell) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  if (state[i] == 0) {
    fcell[i] = -1;
    return;
  }
  const int I = locate_cell(xe, ni, 1.0 / (xe[1] - xe[0]), pos[i]);
  const int J = locate_cell(yn, nj, 1.0 / (yn[1] - yn[0]), pos[S + i]);
  const int K = locate_cell(zt, nk, 1.0 / (zt[1] - zt[0]), pos[2 * S + i]);
  fcell[i] = (I && J && K) ? pack_ijk(I, J, K) : -1;
}

// PIC counts; deletes particles outside the fluid grid (particles_in_cell.f:175-184)
__global__ __launch_bounds__(kBS) void k_pic_count(int n, int S, int ni, int nj, const int* fcell,
                                                   uint8_t* state, uint8_t* wmask, double* pos,
                                                   double* vel, double* omega, double* radius,
                                                   int* cnt, int* flags) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || state[i] == 0) return;
  const int pc = fcell[i];
  if (pc < 0) {
    state[i] = 0;
    wmask[i] = 0;
    for (int d = 0; d < 3; d++) {
      pos[d * S + i] = 0.0;
      vel[d * S + i] = 0.0;
      omega[d * S + i] = 0.0;
    }
    radius[i] = 0.0;
    atomicAdd(&flags[kFlagOutOfDomain], 1);
    return;
  }
  int I, J, K;
  unpack_ijk(pc, I, J, K);
  atomicAdd(&cnt[lin_ijk(I, J, K, ni, nj)], 1);
}

__global__ __launch_bounds__(kBS) void k_pic_scatter(int n, int ni, int nj, const int* fcell,
                                                     const uint8_t* state, const int* off,
                                                     int* cur, int* pic) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || state[i] == 0) return;
  int I, J, K;
  unpack_ijk(fcell[i], I, J, K);
  const int c = lin_ijk(I, J, K, ni, nj);
  const int s = atomicAdd(&cur[c], 1);
  pic[off[c] + s] = i;
}

// PIC(IJK)%p in ascending particle id (serial fill order of the reference)
__global__ __launch_bounds__(kBS) void k_pic_sort(int ncell, const int* off, const int* gid,
                                                  int* pic) {
  const int c = blockIdx.x * kBS + threadIdx.x;
  if (c >= ncell) return;
  const int s = off[c], e = off[c + 1];
  for (int a = s + 1; a < e; a++) {
    const int p = pic[a], g = gid[p];
    int q = a;
    while (q > s && gid[pic[q - 1]] > g) {
      pic[q] = pic[q - 1];
      q--;
    }
    pic[q] = p;
  }
}

// COMP_MEAN_FIELDS0 particle -> node deposition (comp_mean_fields0.f:167-210),
// gathered per node in the reference's summation order: the 8 cells sharing
// the node in ascending IJK (K, then I, then J), particles in ascending id.
__global__ __launch_bounds__(kBS) void k_deposit_nodes(int imax, int jmax, int kmax, int ni,
                                                       int nj, int nph, int S,
                                                       const double* xe, const double* yn,
                                                       const double* zt, const uint8_t* fluid,
                                                       const int* off, const int* pic,
                                                       const double* pos, const double* pmass,
                                                       const uint8_t* phase, double* rops_node,
                                                       int ncell) {
  const int t = blockIdx.x * kBS + threadIdx.x;
  const int nI = imax + 1, nJ = jmax + 1;
  if (t >= nI * nJ * (kmax + 1)) return;
  const int J = t % nJ + 1;
  const int I = (t / nJ) % nI + 1;
  const int K = t / (nJ * nI) + 1;
  const int node = (J - 1) + (I - 1) * nj + (K - 1) * nj * ni;
  double sum[kMaxPh] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < 2; c++)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const int ci = I + a, cj = J + b, ck = K + c;
        const int cell = (cj - 1) + (ci - 1) * nj + (ck - 1) * nj * ni;
        if (!fluid[cell]) continue;
        const int s = off[cell], e = off[cell + 1];
        if (s == e) continue;
        const double x0 = xe[ci - 1], y0 = yn[cj - 1], z0 = zt[ck - 1];
        const double dxx = xe[ci] - x0, dyy = yn[cj] - y0, dzz = zt[ck] - z0;
        for (int q = s; q < e; q++) {
          const int p = pic[q];
          double zx = pos[p] - x0, zy = pos[S + p] - y0;
          zx = zx / dxx;
          zy = zy / dyy;
          double zz = pos[2 * S + p] - z0;
          zz = zz / dzz;
          // this node is corner (1-a, 1-b, 1-c) of the cell
          const double xv = a ? 1.0 - zx : zx;
          const double yv = b ? 1.0 - zy : zy;
          const double zv = c ? 1.0 - zz : zz;
          const double w = xv * yv * zv;
          const int m = phase[p];
          sum[m] = sum[m] + w * pmass[p];
        }
      }
  for (int m = 0; m < nph; m++) rops_node[(size_t)m * ncell + node] = sum[m];
}

// DRAG_GS_GAS0 (drag_gs_des0.f:235-497), the gas side of the GARG_2012 drag,
// in three passes.  k_gas_fgp: each particle's drag coefficient F_GP
// (DES_DRAG_GP with the gas velocity interpolated from the node velocities
// and EP_G of its fluid cell, as on the particle side) into a 64-byte record
// rec[8 i ..]: the normalised position in its cell (zx, zy, zz), F_GP, and
// F_GP * v_p (the first product of the reference's DRAG_BM term); all 0
// for a particle outside a fluid cell or removed, which adds nothing.
constexpr int kGasRec = 8;
__global__ __launch_bounds__(kBS) void k_gas_fgp(DevParams P, StepPtrs a, double* rec) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= P.n) return;
  const int S = P.stride;
  const int pc = a.fcell[i];
  double r[kGasRec] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (a.state[i] != 0 && pc >= 0) {
    int ci, cj, ck;
    unpack_ijk(pc, ci, cj, ck);
    const int c = lin_ijk(ci, cj, ck, P.ni, P.nj);
    if (a.fluid[c]) {
      double pos[3], vel[3];
      load3(a.pos, S, i, pos);
      load3(a.vel, S, i, vel);
      const double x0 = a.xe[ci - 1], y0 = a.yn[cj - 1], z0 = a.zt[ck - 1];
      const double dxx = a.xe[ci] - x0, dyy = a.yn[cj] - y0, dzz = a.zt[ck] - z0;
      double zx = pos[0] - x0, zy = pos[1] - y0;
      zx = zx / dxx;
      zy = zy / dyy;
      double zz = pos[2] - z0;
      zz = zz / dzz;
      const double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
      double velfp[3] = {0.0, 0.0, 0.0};
      const int base = (cj - 2) + (ci - 2) * P.nj + (ck - 2) * P.nj * P.ni;
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++) {
            const double w = xv[ii] * yv[jj] * zv[kk];
            const double* vn = a.vnode + 3 * (base + jj + ii * P.nj + kk * P.nj * P.ni);
            for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + vn[d] * w;
          }
      const double vs[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
      const double vrel = sqrt(dot3(vs, vs));
      const double f = drag_fgp_fp64(P.drag_type, a.ep_g[c], P.mu_g0, P.ro_g0, vrel,
                                     2.0 * a.radius[i], a.pvol[i]);
      r[0] = zx;
      r[1] = zy;
      r[2] = zz;
      r[3] = f;
      for (int d = 0; d < 3; d++) r[4 + d] = f * vel[d];
    }
  }
  double2* o = reinterpret_cast<double2*>(rec) + (kGasRec / 2) * (size_t)i;
  for (int k = 0; k < kGasRec / 2; k++) o[k] = make_double2(r[2 * k], r[2 * k + 1]);
}

// k_gas_nodes: DRAG_AM and DRAG_BM of every node, gathered in the
// reference's order -- the particle's cells in ascending IJK (K, then I,
// then J fastest), each cell's particles in PIC order (ascending id) --
// so each node's sum is the reference's: F_GP*w*OVOL and (F_GP*v)*w*OVOL
// with OVOL = 1/DES_VOL_NODE, w from the particle's record.  Nodes outside
// [1, imax+1] x [1, jmax+1] x [1, kmax+1] get 0.
__global__ __launch_bounds__(kBS) void k_gas_nodes(int imax, int jmax, int kmax, int ni, int nj,
                                                   int nk, const uint8_t* fluid, const int* off,
                                                   const int* pic, const double* rec,
                                                   const double* vol_node, double* drag_am,
                                                   double* drag_bm) {
  const int node = blockIdx.x * kBS + threadIdx.x;
  const int ncell = ni * nj * nk;
  if (node >= ncell) return;
  const int J = node % nj + 1;
  const int I = (node / nj) % ni + 1;
  const int K = node / (nj * ni) + 1;
  double am = 0.0, bm[3] = {0.0, 0.0, 0.0};
  if (I <= imax + 1 && J <= jmax + 1 && K <= kmax + 1) {
    const double ovol = 1.0 / vol_node[node];
    for (int c = 0; c < 2; c++)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          const int ci = I + a, cj = J + b, ck = K + c;
          const int cell = (cj - 1) + (ci - 1) * nj + (ck - 1) * nj * ni;
          if (!fluid[cell]) continue;
          const int s0 = off[cell], e0 = off[cell + 1];
          for (int q = s0; q < e0; q++) {
            const double2* r = reinterpret_cast<const double2*>(rec) + (kGasRec / 2) *
                                                                        (size_t)pic[q];
            const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            // this node is corner (1-a, 1-b, 1-c) of the cell
            const double xv = a ? 1.0 - r0.x : r0.x;
            const double yv = b ? 1.0 - r0.y : r0.y;
            const double zv = c ? 1.0 - r1.x : r1.x;
            const double w = xv * yv * zv;
            am = am + r1.y * w * ovol;
            bm[0] = bm[0] + r2.x * w * ovol;
            bm[1] = bm[1] + r2.y * w * ovol;
            bm[2] = bm[2] + r3.x * w * ovol;
          }
        }
  }
  drag_am[node] = am;
  for (int d = 0; d < 3; d++) drag_bm[(size_t)d * ncell + node] = bm[d];
}

// DRAG_GS_GAS1 (drag_gs_des1.f:179-301), the gas side of the drag for the
// filter schemes: each normal particle in a fluid cell interpolates EP_G and
// the cell-centre gas velocity with its filter weights (as add_drag_des1
// does), takes F_GP from DES_DRAG_GP, and adds F_GP*w/VOL to F_GDS and
// (F_GP*v_p)*w/VOL to DRAG_BM of each filter cell (its own cell with weight
// 1 without DES_INTERP_ON).  The reference adds with OpenMP atomics, in no
// fixed order; so does this kernel, with FP64 atomics, so the cell sums
// equal the reference's to rounding.
__global__ __launch_bounds__(kBS) void k_gas_filter(DevParams P, StepPtrs a, const double* vol,
                                                    int ncell, double* f_gds, double* drag_bm) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= P.n || a.state[i] != 1) return;  // IS_NORMAL
  const int pc = a.fcell[i];
  if (pc < 0) return;
  int ci, cj, ck;
  unpack_ijk(pc, ci, cj, ck);
  const int c0 = lin_ijk(ci, cj, ck, P.ni, P.nj);
  if (!a.fluid[c0]) return;
  const int S = P.stride;
  double pos[3], vel[3];
  load3(a.pos, S, i, pos);
  load3(a.vel, S, i, vel);
  const bool filt = P.interp_on && filter_scheme(P.interp);
  double lepg = 0.0, velfp[3] = {0.0, 0.0, 0.0};
  if (filt) {
    for_filter(P, a.xe, a.yn, a.zt, a.csz, a.fluid, ci, cj, ck, pos, [&](int c, double w) {
      lepg = lepg + a.ep_g[c] * w;
      for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + a.ugc[3 * c + d] * w;
    });
  } else {
    lepg = a.ep_g[c0];
    for (int d = 0; d < 3; d++) velfp[d] = a.ugc[3 * c0 + d];
  }
  if (lepg < 1.0e-6) lepg = a.ep_g[c0];
  const double vs[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
  const double vrel = sqrt(dot3(vs, vs));
  const double f = drag_fgp_fp64(P.drag_type, lepg, P.mu_g0, P.ro_g0, vrel, 2.0 * a.radius[i],
                                 a.pvol[i]);
  if (a.f_gp) a.f_gp[a.gid[i]] = f;  // F_GP(NP) for explicit coupling
  const double lbm[3] = {f * vel[0], f * vel[1], f * vel[2]};
  auto add = [&](int c, double wt) {
    for (int d = 0; d < 3; d++) atomicAdd(&drag_bm[(size_t)d * ncell + c], lbm[d] * wt);
    atomicAdd(&f_gds[c], f * wt);
  };
  if (filt)
    for_filter(P, a.xe, a.yn, a.zt, a.csz, a.fluid, ci, cj, ck, pos,
               [&](int c, double w) { add(c, w / vol[c]); });
  else
    add(c0, 1.0 / vol[c0]);
}

// DES_TIME_INIT with explicit coupling, per normal particle in a fluid cell:
// DRAG_FC = F_GP*(VELFP - v) (drag_gs_des1.f:126-137), VELFP from the
// current cell-centre gas velocity with the particle's filter weights (or
// its cell's value), then DRAG_FC += P_FORCE*PVOL with P_FORCE interpolated
// the same way (calc_pg_grad.f:74-105); DRAG_FC = F_GP = 0 outside a fluid
// cell.  Results in global-id order.
__global__ __launch_bounds__(kBS) void k_drag_fc(DevParams P, StepPtrs a) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= P.n || a.state[i] != 1) return;  // IS_NORMAL
  const int pc = a.fcell[i];
  if (pc < 0) return;
  int ci, cj, ck;
  unpack_ijk(pc, ci, cj, ck);
  const int c0 = lin_ijk(ci, cj, ck, P.ni, P.nj);
  const int g = a.gid[i];
  double* dfc = a.drag_fc + 3 * (size_t)g;
  if (!a.fluid[c0]) {
    for (int d = 0; d < 3; d++) dfc[d] = 0.0;
    a.f_gp[g] = 0.0;
    return;
  }
  const int S = P.stride;
  double pos[3], vel[3];
  load3(a.pos, S, i, pos);
  load3(a.vel, S, i, vel);
  double velfp[3] = {0.0, 0.0, 0.0}, lpf[3] = {0.0, 0.0, 0.0};
  if (P.interp_on && filter_scheme(P.interp)) {
    for_filter(P, a.xe, a.yn, a.zt, a.csz, a.fluid, ci, cj, ck, pos, [&](int c, double w) {
      for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + a.ugc[3 * c + d] * w;
    });
    for_filter(P, a.xe, a.yn, a.zt, a.csz, a.fluid, ci, cj, ck, pos, [&](int c, double w) {
      for (int d = 0; d < 3; d++) lpf[d] = lpf[d] + a.p_force[3 * c + d] * w;
    });
  } else {
    for (int d = 0; d < 3; d++) velfp[d] = a.ugc[3 * c0 + d];
    for (int d = 0; d < 3; d++) lpf[d] = a.p_force[3 * c0 + d];
  }
  const double f = a.f_gp[g], pv = a.pvol[i];
  double out[3];
  for (int d = 0; d < 3; d++) out[d] = f * (velfp[d] - vel[d]);
  for (int d = 0; d < 3; d++) out[d] = out[d] + lpf[d] * pv;
  for (int d = 0; d < 3; d++) dfc[d] = out[d];
}

// k_gas_fgds: F_GDS of a fluid cell, 1/8 of the DRAG_AM of its 8 corner
// nodes in the reference's sum order (drag_gs_des0.f:455-494), else 0.
__global__ __launch_bounds__(kBS) void k_gas_fgds(int ni, int nj, int nk, const uint8_t* fluid,
                                                  const double* am, double* f_gds) {
  const int c = blockIdx.x * kBS + threadIdx.x;
  if (c >= ni * nj * nk) return;
  if (!fluid[c]) {
    f_gds[c] = 0.0;
    return;
  }
  // a fluid cell is interior, so its west, south and bottom nodes exist
  const int di = nj, dk = nj * ni;
  double f = 0.125 * (am[c] + am[c - 1] + am[c - di - 1] + am[c - di]);
  f = f + 0.125 * (am[c - dk] + am[c - dk - 1] + am[c - dk - di - 1] + am[c - dk - di]);
  f_gds[c] = f;
}

// DES_VOL_NODE of the uniform grid (compute_volume_of_nodes,
// cfassign.f:100-230): 1/8 of the fluid volume of the 8 cells sharing node
// (I, J, K) for I <= imax+1, J <= jmax+1, K <= kmax+1, summed in the
// reference's order; 0 at every other index.
__global__ __launch_bounds__(kBS) void k_vol_node(int imax, int jmax, int kmax, int ni, int nj,
                                                  int nk, const uint8_t* fluid, const double* csz,
                                                  double* vol_node) {
  const int node = blockIdx.x * kBS + threadIdx.x;
  if (node >= ni * nj * nk) return;
  const int J = node % nj + 1;
  const int I = (node / nj) % ni + 1;
  const int K = node / (nj * ni) + 1;
  double a = 0.0;
  if (I <= imax + 1 && J <= jmax + 1 && K <= kmax + 1) {
    // dx(I)*dy(J)*dz(K) of each cell (cfassign.f:154-157)
    auto volf = [&](int i, int j, int k) {
      return fluid[(j - 1) + (i - 1) * nj + (k - 1) * nj * ni]
                 ? csz[i - 1] * csz[ni + j - 1] * csz[ni + nj + k - 1]
                 : 0.0;
    };
    const double af = 0.125;
    a = af * (volf(I, J, K) + volf(I + 1, J, K) + volf(I, J + 1, K) + volf(I + 1, J + 1, K));
    a = a + af * (volf(I, J, K + 1) + volf(I + 1, J + 1, K + 1) + volf(I, J + 1, K + 1) +
                  volf(I + 1, J, K + 1));
  }
  vol_node[node] = a;
}

// node -> cell ROP_S (comp_mean_fields0.f:298-345), ROP_S/VOL (:347-370) and
// CALC_EPG_DES (calc_epg_des.f:95-150).  Nodes visited K, J, I ascending.
// FIXED: node sums come from the fixed-point accumulators (converted exactly
// by the inverse power-of-two scale after one rounding to FP64), and the
// other accumulator buffer is cleared for the next step.
struct FieldPtrs {
  const double* rops_node;               // FP64 node sums (reference-order path)
  const unsigned long long* fix;         // fixed-point node sums
  unsigned long long* fix_clear;         // buffer to zero for the next step
  const double *vol, *vol_surr, *ro_s;
  const uint8_t* fluid;
  double *rop_s, *ep_g;
  int* flags;
  // removals queued by the fused step (null: none): count, slot ids, and
  // the other step's count, cleared for the next step
  const int* rm_cnt;
  int* rm_cnt_clear;
  const int* rm_list;
  uint8_t* state;
  double* radius;
};

// Completes the removals the fused step queued (particles that left the
// fluid grid): NONEXISTENT state and zero radius.  Grid-stride over the
// queue; any thread may run it.
__device__ __forceinline__ void apply_removals(const int* cnt, const int* list, uint8_t* state,
                                               double* radius, int tid, int nthreads) {
  const int m = *cnt;
  for (int q = tid; q < m; q += nthreads) {
    const int i = list[q];
    state[i] = 0;
    radius[i] = 0.0;
  }
}
// The same as a kernel of its own, ahead of a neighbour rebuild (which
// reorders the slots the queue names); one workgroup, which then empties
// the queue.
__global__ __launch_bounds__(kBS) void k_apply_removals(int* cnt, const int* list,
                                                        uint8_t* state, double* radius) {
  apply_removals(cnt, list, state, radius, threadIdx.x, kBS);
  __syncthreads();
  if (threadIdx.x == 0) *cnt = 0;
}
// Node sums -> ROP_S, EP_G per fluid cell (comp_mean_fields0.f:339-432,
// calc_epg_des.f:133-145).  A workgroup takes kFieldCells consecutive
// cells (J fastest), kFieldCPT per thread (two per thread measured the same
// as one at 64x128x64 cells).  The corner nodes of those cells form four runs of
// kFieldCells + 1 consecutive nodes, one per (I, K) corner pair, so the
// workgroup reads each node once, coalesced, and computes its density (node
// sum / VOL_SURR) once into LDS instead of once per adjacent cell.  Each
// cell then sums its corners in the reference's order (I fastest, then J,
// then K), skipping corners outside the node range and nodes with VOL_SURR
// = 0 (their density is stored as 0, and adding 0 leaves the non-negative
// sum as is).
constexpr int kFieldCPT = 1, kFieldCells = kBS * kFieldCPT;
template <bool FIXED>
__global__ __launch_bounds__(kBS) void k_cell_fields(int imax, int jmax, int kmax, int ni, int nj,
                                                     int nk, int nph, double packed_eps,
                                                     double inv_scale0, double inv_scale1,
                                                     double inv_scale2, double inv_scale3,
                                                     FieldPtrs f) {
  constexpr int kRun = kFieldCells + 1, kNodes = 4 * kRun, kPer = (kNodes + kBS - 1) / kBS;
  __shared__ double s_dens[4][kRun];
  const int ncell = ni * nj * nk, plane = nj * ni;
  // XCD-aware order: the workgroups of one XCD take a contiguous range of
  // cells, so the node rows a K-neighbour reads again are in that XCD's L2
  const int c0 = xcd_tile(blockIdx.x, gridDim.x) * kFieldCells;
  if (FIXED && f.rm_cnt) {
    apply_removals(f.rm_cnt, f.rm_list, f.state, f.radius, blockIdx.x * kBS + threadIdx.x,
                   gridDim.x * kBS);
    if (blockIdx.x == 0 && threadIdx.x == 0) *f.rm_cnt_clear = 0;
  }
  const double inv[kMaxPh] = {inv_scale0, inv_scale1, inv_scale2, inv_scale3};
  // this thread's cells c0 + tid + v kBS: corner mask (bit q: the corner is
  // inside the node range), volume, fluid flag
  unsigned okm[kFieldCPT];
  double vc[kFieldCPT];
  uint8_t fl[kFieldCPT];
#pragma unroll
  for (int v = 0; v < kFieldCPT; v++) {
    const int c = c0 + threadIdx.x + v * kBS;
    const bool live = c < ncell;
    const int cj = c % nj + 1;
    const int ci = (c / nj) % ni + 1;
    const int ck = c / plane + 1;
    okm[v] = 0u;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int I = ci - 1 + (q & 1), J = cj - 1 + ((q >> 1) & 1), K = ck - 1 + (q >> 2);
      if (live && I >= 1 && I <= imax + 1 && J >= 1 && J <= jmax + 1 && K >= 1 &&
          K <= kmax + 1)
        okm[v] |= 1u << q;
    }
    vc[v] = live ? f.vol[c] : 1.0;
    fl[v] = live ? f.fluid[c] : 0;
  }
  // loader nodes: entry e of run r = qi + 2 qk is node
  // c0 - 1 - (1 - qi) nj - (1 - qk) ni nj + e, the (qi, qj, qk) corner of
  // cell c0 + e - qj
  int nidx[kPer];
  double vsl[kPer];
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int e = threadIdx.x + u * kBS;
    nidx[u] = -1;
    if (e < kNodes) {
      const int r = e / kRun, qi = r & 1, qk = r >> 1;
      const int nd = c0 - 1 - (1 - qi) * nj - (1 - qk) * plane + (e - r * kRun);
      if (nd >= 0 && nd < ncell) nidx[u] = nd;
    }
    vsl[u] = nidx[u] >= 0 ? f.vol_surr[nidx[u]] : 0.0;
  }
  double acc[kFieldCPT][kMaxPh] = {};
  for (int m = 0; m < nph; m++) {
    if (m > 0) __syncthreads();  // the previous phase's densities are read
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = threadIdx.x + u * kBS;
      if (e >= kNodes) continue;
      double d = 0.0;
      if (nidx[u] >= 0) {
        double nodev;
        if (FIXED) {
          const unsigned long long x = f.fix[(size_t)m * ncell + nidx[u]];
          if (x >> 62) atomicOr(&f.flags[kFlagEpg], 1 << 30);  // accumulator range exceeded
          nodev = (double)x * inv[m];
        } else {
          nodev = f.rops_node[(size_t)m * ncell + nidx[u]];
        }
        if (vsl[u] != 0.0) d = nodev / vsl[u];
      }
      s_dens[e / kRun][e % kRun] = d;
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < kFieldCPT; v++) {
      const int o = threadIdx.x + v * kBS;
      double a = 0.0;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if (!((okm[v] >> q) & 1u)) continue;
        const double dens = s_dens[(q & 1) + 2 * (q >> 2)][o + ((q >> 1) & 1)];
        a = a + dens * vc[v];
      }
      acc[v][m] = a;
    }
  }
#pragma unroll
  for (int v = 0; v < kFieldCPT; v++) {
    const int c = c0 + threadIdx.x + v * kBS;
    if (c >= ncell) continue;
    double* am = acc[v];
    if (FIXED)
      for (int m = 0; m < nph; m++) f.fix_clear[(size_t)m * ncell + c] = 0ull;
    if (fl[v])
      for (int m = 0; m < nph; m++)
        if (am[m] > 0.0) am[m] = am[m] / vc[v];
    for (int m = 0; m < nph; m++) f.rop_s[(size_t)m * ncell + c] = am[m];
    if (fl[v]) {
      double sum_eps = 0.0;
      for (int m = 0; m < nph; m++) sum_eps = sum_eps + am[m] / f.ro_s[m];
      const double e = 1.0 - (packed_eps < sum_eps ? packed_eps : sum_eps);
      f.ep_g[c] = e;
      if (e <= 0.0 || e > 1.0) atomicAdd(&f.flags[kFlagEpg], 1);
    }
  }
}

// SOLVOLINC -> ROP_S = RO_S*SOLVOLINC/VOL in fluid cells, 0 elsewhere
// (comp_mean_fields1.f:66-80), then EP_G (calc_epg_des.f:133-145); clears
// the other accumulator buffer for the next step.
__global__ __launch_bounds__(kBS) void k_cell_fields_des1(int ncell, int nph, double packed_eps,
                                                          double inv_scale0, double inv_scale1,
                                                          double inv_scale2, double inv_scale3,
                                                          FieldPtrs f) {
  const int c = blockIdx.x * kBS + threadIdx.x;
  if (c >= ncell) return;
  const double inv[kMaxPh] = {inv_scale0, inv_scale1, inv_scale2, inv_scale3};
  const uint8_t fl = f.fluid[c];
  const double vc = f.vol[c];
  double acc[kMaxPh] = {0.0, 0.0, 0.0, 0.0};
  for (int m = 0; m < nph; m++) {
    const unsigned long long u = f.fix[(size_t)m * ncell + c];
    f.fix_clear[(size_t)m * ncell + c] = 0ull;
    if (u >> 62) atomicOr(&f.flags[kFlagEpg], 1 << 30);  // accumulator range exceeded
    if (fl) acc[m] = f.ro_s[m] * ((double)u * inv[m]) / vc;
  }
  for (int m = 0; m < nph; m++) f.rop_s[(size_t)m * ncell + c] = acc[m];
  if (fl) {
    double sum_eps = 0.0;
    for (int m = 0; m < nph; m++) sum_eps = sum_eps + acc[m] / f.ro_s[m];
    const double e = 1.0 - (packed_eps < sum_eps ? packed_eps : sum_eps);
    f.ep_g[c] = e;
    if (e <= 0.0 || e > 1.0) atomicAdd(&f.flags[kFlagEpg], 1);
  }
}

// Gas velocities the drag reads, from the staggered U_G/V_G/W_G (one thread
// per cell, MFiX serial layout):
//  * node velocities of DRAG_GS_DES0 (drag_gs_des0.f:148-179) at nodes
//    (I,J,K) = (1..IMAX+1, 1..JMAX+1, 1..KMAX+1), averages of four face
//    values with AVG_FACTOR = 1/4, in the reference's sum order; 0 at the
//    other indices;
//  * cell-centre velocities of CALC_CELL_CENTER_GAS_VEL
//    (drag_gs_des1.f:313-370) in fluid cells, AVG_X_E with FX = 1/2 and
//    HALF*(m + p) for AVG_Y_N / AVG_Z_T (fun_avg.inc, set_geometry.f:344);
//    0 outside fluid cells.
__global__ __launch_bounds__(kBS) void k_gas_velocities(int ni, int nj, int nk, int imax,
                                                        int jmax, int kmax,
                                                        const uint8_t* fluid, const double* ug,
                                                        const double* vg, const double* wg,
                                                        double* vnode, double* ugc) {
  const int c = blockIdx.x * kBS + threadIdx.x;
  if (c >= ni * nj * nk) return;
  const int J = c % nj + 1, I = (c / nj) % ni + 1, K = c / (nj * ni) + 1;
  const int si = nj, sk = nj * ni;  // strides of I and K (J is 1)
  double vx = 0.0, vy = 0.0, vz = 0.0;
  if (I <= imax + 1 && J <= jmax + 1 && K <= kmax + 1) {
    const double af = 0.25;
    const int ipjk = c + si, ijpk = c + 1, ipjpk = c + si + 1;
    const int ijkp = c + sk, ijpkp = c + 1 + sk, ipjkp = c + si + sk;
    vx = af * (ug[c] + ug[ijpk]);
    vy = af * (vg[c] + vg[ipjk]);
    vx = vx + af * (ug[ijkp] + ug[ijpkp]);
    vy = vy + af * (vg[ijkp] + vg[ipjkp]);
    vz = af * (wg[c] + wg[ijpk] + wg[ipjk] + wg[ipjpk]);
  }
  vnode[3 * c] = vx;
  vnode[3 * c + 1] = vy;
  vnode[3 * c + 2] = vz;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (fluid[c]) {  // fluid cells are interior: the west/south/bottom faces exist
    gx = 0.5 * ug[c - si] + (1.0 - 0.5) * ug[c];
    gy = 0.5 * (vg[c - 1] + vg[c]);
    gz = 0.5 * (wg[c - sk] + wg[c]);
  }
  ugc[3 * c] = gx;
  ugc[3 * c + 1] = gy;
  ugc[3 * c + 2] = gz;
}

// P_FORCE(3, IJK) from the caller's array, or 0 when none is given.
__global__ void k_copy_or_zero_f64(double* dst, const double* src, int n) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) dst[i] = src ? src[i] : 0.0;
}

// calc_del_phi (calc_grad_des.f:104-133): 3-point derivative, a neighbour
// that is not a valid pressure cell replaced by extrapolation
__device__ __forceinline__ double calc_del_phi(bool valid_lo, bool valid_hi, double phi,
                                               double dx, double phi_lo, double dx_lo,
                                               double phi_hi, double dx_hi) {
  if (!valid_lo && !valid_hi) return 0.0;
  if (!valid_lo) phi_lo = phi_hi + (phi - phi_hi) * (dx_hi + 2.0 * dx + dx_lo) / (dx + dx_hi);
  if (!valid_hi) phi_hi = phi_lo + (phi - phi_lo) * (dx_hi + 2.0 * dx + dx_lo) / (dx + dx_lo);
  const double x = 0.5 * (dx_lo + dx);
  const double x_hi = 0.5 * (dx_hi + dx) + x;
  return phi_lo * (x - x_hi) / (x_hi * x) + phi * (2.0 * x - x_hi) / (x * (x - x_hi)) +
         phi_hi * (x) / (x_hi * (x_hi - x));
}

// CALC_PG_GRAD (calc_pg_grad.f:20-110) / CALC_GRAD_DES (calc_grad_des.f:21-102):
// P_FORCE = -grad P_G in fluid cells, 0 elsewhere (box walls: the valid
// pressure cells are the fluid cells; no cyclic pressure drop).
__global__ __launch_bounds__(kBS) void k_pg_grad(int ni, int nj, int nk, const double* csz,
                                                 const uint8_t* fluid, const double* p_g,
                                                 double* p_force) {
  const int c = blockIdx.x * kBS + threadIdx.x;
  if (c >= ni * nj * nk) return;
  double* f = p_force + 3 * c;
  if (!fluid[c]) {
    f[0] = f[1] = f[2] = 0.0;
    return;
  }
  // fluid cells are interior, so every neighbour index is on the grid;
  // DX(I), DX of the west and east neighbours (calc_grad_des.f:85-87)
  const int sj = 1, si = nj, sk = nj * ni;
  const int J = c % nj, I = (c / nj) % ni, K = c / sk;  // 0-based
  const double *dx = csz + I, *dy = csz + ni + J, *dz = csz + ni + nj + K;
  const double g0 = calc_del_phi(fluid[c - si], fluid[c + si], p_g[c], dx[0], p_g[c - si], dx[-1],
                                 p_g[c + si], dx[1]);
  const double g1 = calc_del_phi(fluid[c - sj], fluid[c + sj], p_g[c], dy[0], p_g[c - sj], dy[-1],
                                 p_g[c + sj], dy[1]);
  const double g2 = calc_del_phi(fluid[c - sk], fluid[c + sk], p_g[c], dz[0], p_g[c - sk], dz[-1],
                                 p_g[c + sk], dz[1]);
  f[0] = -g0;
  f[1] = -g1;
  f[2] = -g2;
}

// PARTICLES_IN_CELL deletion of particles outside the fluid grid
// (particles_in_cell.f:175-184), for particles the fused step flagged
__global__ __launch_bounds__(kBS) void k_delete_outside(int n, int S, const int* fcell,
                                                        uint8_t* state, uint8_t* wmask,
                                                        double* pos, double* vel, double* omega,
                                                        double* radius, int* flags) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n || state[i] == 0 || fcell[i] >= 0) return;
  state[i] = 0;
  wmask[i] = 0;
  for (int d = 0; d < 3; d++) {
    pos[d * S + i] = 0.0;
    vel[d * S + i] = 0.0;
    omega[d * S + i] = 0.0;
  }
  radius[i] = 0.0;
  atomicAdd(&flags[kFlagOutOfDomain], 1);
}

// ------------------------------------------------------------- exports
__global__ void k_scatter_gid_d3(int n, int S, int ld, const int* gid, const uint8_t* state,
                                 const double* src, double* dst) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const int g = gid[i];
  for (int c = 0; c < 3; c++) dst[(size_t)c * ld + g] = src[c * S + i];
}
__global__ void k_scatter_gid_i(int n, const int* gid, const int* src, int add, int* dst) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  dst[gid[i]] = src[i] + add;
}
// Holes at set-up: slots q < n whose PARTICLE_STATE st[q] is 0 become
// nonexistent with zero radius and state (slot order = global-id order
// before the first sort).
__global__ __launch_bounds__(kBS) void k_set_holes(int n, const int8_t* st, uint8_t* state,
                                                   double* radius, int S, double* pos,
                                                   double* vel, double* omega) {
  const int q = blockIdx.x * kBS + threadIdx.x;
  if (q >= n || st[q] != 0) return;
  state[q] = 0;
  radius[q] = 0.0;
  for (int d = 0; d < 3; d++) {
    pos[d * S + q] = 0.0;
    vel[d * S + q] = 0.0;
    omega[d * S + q] = 0.0;
  }
}

// PARTICLE_STATE in global-id order (a seeded particle is ENTERING)
__global__ void k_scatter_gid_u8(int n, const int* gid, const uint8_t* src, int8_t* dst) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const uint8_t v = src[i];
  dst[gid[i]] = (int8_t)(v == kSeeded ? kEntering : v);
}

}  // namespace mfixdem

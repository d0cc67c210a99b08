/*
 * dem_oracle.c -- CPU restatement of MFiX 25.2.1's DEM particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product library
 * links or calls this file.
 *
 * Each function follows the loop order and FP64 expression order of the
 * reference routine it cites (paths relative to the reference's model/),
 * so that a serial build is bit-reproducible and can serve as the parity
 * oracle for the HIP path.  Fortran `a*b*c` is evaluated left to right,
 * `x**2` / `x**3` as repeated products, DOT_PRODUCT as ((a1*b1+a2*b2)+a3*b3);
 * the file is compiled with -ffp-contract=off so no FMA is formed.
 *
 * Parity pins (see tests/): DEM01 golden trajectories (Euler and
 * Adams-Bashforth, LSD wall) from the reference's own AUTOTEST files;
 * brute-force NSQUARE pair sets; closed-form and invariant checks for the
 * contact laws, drag correlations and deposition (see tests/test_oracle*.py).
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.141592653589793 /* 4.D0*ATAN(1.D0), constant_mod.f:168 */
#define SMALL_NUMBER 1.0e-15   /* param1_mod.f:26 */
#define LARGE_NUMBER 1.0e32    /* param1_mod.f:25 */
#define UNDEF MFIXDEM_UNDEFINED
#define NORMAL_PARTICLE 1
#define NONEXISTENT 0
#define ENTERING_PARTICLE 2
#define EXITING_PARTICLE 3
#define NUMFRAC_LIMIT 10000
#define COLLISION_ARRAY_MAX 8  /* discretelement_mod.f:637 */
#define MP MFIXDEM_MAX_PHASES
#define FILTER_N 27            /* FILTER_SIZE = merge(27, 9, DO_K), 3-D */

struct orc_sys {
  mfixdem_params p;
  mfixdem_derived d;
  double kt, kt_w;
  /* fluid grid (MFiX serial: cells 1..IMAX+2, faces XE(0:IMAX+2)) */
  int ni, nj, nk, ncell;
  /* DX(I) of cell I = 1..ni at cdx[I-1] (ghosts included); smallest
     interior sizes */
  double *cdx, *cdy, *cdz;
  double dx_min, dy_min, dz_min;
  double *xe, *yn, *zt;
  double *vol, *vol_surr;
  unsigned char* fluid;
  double x_max, y_max, z_max;
  /* DES grid */
  int dg_isize, dg_jsize, dg_ksize;
  int dg_ni, dg_nj, dg_nk, dg_ncell;
  int dg_istart1, dg_iend1, dg_jstart1, dg_jend1, dg_kstart1, dg_kend1;
  int dg_istart2, dg_iend2, dg_jstart2, dg_jend2, dg_kstart2, dg_kend2;
  double dg_xstart, dg_ystart, dg_zstart, dg_dxinv, dg_dyinv, dg_dzinv;
  int* dg_cnt;   /* per DES cell */
  int* dg_off;   /* CSR offsets, dg_ncell+1 */
  int* dg_list;  /* particle indices (0-based), ascending within a cell */
  /* facets */
  int nf;
  double (*vert)[3][3];    /* VERTEX(v, c, nf) -> vert[nf][v][c] */
  double (*norm)[3];
  unsigned char (*ignore_edge)[4];
  double (*wall_vt)[3];    /* STL_TANGENTIAL_VEL(:, nf) */
  double (*wall_vn)[3];    /* STL_NORMAL_VEL(:, nf) */
  int* fc_count;           /* facets_at_dg count */
  int* fc_off;             /* list of DES cell c: [fc_off[c], fc_off[c] + fc_count[c]) */
  int* fc_id;              /* facet ids, ascending within a cell */
  int* fc_dir;
  double* fc_min;
  double* fc_max;
  int wall_slot_overflow;  /* GET_COLLISION found no slot (the reference grows the table) */
  /* particles */
  int64_t max_pip, ld;
  int64_t pip;
  double *pos, *vel, *omega;          /* [c*ld + p] */
  double *radius, *ro_sol, *pvol, *pmass, *omoi;
  int* phase;                         /* 0-based */
  signed char* state;
  double *fc, *tow, *fc_last, *tow_last;
  double *pos_old, *vel_old, *omega_old, *acc_old, *rot_acc_old;
  int first_pass;
  int* dg_pijk;                       /* 1-based DES cell */
  int* pijk;                          /* [c*ld + p], c = 0..3: I,J,K,IJK (1-based) */
  /* neighbour CSR (1-based positions, reference layout) */
  int64_t cap;
  int *neighbor_index, *neighbor_index_old;
  int *neighbors, *neighbors_old;
  double *pft, *pft_old;              /* [3*c + d] */
  /* wall collision slots (calc_collision_wall_mod.f:651-792) */
  int* wall_fid;                      /* [slot*ld + p] */
  double* wall_pft;                   /* [(slot*3 + d)*ld + p] */
  /* gas */
  int gas_set;
  double *u_g, *v_g, *w_g, *p_force, *ep_g, *rop_s, *rops_node;
  /* filter schemes (DES1 path) */
  double fw, fw3, oofv;               /* FILTER_WIDTH_INTERP, x3, OoFILTER_VOL */
  double* ugc;                        /* cell-centre gas velocity [3*ijk + d] */
  int* filter_cell;                   /* FILTER_CELL(27, MAX_PIP) -> [lc*ld + p], 0-based */
  double* filter_weight;              /* FILTER_WEIGHT(27, MAX_PIP) */
  double* solvolinc;                  /* SOLVOLINC(ijk, m) -> [m*ncell + ijk] */
  double* f_gp;                       /* F_GP(MAX_PIP): the gas side's drag coefficient */
  double* drag_fc;                    /* DRAG_FC(MAX_PIP,3) -> [d*ld + p] */
  int* pinc;
  int* pic_off;
  int* pic_list;
  mfixdem_status st;
  int inited;
  /* DEM mass inlets and outlets (des_bc_mod.f) */
  int nbc;
  struct orc_bc* bcs;
  uint64_t rng;                       /* stands in for RANDOM_NUMBER */
  double s_time;                      /* S_TIME of the step being taken */
};

/* One DEM boundary: the input, its axes, the DES cells whose particles it
   looks at (DEM_BCMI_IJK / DEM_BCMO_IJK as an index box), and for an inlet
   the layout of LAYOUT_MI_DEM and the seeding schedule of SET_BC_DEM_MI. */
typedef struct orc_bc {
  mfixdem_dem_bc bc;
  int ax, pa, qa;                     /* normal axis; in-plane axes P, Q */
  int low;                            /* boundary at the low end of ax */
  double face;                        /* XE/YN/ZT of the boundary plane */
  int clo[3], chi[3];
  int occupants, vacancy, pi_count, polydisperse, L;
  double pi_factor, mi_time, window, offset;
  double *P, *Q;
  int *W, *H;
  int* poly_layout;                   /* DEM_BC_POLY_LAYOUT, NUMFRAC_LIMIT entries, 0-based phase */
  double vel[MP][3];                  /* BC_U_s, BC_V_s, BC_W_s after SET_BC_DEM_MI */
} orc_bc;

static void free_bcs(orc_sys* s);

/* ---------------------------------------------------------------- helpers */
static inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* CROSS, discretelement_mod.f:906-922 */
static inline void cross3(const double x[3], const double y[3], double c[3]) {
  c[0] = x[1] * y[2] - x[2] * y[1];
  c[1] = x[2] * y[0] - x[0] * y[2];
  c[2] = x[0] * y[1] - x[1] * y[0];
}
static inline int fluid_ijk(const orc_sys* s, int i, int j, int k) {
  /* FUNIJK (functions.inc:37-42), serial: J fastest; 1-based result */
  return j + (i - 1) * s->nj + (k - 1) * s->nj * s->ni;
}
static inline int dg_funijk(const orc_sys* s, int i, int j, int k) {
  /* dg_funijk (desgrid_mod.f:184-188) with istart2=jstart2=kstart2=1 */
  return j + (i - 1) * s->dg_nj + (k - 1) * s->dg_nj * s->dg_ni;
}
static inline void dg_ijk_of(const orc_sys* s, int ijk, int* i, int* j, int* k) {
  int z = ijk - 1;
  *j = z % s->dg_nj + 1;
  z /= s->dg_nj;
  *i = z % s->dg_ni + 1;
  *k = z / s->dg_ni + 1;
}
/* floor of a double, saturated to the int range (NaN -> very negative) */
static inline int sat_floor(double v) {
  if (!(v > -1.0e9)) return -1000000000;
  if (v > 1.0e9) return 1000000000;
  return (int)floor(v);
}
static inline int iofpos(const orc_sys* s, double x) {
  return sat_floor((x - s->dg_xstart) * s->dg_dxinv) + s->dg_istart1;
}
static inline int jofpos(const orc_sys* s, double y) {
  return sat_floor((y - s->dg_ystart) * s->dg_dyinv) + s->dg_jstart1;
}
static inline int kofpos(const orc_sys* s, double z) {
  return sat_floor((z - s->dg_zstart) * s->dg_dzinv) + s->dg_kstart1;
}
static inline int imin_i(int a, int b) { return a < b ? a : b; }
static inline int imax_i(int a, int b) { return a > b ? a : b; }

int orc_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Thread count of the OpenMP build's later parallel regions (the CPU
   baseline times 1 thread and all threads); no effect on the serial build. */
void orc_set_omp_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------ parameter derivation */
/* CHECK_SOLIDS_DEM_COLL_LSD, check_data/check_solids_dem.f:519-782 and
   REF_MASS / REF_RADIUS at :1162-1163. */
static void derive_lsd(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  int n = p->nphases, lc = 0;
  double tcoll = UNDEF;
  double kt_fac = (p->kt_fac == UNDEF) ? 2.0 / 7.0 : p->kt_fac;
  double kt_w_fac = (p->kt_w_fac == UNDEF) ? 2.0 / 7.0 : p->kt_w_fac;
  double etat_fac = (p->des_etat_fac == UNDEF) ? 0.5 : p->des_etat_fac;
  double etat_w_fac = (p->des_etat_w_fac == UNDEF) ? 0.5 : p->des_etat_w_fac;
  s->kt = kt_fac * p->kn;
  s->kt_w = kt_w_fac * p->kn_w;
  for (int m = 0; m < n; m++) {
    double dm = p->d_p0[m];
    double mass_m = (PI_D / 6.0) * (dm * dm * dm) * p->ro_s0[m];
    for (int l = m; l < n; l++) {
      double en = p->des_en_input[lc++];
      double dl = p->d_p0[l];
      double mass_l = (PI_D / 6.0) * (dl * dl * dl) * p->ro_s0[l];
      double mass_eff = mass_m * mass_l / (mass_m + mass_l);
      double etan, tc;
      if (en != 0.0) {
        double lg = log(en);
        etan = 2.0 * sqrt(p->kn * mass_eff) * fabs(lg);
        etan = etan / sqrt(PI_D * PI_D + (lg * lg));
        double q = etan / mass_eff;
        tc = PI_D / sqrt(p->kn / mass_eff - (q * q) / 4.0);
      } else {
        etan = 2.0 * sqrt(p->kn * mass_eff);
        tc = PI_D / sqrt(p->kn / mass_eff);
      }
      if (tc < tcoll) tcoll = tc;
      etan = etan / sqrt(mass_eff);
      s->d.des_etan[m][l] = etan;
      s->d.des_etat[m][l] = etat_fac * etan;
      s->d.des_etan[l][m] = s->d.des_etan[m][l];
      s->d.des_etat[l][m] = s->d.des_etat[m][l];
    }
    double en = p->des_en_wall_input[m];
    double mass_eff = mass_m, etan;
    if (en != 0.0) {
      double lg = log(en);
      etan = 2.0 * sqrt(p->kn_w * mass_eff) * fabs(lg);
      etan = etan / sqrt(PI_D * PI_D + (lg * lg));
    } else {
      etan = 2.0 * sqrt(p->kn_w * mass_eff);
    }
    etan = etan / sqrt(mass_eff);
    s->d.des_etan_wall[m] = etan;
    s->d.des_etat_wall[m] = etat_w_fac * etan;
  }
  s->d.tcoll = tcoll;
}

/* CHECK_SOLIDS_DEM_COLL_HERTZ, check_data/check_solids_dem.f:798-1122 */
static void derive_hertz(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  int n = p->nphases, lc = 0;
  double tcoll = UNDEF;
  double g_mod[MP];
  double g_mod_wall = 0.5 * p->ew_young / (1.0 + p->vw_poisson);
  for (int m = 0; m < n; m++) g_mod[m] = 0.5 * p->e_young[m] / (1.0 + p->v_poisson[m]);
  for (int m = 0; m < n; m++) {
    double dm = p->d_p0[m];
    double mass_m = (PI_D / 6.0) * (dm * dm * dm) * p->ro_s0[m];
    double rm = 0.5 * dm;
    for (int l = m; l < n; l++) {
      double en = p->des_en_input[lc];
      double et = p->des_et_input[lc];
      lc++;
      double dl = p->d_p0[l];
      double mass_l = (PI_D / 6.0) * (dl * dl * dl) * p->ro_s0[l];
      double rl = 0.5 * dl;
      double mass_eff = (mass_m * mass_l) / (mass_m + mass_l);
      double red_mass_eff = (2.0 / 7.0) * mass_eff;
      double r_eff = rm * rl / (rm + rl);
      double vm = p->v_poisson[m], vl = p->v_poisson[l];
      double e_eff = p->e_young[m] * p->e_young[l] /
                     (p->e_young[m] * (1.0 - vl * vl) + p->e_young[l] * (1.0 - vm * vm));
      double g_eff = g_mod[m] * g_mod[l] /
                     (g_mod[m] * (2.0 - vl) + g_mod[l] * (2.0 - vm));
      double hkn = (4.0 / 3.0) * sqrt(r_eff) * e_eff;
      double hkt = 8.0 * sqrt(r_eff) * g_eff;
      double etan, etat, tc;
      if (en != 0.0) {
        double lg = log(en);
        etan = sqrt(5.0) * sqrt(hkn * mass_eff) * fabs(lg);
        etan = etan / sqrt(PI_D * PI_D + (lg * lg));
      } else {
        etan = sqrt(5.0) * sqrt(hkn * mass_eff);
      }
      if (et != 0.0) {
        double lg = log(et);
        etat = sqrt(5.0) * sqrt(hkt * red_mass_eff) * fabs(lg);
        etat = etat / sqrt(PI_D * PI_D + (lg * lg));
      } else {
        etat = sqrt(5.0) * sqrt(hkt * red_mass_eff);
      }
      if (en != 0.0) {
        double q = etan / mass_eff;
        tc = PI_D / sqrt(hkn / mass_eff - (q * q) / 5.0);
      } else {
        tc = PI_D / sqrt(hkn / mass_eff);
      }
      if (tc < tcoll) tcoll = tc;
      etan = etan / sqrt(hkn * mass_eff);
      etat = etat / sqrt(hkt * mass_eff);
      hkn = hkn / sqrt(r_eff);
      hkt = hkt / sqrt(r_eff);
      s->d.hert_kn[m][l] = s->d.hert_kn[l][m] = hkn;
      s->d.hert_kt[m][l] = s->d.hert_kt[l][m] = hkt;
      s->d.des_etan[m][l] = s->d.des_etan[l][m] = etan;
      s->d.des_etat[m][l] = s->d.des_etat[l][m] = etat;
    }
    /* wall */
    double en = p->des_en_wall_input[m], et = p->des_et_wall_input[m];
    double mass_eff = mass_m;
    double red_mass_eff = (2.0 / 7.0) * mass_eff;
    double r_eff = rm;
    double vm = p->v_poisson[m], vw = p->vw_poisson;
    double e_eff = p->e_young[m] * p->ew_young /
                   (p->e_young[m] * (1.0 - vw * vw) + p->ew_young * (1.0 - vm * vm));
    double g_eff = g_mod[m] * g_mod_wall /
                   (g_mod[m] * (2.0 - vw) + g_mod_wall * (2.0 - vm));
    double hkwn = (4.0 / 3.0) * sqrt(r_eff) * e_eff;
    double hkwt = 8.0 * sqrt(r_eff) * g_eff;
    double etan, etat;
    if (en != 0.0) {
      double lg = log(en);
      etan = sqrt(5.0) * sqrt(hkwn * mass_eff) * fabs(lg);
      etan = etan / sqrt(PI_D * PI_D + (lg * lg));
    } else {
      etan = sqrt(5.0) * sqrt(hkwn * mass_eff);
    }
    if (et != 0.0) {
      double lg = log(et);
      etat = sqrt(5.0) * sqrt(hkwt * red_mass_eff) * fabs(lg);
      etat = etat / sqrt(PI_D * PI_D + (lg * lg));
    } else {
      etat = sqrt(5.0) * sqrt(hkwt * red_mass_eff);
    }
    etan = etan / sqrt(hkwn * mass_eff);
    etat = etat / sqrt(hkwt * mass_eff);
    hkwn = hkwn / sqrt(r_eff);
    hkwt = hkwt / sqrt(r_eff);
    s->d.hert_kwn[m] = hkwn;
    s->d.hert_kwt[m] = hkwt;
    s->d.des_etan_wall[m] = etan;
    s->d.des_etat_wall[m] = etat;
  }
  s->d.tcoll = tcoll;
}

/* -------------------------------------------------------- geometry set-up */
/* Cell sizes of one axis: the interior sizes given (DX(0:IMAX-1) of the
   namelist, shifted to DX(2..IMAX+1), shift_dxyz.f:36-44) or CHECK_AXIS's
   uniform DA = ALENGTH/DBLE(NA) (check_axis.f:152-156); the ghost cells
   take their neighbour's size.  Returns the array (n + 2 entries). */
static double* cell_sizes(int n, const double* in, double len, double* mn) {
  double* c = calloc((size_t)n + 2, sizeof(double));
  for (int i = 1; i <= n; i++) c[i] = in ? in[i - 1] : len / (double)n;
  c[0] = c[1];
  c[n + 1] = c[n];
  *mn = c[1];
  for (int i = 1; i <= n; i++) *mn = c[i] < *mn ? c[i] : *mn;
  return c;
}

/* SET_GEOMETRY_DES cumulative faces XE(0) = X_MIN - DX(1), XE(I) = XE(I-1)
   + DX(I) (set_geometry_des.f:72-87); VOL = DX*DY*DZ (set_geometry1.f:125
   with X=1); VOL_SURR (make_arrays_des.f:78-101). */
static void setup_fluid_grid(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  s->ni = p->imax + 2;
  s->nj = p->jmax + 2;
  s->nk = p->kmax + 2;
  s->ncell = s->ni * s->nj * s->nk;
  s->cdx = cell_sizes(p->imax, p->dx, p->xlength, &s->dx_min);
  s->cdy = cell_sizes(p->jmax, p->dy, p->ylength, &s->dy_min);
  s->cdz = cell_sizes(p->kmax, p->dz, p->zlength, &s->dz_min);
  s->p.dx = s->p.dy = s->p.dz = NULL;  /* copied */
  s->x_max = p->x_min + p->xlength;
  s->y_max = p->y_min + p->ylength;
  s->z_max = p->z_min + p->zlength;
  s->xe = calloc(s->ni + 1, sizeof(double));
  s->yn = calloc(s->nj + 1, sizeof(double));
  s->zt = calloc(s->nk + 1, sizeof(double));
  s->xe[0] = p->x_min - s->cdx[0];
  for (int i = 1; i <= s->ni; i++) s->xe[i] = s->xe[i - 1] + s->cdx[i - 1];
  s->yn[0] = p->y_min - s->cdy[0];
  for (int j = 1; j <= s->nj; j++) s->yn[j] = s->yn[j - 1] + s->cdy[j - 1];
  s->zt[0] = p->z_min - s->cdz[0];
  for (int k = 1; k <= s->nk; k++) s->zt[k] = s->zt[k - 1] + s->cdz[k - 1];
  s->vol = calloc(s->ncell, sizeof(double));
  s->vol_surr = calloc(s->ncell, sizeof(double));
  s->fluid = calloc(s->ncell, 1);
  for (int k = 1; k <= s->nk; k++)
    for (int i = 1; i <= s->ni; i++)
      for (int j = 1; j <= s->nj; j++) {
        int ijk = fluid_ijk(s, i, j, k) - 1;
        s->vol[ijk] = s->cdx[i - 1] * s->cdy[j - 1] * (1.0 * s->cdz[k - 1]);
        s->fluid[ijk] = (i >= 2 && i <= p->imax + 1 && j >= 2 && j <= p->jmax + 1 &&
                         k >= 2 && k <= p->kmax + 1);
      }
  for (int k = 1; k <= p->kmax + 1; k++)
    for (int j = 1; j <= p->jmax + 1; j++)
      for (int i = 1; i <= p->imax + 1; i++) {
        int ijk = fluid_ijk(s, i, j, k) - 1;
        for (int kk = k; kk <= k + 1; kk++)
          for (int jj = j; jj <= j + 1; jj++)
            for (int ii = i; ii <= i + 1; ii++) {
              int ijk2 = fluid_ijk(s, ii, jj, kk) - 1;
              if (s->fluid[ijk2]) s->vol_surr[ijk] = s->vol_surr[ijk] + s->vol[ijk2];
            }
      }
}

/* DES grid: default size check_geometry.f:134-170; desgrid_init
   desgrid_mod.f:572-574 (sizes) and :833-847 (xstart, dxinv). */
static void setup_des_grid(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  double max_diam = 2.0 * p->max_radius;
  double width = 3.0 * max_diam;
  int gi = p->desgridsearch_imax, gj = p->desgridsearch_jmax, gk = p->desgridsearch_kmax;
  /* WIDTH_X = min(WIDTH, minval(DX)) (check_geometry.f:137-144) */
  double wx = width < s->dx_min ? width : s->dx_min;
  double wy = width < s->dy_min ? width : s->dy_min;
  double wz = width < s->dz_min ? width : s->dz_min;
  if (gi <= 0) { gi = (int)(p->xlength / wx); if (gi < 1) gi = 1; }
  if (gj <= 0) { gj = (int)(p->ylength / wy); if (gj < 1) gj = 1; }
  if (gk <= 0) { gk = (int)(p->zlength / wz); if (gk < 1) gk = 1; }
  double ltx = p->xlength / gi, lty = p->ylength / gj, ltz = p->zlength / gk;
  /* NINT((xe(iend1)-xe(istart1-1))/ltempdx), iend1 = IMAX+1, istart1-1 = 1 */
  double ex = s->xe[p->imax + 1] - s->xe[1];
  double ey = s->yn[p->jmax + 1] - s->yn[1];
  double ez = s->zt[p->kmax + 1] - s->zt[1];
  s->dg_isize = (int)lround(ex / ltx);
  s->dg_jsize = (int)lround(ey / lty);
  s->dg_ksize = (int)lround(ez / ltz);
  s->dg_istart1 = 2; s->dg_iend1 = s->dg_isize + 1;
  s->dg_jstart1 = 2; s->dg_jend1 = s->dg_jsize + 1;
  s->dg_kstart1 = 2; s->dg_kend1 = s->dg_ksize + 1;
  s->dg_istart2 = 1; s->dg_iend2 = s->dg_isize + 2;
  s->dg_jstart2 = 1; s->dg_jend2 = s->dg_jsize + 2;
  s->dg_kstart2 = 1; s->dg_kend2 = s->dg_ksize + 2;
  s->dg_ni = s->dg_isize + 2;
  s->dg_nj = s->dg_jsize + 2;
  s->dg_nk = s->dg_ksize + 2;
  s->dg_ncell = s->dg_ni * s->dg_nj * s->dg_nk;
  s->dg_xstart = s->xe[1];
  s->dg_ystart = s->yn[1];
  s->dg_zstart = s->zt[1];
  s->dg_dxinv = (double)(s->dg_iend1 - s->dg_istart1 + 1) / (s->xe[p->imax + 1] - s->dg_xstart);
  s->dg_dyinv = (double)(s->dg_jend1 - s->dg_jstart1 + 1) / (s->yn[p->jmax + 1] - s->dg_ystart);
  s->dg_dzinv = (double)(s->dg_kend1 - s->dg_kstart1 + 1) / (s->zt[p->kmax + 1] - s->dg_zstart);
  s->dg_cnt = calloc(s->dg_ncell, sizeof(int));
  s->dg_off = calloc(s->dg_ncell + 1, sizeof(int));
  s->d.dg_imax = s->dg_isize;
  s->d.dg_jmax = s->dg_jsize;
  s->d.dg_kmax = s->dg_ksize;
  s->d.dg_xstart = s->dg_xstart;
  s->d.dg_ystart = s->dg_ystart;
  s->d.dg_zstart = s->dg_zstart;
  s->d.dg_dxinv = s->dg_dxinv;
  s->d.dg_dyinv = s->dg_dyinv;
  s->d.dg_dzinv = s->dg_dzinv;
}

/* TRI_BOX_OVERLAP (stl_functions_des_mod.f:280-536): separating-axis
   triangle/box test of Akenine-Moller, restated with the reference's
   axis order and comparisons. */
static int tri_box_overlap(const double c[3], const double h[3], const double vt[3][3]) {
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
#define AXTEST(pa, pb, ra) do { double lo_ = (pa) < (pb) ? (pa) : (pb); \
    double hi_ = (pa) < (pb) ? (pb) : (pa); if (lo_ > (ra) || hi_ < -(ra)) return 0; } while (0)
  {
    double a = e0[2], b = e0[1], fa = fabs(e0[2]), fb = fabs(e0[1]);
    double p0 = a * v0[1] - b * v0[2], p2 = a * v2[1] - b * v2[2];
    AXTEST(p0, p2, fa * h[1] + fb * h[2]);
    a = e0[2]; b = e0[0]; fa = fabs(e0[2]); fb = fabs(e0[0]);
    p0 = -a * v0[0] + b * v0[2]; p2 = -a * v2[0] + b * v2[2];
    AXTEST(p0, p2, fa * h[0] + fb * h[2]);
    a = e0[1]; b = e0[0]; fa = fabs(e0[1]); fb = fabs(e0[0]);
    double p1 = a * v1[0] - b * v1[1]; p2 = a * v2[0] - b * v2[1];
    { double lo = p2 < p1 ? p2 : p1, hi = p2 < p1 ? p1 : p2, r = fa * h[0] + fb * h[1];
      if (lo > r || hi < -r) return 0; }
  }
  {
    double a = e1[2], b = e1[1], fa = fabs(e1[2]), fb = fabs(e1[1]);
    double p0 = a * v0[1] - b * v0[2], p2 = a * v2[1] - b * v2[2];
    AXTEST(p0, p2, fa * h[1] + fb * h[2]);
    a = e1[2]; b = e1[0]; fa = fabs(e1[2]); fb = fabs(e1[0]);
    p0 = -a * v0[0] + b * v0[2]; p2 = -a * v2[0] + b * v2[2];
    AXTEST(p0, p2, fa * h[0] + fb * h[2]);
    a = e1[1]; b = e1[0]; fa = fabs(e1[1]); fb = fabs(e1[0]);
    p0 = a * v0[0] - b * v0[1]; double p1 = a * v1[0] - b * v1[1];
    AXTEST(p0, p1, fa * h[0] + fb * h[1]);
  }
  {
    double a = e2[2], b = e2[1], fa = fabs(e2[2]), fb = fabs(e2[1]);
    double p0 = a * v0[1] - b * v0[2], p1 = a * v1[1] - b * v1[2];
    AXTEST(p0, p1, fa * h[1] + fb * h[2]);
    a = e2[2]; b = e2[0]; fa = fabs(e2[2]); fb = fabs(e2[0]);
    p0 = -a * v0[0] + b * v0[2]; p1 = -a * v1[0] + b * v1[2];
    AXTEST(p0, p1, fa * h[0] + fb * h[2]);
    a = e2[1]; b = e2[0]; fa = fabs(e2[1]); fb = fabs(e2[0]);
    p1 = a * v1[0] - b * v1[1]; double p2 = a * v2[0] - b * v2[1];
    { double lo = p2 < p1 ? p2 : p1, hi = p2 < p1 ? p1 : p2, r = fa * h[0] + fb * h[1];
      if (lo > r || hi < -r) return 0; }
  }
#undef AXTEST
  for (int i = 0; i < 3; i++) {
    double mn = v0[i], mx = v0[i];
    if (v1[i] < mn) mn = v1[i];
    if (v2[i] < mn) mn = v2[i];
    if (v1[i] > mx) mx = v1[i];
    if (v2[i] > mx) mx = v2[i];
    if (mn > h[i] || mx < -h[i]) return 0;
  }
  double nrm[3];
  nrm[0] = e0[1] * e1[2] - e0[2] * e1[1];
  nrm[1] = e0[2] * e1[0] - e0[0] * e1[2];
  nrm[2] = e0[0] * e1[1] - e0[1] * e1[0];
  double vmin[3], vmax[3];
  for (int q = 0; q < 3; q++) {
    double v = v0[q];
    if (nrm[q] > 0.0) {
      vmin[q] = -h[q] - v;
      vmax[q] = h[q] - v;
    } else {
      vmin[q] = h[q] - v;
      vmax[q] = -h[q] - v;
    }
  }
  if (dot3(nrm, vmin) > 0.0) return 0;
  if (dot3(nrm, vmax) >= 0.0) return 1;
  return 0;
}

/* DES_STL_PREPROCESSING (stl_preproc_des_mod.f:80-124): the user's STL
   facets first (ids 0..nu-1), then, with default walls,
   CONVERT_DEFAULT_WALLS_TO_STL (:973-1047): one large triangle per box face,
   W E S N B T; BIN_FACETS_TO_DG (:135-203) with ADD_FACET_FOR_DES (:210-260)
   and ADD_FACET (:268-345), each cell's facets in ascending id;
   FLAG_FACETS_WHERE_EDGE_COLLISION_IS_IGNORED (:362-480).  user_vert is
   VERTEX(3,3,nu) as MFiX stores it (vertex index fastest, then coordinate),
   user_norm NORM_FACE(3,nu). */
static void free_facets(orc_sys* s) {
  void* ptrs[] = {s->vert, s->norm, s->ignore_edge, s->wall_vt, s->wall_vn, s->fc_count,
                  s->fc_off, s->fc_id, s->fc_dir, s->fc_min, s->fc_max};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  s->vert = NULL; s->norm = NULL; s->ignore_edge = NULL; s->wall_vt = NULL; s->wall_vn = NULL;
  s->fc_count = NULL; s->fc_off = NULL; s->fc_id = NULL; s->fc_dir = NULL;
  s->fc_min = NULL; s->fc_max = NULL;
}

static void setup_facets(orc_sys* s, int nu, const double* user_vert, const double* user_norm,
                         int default_walls) {
  const mfixdem_params* p = &s->p;
  double xn = p->x_min, xx = s->x_max, yn = p->y_min, yx = s->y_max, zn = p->z_min, zx = s->z_max;
  double XL = p->xlength, YL = p->ylength, ZL = p->zlength;
  free_facets(s);
  int nt = nu + (default_walls ? 6 : 0);
  int na = nt > 0 ? nt : 1;
  s->vert = calloc(na, sizeof(*s->vert));
  s->norm = calloc(na, sizeof(*s->norm));
  s->ignore_edge = calloc(na, sizeof(*s->ignore_edge));
  s->wall_vt = calloc(na, sizeof(*s->wall_vt));
  s->wall_vn = calloc(na, sizeof(*s->wall_vn));
  int n = 0;
  for (; n < nu; n++)
    for (int c = 0; c < 3; c++) {
      for (int v = 0; v < 3; v++) s->vert[n][v][c] = user_vert[(n * 3 + c) * 3 + v];
      s->norm[n][c] = user_norm[n * 3 + c];
    }
#define SETF(a0, a1, a2, b0, b1, b2, c0, c1, c2, n0, n1, n2) do { \
    s->vert[n][0][0] = a0; s->vert[n][0][1] = a1; s->vert[n][0][2] = a2; \
    s->vert[n][1][0] = b0; s->vert[n][1][1] = b1; s->vert[n][1][2] = b2; \
    s->vert[n][2][0] = c0; s->vert[n][2][1] = c1; s->vert[n][2][2] = c2; \
    s->norm[n][0] = n0; s->norm[n][1] = n1; s->norm[n][2] = n2; n++; } while (0)
  if (default_walls) {
    SETF(xn, yn, zn, xn, yx + YL, zn, xn, yn, zx + ZL, 1.0, 0.0, 0.0);   /* west   */
    SETF(xx, yn, zn, xx, yx + YL, zn, xx, yn, zx + ZL, -1.0, 0.0, 0.0);  /* east   */
    SETF(xn, yn, zn, xx + XL, yn, zn, xn, yn, zx + ZL, 0.0, 1.0, 0.0);   /* south  */
    SETF(xn, yx, zn, xx + XL, yx, zn, xn, yx, zx + ZL, 0.0, -1.0, 0.0);  /* north  */
    SETF(xn, yn, zn, xx + XL, yn, zn, xn, yx + YL, zn, 0.0, 0.0, 1.0);   /* bottom */
    SETF(xn, yn, zx, xx + XL, yn, zx, xn, yx + YL, zx, 0.0, 0.0, -1.0);  /* top    */
  }
#undef SETF
  s->nf = n;
  s->d.n_facets = n;
  s->fc_count = calloc(s->dg_ncell, sizeof(int));
  s->fc_off = calloc(s->dg_ncell + 1, sizeof(int));
  const double tol_stl = 1.0e-6; /* cartesian_grid_init_namelist.f:748 */
  double buffer = 1.1 * p->max_radius;
  double ldx = 1.0 / s->dg_dxinv, ldy = 1.0 / s->dg_dyinv, ldz = 1.0 / s->dg_dzinv;
  /* two passes over the (facet, cell) overlaps: count, then fill */
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int c = 0; c < s->dg_ncell; c++) s->fc_off[c + 1] = s->fc_off[c] + s->fc_count[c];
      int tot = s->fc_off[s->dg_ncell] > 0 ? s->fc_off[s->dg_ncell] : 1;
      s->fc_id = calloc(tot, sizeof(int));
      s->fc_dir = calloc(tot, sizeof(int));
      s->fc_min = calloc(tot, sizeof(double));
      s->fc_max = calloc(tot, sizeof(double));
      memset(s->fc_count, 0, sizeof(int) * s->dg_ncell);
    }
    for (int nn = 0; nn < n; nn++) {
      double X1 = s->vert[nn][0][0], X2 = X1, Y1 = s->vert[nn][0][1], Y2 = Y1,
             Z1 = s->vert[nn][0][2], Z2 = Z1;
      for (int v = 1; v < 3; v++) {
        if (s->vert[nn][v][0] < X1) X1 = s->vert[nn][v][0];
        if (s->vert[nn][v][0] > X2) X2 = s->vert[nn][v][0];
        if (s->vert[nn][v][1] < Y1) Y1 = s->vert[nn][v][1];
        if (s->vert[nn][v][1] > Y2) Y2 = s->vert[nn][v][1];
        if (s->vert[nn][v][2] < Z1) Z1 = s->vert[nn][v][2];
        if (s->vert[nn][v][2] > Z2) Z2 = s->vert[nn][v][2];
      }
      int I1 = s->dg_iend2, I2 = s->dg_istart2;
      if (X2 >= p->x_min - tol_stl && X1 <= s->x_max + tol_stl) {
        I1 = imax_i(iofpos(s, X1) - 1, s->dg_istart2);
        I2 = imin_i(iofpos(s, X2) + 1, s->dg_iend2);
      }
      int J1 = s->dg_jend2, J2 = s->dg_jstart2;
      if (Y2 >= p->y_min - tol_stl && Y1 <= s->y_max + tol_stl) {
        J1 = imax_i(jofpos(s, Y1) - 1, s->dg_jstart2);
        J2 = imin_i(jofpos(s, Y2) + 1, s->dg_jend2);
      }
      int K1 = s->dg_kend2, K2 = s->dg_kstart2;
      if (Z2 >= p->z_min - tol_stl && Z1 <= s->z_max + tol_stl) {
        K1 = imax_i(kofpos(s, Z1) - 1, s->dg_kstart2);
        K2 = imin_i(kofpos(s, Z2) + 1, s->dg_kend2);
      }
      for (int kk = K1; kk <= K2; kk++)
        for (int jj = J1; jj <= J2; jj++)
          for (int ii = I1; ii <= I2; ii++) {
            int ijk = dg_funijk(s, ii, jj, kk) - 1;
            double c[3], h[3];
            c[0] = s->dg_xstart + ((double)(ii - s->dg_istart1) + 0.5) * ldx;
            h[0] = 0.5 * ldx + buffer;
            c[1] = s->dg_ystart + ((double)(jj - s->dg_jstart1) + 0.5) * ldy;
            h[1] = 0.5 * ldy + buffer;
            c[2] = s->dg_zstart + ((double)(kk - s->dg_kstart1) + 0.5) * ldz;
            h[2] = 0.5 * ldz + buffer;
            if (!tri_box_overlap(c, h, s->vert[nn])) continue;
            int q = s->fc_count[ijk]++;
            if (pass == 0) continue;
            q += s->fc_off[ijk];
            s->fc_id[q] = nn;
            double smallest = 3.4028234663852886e38; /* HUGE(0.0) */
            for (int ax = 0; ax < 3; ax++) {
              double mn = s->vert[nn][0][ax], mx = mn;
              for (int v = 1; v < 3; v++) {
                if (s->vert[nn][v][ax] < mn) mn = s->vert[nn][v][ax];
                if (s->vert[nn][v][ax] > mx) mx = s->vert[nn][v][ax];
              }
              if (fabs(mx - mn) < smallest) {
                s->fc_dir[q] = ax;
                s->fc_min[q] = mn;
                s->fc_max[q] = mx;
                smallest = fabs(mx - mn);
              }
            }
          }
    }
  }
  /* edge-collision flags */
  const double tol_edge = 1.0e-9;
  double cos_nb = cos(5.0 / 180.0 * PI_D) + 1.0e-6;
  for (int ijk = 0; ijk < s->dg_ncell; ijk++) {
    int cnt = s->fc_count[ijk], o = s->fc_off[ijk];
    for (int f1 = 0; f1 < cnt; f1++) {
      int fid = s->fc_id[o + f1];
      double ec[3][3];
      for (int c = 0; c < 3; c++) {
        ec[0][c] = 0.5 * (s->vert[fid][0][c] + s->vert[fid][1][c]);
        ec[1][c] = 0.5 * (s->vert[fid][0][c] + s->vert[fid][2][c]);
        ec[2][c] = 0.5 * (s->vert[fid][1][c] + s->vert[fid][2][c]);
      }
      for (int e = 0; e < 3; e++)
        for (int f2 = 0; f2 < cnt; f2++) {
          if (f1 == f2) continue;
          int nb = s->fc_id[o + f2];
          double nec[3][3];
          for (int c = 0; c < 3; c++) {
            nec[0][c] = 0.5 * (s->vert[nb][0][c] + s->vert[nb][1][c]);
            nec[1][c] = 0.5 * (s->vert[nb][0][c] + s->vert[nb][2][c]);
            nec[2][c] = 0.5 * (s->vert[nb][1][c] + s->vert[nb][2][c]);
          }
          for (int ne = 0; ne < 3; ne++) {
            double dd[3] = {nec[ne][0] - ec[e][0], nec[ne][1] - ec[e][1], nec[ne][2] - ec[e][2]};
            if (dot3(dd, dd) < tol_edge) {
              if (dot3(s->norm[fid], s->norm[nb]) > cos_nb) {
                s->ignore_edge[fid][e + 1] = 1;
                s->ignore_edge[nb][ne + 1] = 1;
              }
            }
          }
        }
    }
  }
}

/* User STL facets (replacing any given before), before set-up; the box
   walls follow them unless default_walls is 0. */
int orc_set_facets(orc_sys* s, int32_t nu, const double* vertex, const double* normal,
                   int32_t default_walls) {
  if (!s || nu < 0 || (nu > 0 && (!vertex || !normal)) || s->inited) return MFIXDEM_ERR_ARG;
  setup_facets(s, nu, vertex, normal, default_walls);
  return 0;
}

/* Moving walls of a BC_STL group (get_stl_data.f:1484-1502): facets
   [first, first+count) move with the wall velocity vel plus omega x (facet
   centre - rot_center); STL_NORMAL_VEL is that velocity's component along
   the facet normal and STL_TANGENTIAL_VEL the rest. */
int orc_set_wall_motion(orc_sys* s, int32_t first, int32_t count, const double* vel,
                        const double* omega, const double* rot_center) {
  if (!s || first < 0 || count < 0 || first + count > s->nf || !vel || !omega || !rot_center)
    return MFIXDEM_ERR_ARG;
  for (int n = first; n < first + count; n++) {
    double center[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) center[c] = center[c] + s->vert[n][v][c];
    for (int c = 0; c < 3; c++) center[c] = center[c] / 3.0;
    double op[3] = {center[0] - rot_center[0], center[1] - rot_center[1],
                    center[2] - rot_center[2]};
    double cr[3], avg[3];
    cross3(omega, op, cr);
    for (int c = 0; c < 3; c++) avg[c] = vel[c] + cr[c];
    double an = dot3(avg, s->norm[n]);
    for (int c = 0; c < 3; c++) s->wall_vn[n][c] = an * s->norm[n][c];
    for (int c = 0; c < 3; c++) s->wall_vt[n][c] = avg[c] - s->wall_vn[n][c];
  }
  return 0;
}

/* --------------------------------------------------------------- create */
/* DES_INTERP_SCHEME checks of check_solids_common_discrete.f:752-860 and
   the width check of set_filter_des.f:70-87 (returns 0 when acceptable). */
static int check_interp(const orc_sys* s) {
  const mfixdem_params* p = &s->p;
  const double w = p->des_interp_width;
  switch (p->interp_scheme) {
    case MFIXDEM_INTERP_GARG_2012: return w != UNDEF;
    case MFIXDEM_INTERP_NONE: return p->des_interp_on || p->des_interp_mean_fields || w != UNDEF;
    case MFIXDEM_INTERP_LINEAR_HAT: return w != UNDEF;
    case MFIXDEM_INTERP_SQUARE_DPVM:
    case MFIXDEM_INTERP_GAUSS_DPVM: {
      if (w == UNDEF) return 1;
      double dmin = fmin(fmin(s->dx_min, s->dy_min), s->dz_min);
      double dgmin = fmin(fmin(1.0 / s->dg_dxinv, 1.0 / s->dg_dyinv), 1.0 / s->dg_dzinv);
      return 0.5 * w > dmin || 0.5 * w > dgmin;
    }
    default: return 1;
  }
}

orc_sys* orc_create(const mfixdem_params* p) {
  if (!p || p->nphases < 1 || p->nphases > MP || p->max_pip < 1) return NULL;
  if (p->imax < 1 || p->jmax < 1 || p->kmax < 1) return NULL;
  if (p->des_neighbor_search != MFIXDEM_NSEARCH_GRID &&
      p->des_neighbor_search != MFIXDEM_NSEARCH_NSQUARE)
    return NULL;
  /* explicit coupling: DRAG_GS_DES1's F_GP path, filter schemes only */
  if (p->des_explicitly_coupled && (!p->coupled || p->interp_scheme == MFIXDEM_INTERP_GARG_2012))
    return NULL;
  /* non-uniform cell sizes: positive, summing to the domain length */
  const double* in[3] = {p->dx, p->dy, p->dz};
  const int na[3] = {p->imax, p->jmax, p->kmax};
  const double len[3] = {p->xlength, p->ylength, p->zlength};
  for (int a = 0; a < 3; a++) {
    if (!in[a]) continue;
    double sum = 0.0;
    for (int i = 0; i < na[a]; i++) {
      if (!(in[a][i] > 0.0) || !isfinite(in[a][i])) return NULL;
      sum += in[a][i];
    }
    if (fabs(sum - len[a]) > 1e-12 * fabs(len[a])) return NULL;
  }
  orc_sys* s = calloc(1, sizeof(orc_sys));
  s->p = *p;
  if (p->coll_model == MFIXDEM_HERTZ) derive_hertz(s);
  else derive_lsd(s);
  s->d.kt = s->kt;
  s->d.kt_w = s->kt_w;
  s->d.dtsolid = (p->dtsolid == UNDEF) ? s->d.tcoll / p->dtsolid_factor : p->dtsolid;
  setup_fluid_grid(s);
  setup_des_grid(s);
  setup_facets(s, 0, NULL, NULL, 1);
  if (check_interp(s)) {
    orc_destroy(s);
    return NULL;
  }
  if (p->interp_scheme == MFIXDEM_INTERP_SQUARE_DPVM ||
      p->interp_scheme == MFIXDEM_INTERP_GAUSS_DPVM) {
    /* set_filter_des.f:85,99-100 */
    s->fw = 0.500 * p->des_interp_width;
    s->oofv = 0.25 / (s->fw * s->fw * s->fw);
    s->fw3 = s->fw * 3;
  }
  int64_t ld = p->max_pip;
  s->max_pip = ld;
  s->ld = ld;
#define A(ptr, cnt, T) ptr = (T*)calloc((size_t)(cnt), sizeof(T))
  A(s->pos, 3 * ld, double); A(s->vel, 3 * ld, double); A(s->omega, 3 * ld, double);
  A(s->radius, ld, double); A(s->ro_sol, ld, double); A(s->pvol, ld, double);
  A(s->pmass, ld, double); A(s->omoi, ld, double);
  A(s->phase, ld, int); A(s->state, ld, signed char);
  A(s->fc, 3 * ld, double); A(s->tow, 3 * ld, double);
  A(s->fc_last, 3 * ld, double); A(s->tow_last, 3 * ld, double);
  A(s->pos_old, 3 * ld, double); A(s->vel_old, 3 * ld, double); A(s->omega_old, 3 * ld, double);
  A(s->acc_old, 3 * ld, double); A(s->rot_acc_old, 3 * ld, double);
  A(s->dg_pijk, ld, int); A(s->pijk, 4 * ld, int);
  A(s->dg_list, ld, int);
  A(s->neighbor_index, ld, int); A(s->neighbor_index_old, ld, int);
  s->cap = 4 * ld + 16;
  A(s->neighbors, s->cap, int); A(s->neighbors_old, s->cap, int);
  A(s->pft, 3 * s->cap, double); A(s->pft_old, 3 * s->cap, double);
  A(s->wall_fid, COLLISION_ARRAY_MAX * ld, int);
  A(s->wall_pft, 3 * COLLISION_ARRAY_MAX * ld, double);
  A(s->u_g, s->ncell, double); A(s->v_g, s->ncell, double); A(s->w_g, s->ncell, double);
  A(s->p_force, 3 * s->ncell, double); A(s->ep_g, s->ncell, double);
  A(s->rop_s, (size_t)s->ncell * p->nphases, double);
  A(s->rops_node, (size_t)s->ncell * p->nphases, double);
  A(s->pinc, s->ncell, int); A(s->pic_off, s->ncell + 1, int); A(s->pic_list, ld, int);
  A(s->ugc, 3 * s->ncell, double); A(s->filter_cell, FILTER_N * ld, int);
  A(s->filter_weight, FILTER_N * ld, double);
  A(s->solvolinc, (size_t)s->ncell * p->nphases, double);
  A(s->f_gp, ld, double); A(s->drag_fc, 3 * ld, double);
#undef A
  for (int64_t q = 0; q < COLLISION_ARRAY_MAX * ld; q++) s->wall_fid[q] = -1;
  for (int c = 0; c < s->ncell; c++) s->ep_g[c] = 1.0;
  s->first_pass = 1;
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  void* ptrs[] = {s->cdx, s->cdy, s->cdz, s->xe, s->yn, s->zt, s->vol, s->vol_surr, s->fluid, s->dg_cnt, s->dg_off,
                  s->dg_list,
                  s->pos, s->vel, s->omega, s->radius, s->ro_sol, s->pvol, s->pmass,
                  s->omoi, s->phase, s->state, s->fc, s->tow, s->fc_last, s->tow_last,
                  s->pos_old, s->vel_old, s->omega_old, s->acc_old, s->rot_acc_old,
                  s->dg_pijk, s->pijk, s->neighbor_index, s->neighbor_index_old,
                  s->neighbors, s->neighbors_old, s->pft, s->pft_old, s->wall_fid,
                  s->wall_pft, s->u_g, s->v_g, s->w_g, s->p_force, s->ep_g, s->rop_s,
                  s->rops_node, s->pinc, s->pic_off, s->pic_list, s->ugc, s->filter_cell,
                  s->filter_weight, s->solvolinc, s->f_gp, s->drag_fc};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free_facets(s);
  free_bcs(s);
  free(s);
}

int orc_get_derived(const orc_sys* s, mfixdem_derived* out) {
  if (!s || !out) return MFIXDEM_ERR_ARG;
  *out = s->d;
  return 0;
}
double orc_dtsolid(const orc_sys* s) { return s->d.dtsolid; }

/* Per-particle PVOL, PMASS, OMOI: make_arrays_des.f:288-291. */
int orc_set_particles(orc_sys* s, int64_t n, const double* pos, const double* vel,
                      const double* omega, const double* radius, const double* ro_sol,
                      const int32_t* phase, int64_t ld) {
  if (!s || n < 0 || n > s->max_pip || ld < n) return MFIXDEM_ERR_ARG;
  int64_t L = s->ld;
  for (int64_t q = 0; q < n; q++) {
    for (int c = 0; c < 3; c++) {
      s->pos[c * L + q] = pos[c * ld + q];
      s->vel[c * L + q] = vel ? vel[c * ld + q] : 0.0;
      s->omega[c * L + q] = omega ? omega[c * ld + q] : 0.0;
    }
    double r = radius[q];
    s->radius[q] = r;
    s->ro_sol[q] = ro_sol[q];
    s->phase[q] = phase ? phase[q] - 1 : 0;
    if (s->phase[q] < 0 || s->phase[q] >= s->p.nphases) return MFIXDEM_ERR_ARG;
    s->pvol[q] = (4.0 / 3.0) * PI_D * (r * r * r);
    s->pmass[q] = s->pvol[q] * s->ro_sol[q];
    s->omoi[q] = 2.5 / (s->pmass[q] * (r * r));
    s->state[q] = NORMAL_PARTICLE;
  }
  for (int64_t q = n; q < s->max_pip; q++) s->state[q] = NONEXISTENT;
  memset(s->f_gp, 0, sizeof(double) * (size_t)L);        /* des_init_arrays.f:143 */
  memset(s->drag_fc, 0, sizeof(double) * 3 * (size_t)L);
  s->pip = n;
  s->first_pass = 1;
  s->inited = 0;
  return 0;
}

int orc_get_particles(const orc_sys* s, double* pos, double* vel, double* omega,
                      int8_t* state, int64_t ld) {
  int64_t L = s->ld, n = s->max_pip;
  if (ld < n) n = ld;
  for (int64_t q = 0; q < n; q++) {
    for (int c = 0; c < 3; c++) {
      if (pos) pos[c * ld + q] = s->pos[c * L + q];
      if (vel) vel[c * ld + q] = s->vel[c * L + q];
      if (omega) omega[c * ld + q] = s->omega[c * L + q];
    }
    if (state) state[q] = s->state[q];
  }
  return 0;
}

/* ------------------------------------------------ filter schemes (DES1) */
/* CALC_CELL_CENTER_GAS_VEL (drag_gs_des1.f:313-370) on a uniform grid
   without cut cells: FX = 1/2, so AVG_X_E = FX*U(IMJK) + FX_BAR*U(IJK)
   (fun_avg.inc:38-44); AVG_Y_N, AVG_Z_T = HALF*(m + p).  Non-fluid cells: 0. */
static void cell_center_gas_vel(orc_sys* s) {
  for (int k = 1; k <= s->nk; k++)
    for (int i = 1; i <= s->ni; i++)
      for (int j = 1; j <= s->nj; j++) {
        int ijk = fluid_ijk(s, i, j, k) - 1;
        double* g = s->ugc + 3 * ijk;
        if (!s->fluid[ijk]) {
          g[0] = g[1] = g[2] = 0.0;
          continue;
        }
        int imjk = fluid_ijk(s, i - 1, j, k) - 1;
        int ijmk = fluid_ijk(s, i, j - 1, k) - 1;
        int ijkm = fluid_ijk(s, i, j, k - 1) - 1;
        g[0] = 0.5 * s->u_g[imjk] + (1.0 - 0.5) * s->u_g[ijk];
        g[1] = 0.5 * (s->v_g[ijmk] + s->v_g[ijk]);
        g[2] = 0.5 * (s->w_g[ijkm] + s->w_g[ijk]);
      }
}

/* CALC_FILTER_WEIGHTS (calc_interp_weights.f:141-156): the mass of the
   cubic (Epanechnikov) kernel of half-width FILTER_WIDTH_INTERP beyond a face. */
static double dpvm_weight(const orc_sys* s, double pos1, double pos2) {
  double overlap = pos1 - pos2;
  if (overlap < s->fw) {
    double height = s->fw - overlap;
    return height * height * (s->fw3 - height) * s->oofv;
  }
  return 0.0;
}

static int on_mype_plus2(const orc_sys* s, int i, int j, int k) {
  /* IS_ON_MYPE_PLUS2LAYERS (functions.inc:169-181), serial: 1..IMAX+2 */
  return i >= 1 && i <= s->ni && j >= 1 && j <= s->nj && k >= 1 && k <= s->nk;
}

/* CALC_INTERP_WEIGHTS1 (calc_interp_weights.f:47-139): SQUARE_DPVM and
   GAUSS_DPVM; loop order KC, IC, JC (J fastest). */
static void interp_weights1(orc_sys* s, int64_t q) {
  int64_t L = s->ld;
  int i = s->pijk[q], j = s->pijk[L + q], k = s->pijk[2 * L + q];
  int ijk = s->pijk[3 * L + q] - 1;
  double x = s->pos[q], y = s->pos[L + q], z = s->pos[2 * L + q];
  double wi[3], wj[3], wk[3];
  wi[0] = dpvm_weight(s, x, s->xe[i - 1]);
  wi[2] = dpvm_weight(s, s->xe[i], x);
  wi[1] = 1.0 - wi[0] - wi[2];
  wj[0] = dpvm_weight(s, y, s->yn[j - 1]);
  wj[2] = dpvm_weight(s, s->yn[j], y);
  wj[1] = 1.0 - wj[0] - wj[2];
  wk[0] = dpvm_weight(s, z, s->zt[k - 1]);
  wk[2] = dpvm_weight(s, s->zt[k], z);
  wk[1] = 1.0 - wk[0] - wk[2];
  int idx = 0;
  for (int kc = -1; kc <= 1; kc++)
    for (int ic = -1; ic <= 1; ic++)
      for (int jc = -1; jc <= 1; jc++, idx++) {
        double w = wi[ic + 1] * wj[jc + 1] * wk[kc + 1];
        int cell = ijk;
        if (on_mype_plus2(s, i + ic, j + jc, k + kc)) {
          int t = fluid_ijk(s, i + ic, j + jc, k + kc) - 1;
          if (s->fluid[t]) cell = t;
        } else {
          w = 0.0;
        }
        s->filter_cell[idx * L + q] = cell;
        s->filter_weight[idx * L + q] = w;
      }
}

/* one axis of CALC_INTERP_WEIGHTS2 (calc_interp_weights.f:365-460): hat
   weights about the cell centre, own cell if the neighbour is not fluid */
/* c = &DX(I) of the particle's cell I, c[-1] and c[1] its neighbours';
   lOoDX = oDX_E(I-1) or oDX_E(I), oDX_E(L) = ONE/(HALF*(DX(L+1)+DX(L)))
   (set_geometry.f:341-343) */
static void lhat_axis(double xpos, double face_lo, const double* c, int fluid_m, int fluid_p,
                      double sw[3]) {
  double xc = face_lo + 0.5 * c[0];
  if (xpos < xc) {
    if (fluid_m) {
      double oo = 1.0 / (0.5 * (c[0] + c[-1]));
      sw[0] = (xc - xpos) * oo;
      sw[1] = 1.0 - sw[0];
      sw[2] = 0.0;
    } else {
      sw[0] = 0.0; sw[1] = 1.0; sw[2] = 0.0;
    }
  } else {
    if (fluid_p) {
      double oo = 1.0 / (0.5 * (c[1] + c[0]));
      sw[0] = 0.0;
      sw[2] = (xpos - xc) * oo;
      sw[1] = 1.0 - sw[2];
    } else {
      sw[0] = 0.0; sw[1] = 1.0; sw[2] = 0.0;
    }
  }
}

static int fluid_at(const orc_sys* s, int i, int j, int k) {
  if (!on_mype_plus2(s, i, j, k)) return 0;
  return s->fluid[fluid_ijk(s, i, j, k) - 1];
}

/* CALC_INTERP_WEIGHTS2 (calc_interp_weights.f:351-528): LINEAR_HAT; loop
   order KC, JC, IC (I fastest), non-fluid target -> own cell, weight 0. */
static void interp_weights2(orc_sys* s, int64_t q) {
  int64_t L = s->ld;
  int i = s->pijk[q], j = s->pijk[L + q], k = s->pijk[2 * L + q];
  int ijk = s->pijk[3 * L + q] - 1;
  double sx[3], sy[3], sz[3];
  lhat_axis(s->pos[q], s->xe[i - 1], s->cdx + (i - 1), fluid_at(s, i - 1, j, k),
            fluid_at(s, i + 1, j, k), sx);
  lhat_axis(s->pos[L + q], s->yn[j - 1], s->cdy + (j - 1), fluid_at(s, i, j - 1, k),
            fluid_at(s, i, j + 1, k), sy);
  lhat_axis(s->pos[2 * L + q], s->zt[k - 1], s->cdz + (k - 1), fluid_at(s, i, j, k - 1),
            fluid_at(s, i, j, k + 1), sz);
  int idx = 0;
  for (int kc = 0; kc <= 2; kc++)
    for (int jc = 0; jc <= 2; jc++)
      for (int ic = 0; ic <= 2; ic++, idx++) {
        double w = sx[ic] * sy[jc] * sz[kc];
        int cell = ijk;
        if (fluid_at(s, i - 1 + ic, j - 1 + jc, k - 1 + kc))
          cell = fluid_ijk(s, i - 1 + ic, j - 1 + jc, k - 1 + kc) - 1;
        else
          w = 0.0;
        s->filter_cell[idx * L + q] = cell;
        s->filter_weight[idx * L + q] = w;
      }
}

/* CALC_INTERP_WEIGHTS (calc_interp_weights.f:22-33) for normal particles */
int orc_calc_interp_weights(orc_sys* s) {
  int sch = s->p.interp_scheme;
  if (sch != MFIXDEM_INTERP_SQUARE_DPVM && sch != MFIXDEM_INTERP_GAUSS_DPVM &&
      sch != MFIXDEM_INTERP_LINEAR_HAT)
    return 0;
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] != NORMAL_PARTICLE) continue;
    if (sch == MFIXDEM_INTERP_LINEAR_HAT) interp_weights2(s, q);
    else interp_weights1(s, q);
  }
  return 0;
}

static int filter_size(const orc_sys* s) {
  int sch = s->p.interp_scheme;
  return (sch == MFIXDEM_INTERP_SQUARE_DPVM || sch == MFIXDEM_INTERP_GAUSS_DPVM ||
          sch == MFIXDEM_INTERP_LINEAR_HAT) ? FILTER_N : 0;
}

/* FILTER_CELL (0-based) and FILTER_WEIGHT of particle q after the last
   CALC_INTERP_WEIGHTS; returns the filter size (0 for GARG_2012 / NONE). */
int orc_get_filter(const orc_sys* s, int64_t q, int32_t* cells, double* weights) {
  int nf = filter_size(s);
  if (q < 0 || q >= s->max_pip) return -1;
  for (int lc = 0; lc < nf; lc++) {
    cells[lc] = s->filter_cell[lc * s->ld + q];
    weights[lc] = s->filter_weight[lc * s->ld + q];
  }
  return nf;
}

/* cell-centre gas velocity of fluid cell ijk (0-based), for tests */
int orc_get_ugc(const orc_sys* s, double* ugc) {
  memcpy(ugc, s->ugc, sizeof(double) * 3 * s->ncell);
  return 0;
}

/* calc_del_phi (calc_grad_des.f:104-133): 3-point derivative with a wall
   neighbour (not a valid pressure cell) replaced by extrapolation */
static double calc_del_phi(int valid_lo, int valid_hi, double phi, double dx, double phi_lo,
                           double dx_lo, double phi_hi, double dx_hi) {
  if (!valid_lo && !valid_hi) return 0.0;
  if (!valid_lo) phi_lo = phi_hi + (phi - phi_hi) * (dx_hi + 2.0 * dx + dx_lo) / (dx + dx_hi);
  if (!valid_hi) phi_hi = phi_lo + (phi - phi_lo) * (dx_hi + 2.0 * dx + dx_lo) / (dx + dx_lo);
  double x = 0.5 * (dx_lo + dx);
  double x_hi = 0.5 * (dx_hi + dx) + x;
  return phi_lo * (x - x_hi) / (x_hi * x) + phi * (2.0 * x - x_hi) / (x * (x - x_hi)) +
         phi_hi * (x) / (x_hi * (x_hi - x));
}

/* CALC_PG_GRAD (calc_pg_grad.f:20-110) with CALC_GRAD_DES
   (calc_grad_des.f:21-102): P_FORCE = -grad P_G in fluid cells, 0 elsewhere;
   no cyclic pressure drop and no impermeable internal surfaces (box walls:
   the valid pressure cells are the fluid cells). */
int orc_set_gas_pressure(orc_sys* s, const double* p_g) {
  for (int k = 1; k <= s->nk; k++)
    for (int i = 1; i <= s->ni; i++)
      for (int j = 1; j <= s->nj; j++) {
        int ijk = fluid_ijk(s, i, j, k) - 1;
        double* f = s->p_force + 3 * ijk;
        if (!s->fluid[ijk]) {
          f[0] = f[1] = f[2] = 0.0;
          continue;
        }
        int lo = fluid_ijk(s, i - 1, j, k) - 1, hi = fluid_ijk(s, i + 1, j, k) - 1;
        /* dx(i), dx of the west and east cells (calc_grad_des.f:85-87) */
        f[0] = calc_del_phi(s->fluid[lo], s->fluid[hi], p_g[ijk], s->cdx[i - 1], p_g[lo],
                            s->cdx[i - 2], p_g[hi], s->cdx[i]);
        lo = fluid_ijk(s, i, j - 1, k) - 1;
        hi = fluid_ijk(s, i, j + 1, k) - 1;
        f[1] = calc_del_phi(s->fluid[lo], s->fluid[hi], p_g[ijk], s->cdy[j - 1], p_g[lo],
                            s->cdy[j - 2], p_g[hi], s->cdy[j]);
        lo = fluid_ijk(s, i, j, k - 1) - 1;
        hi = fluid_ijk(s, i, j, k + 1) - 1;
        f[2] = calc_del_phi(s->fluid[lo], s->fluid[hi], p_g[ijk], s->cdz[k - 1], p_g[lo],
                            s->cdz[k - 2], p_g[hi], s->cdz[k]);
        for (int d = 0; d < 3; d++) f[d] = -f[d];
      }
  return 0;
}

int orc_get_p_force(const orc_sys* s, double* p_force) {
  memcpy(p_force, s->p_force, sizeof(double) * 3 * s->ncell);
  return 0;
}

int orc_set_gas(orc_sys* s, const double* u_g, const double* v_g, const double* w_g,
                const double* p_force) {
  for (int c = 0; c < s->ncell; c++) {
    s->u_g[c] = u_g[c];
    s->v_g[c] = v_g[c];
    s->w_g[c] = w_g[c];
    for (int d = 0; d < 3; d++) s->p_force[3 * c + d] = p_force ? p_force[3 * c + d] : 0.0;
  }
  cell_center_gas_vel(s);
  s->gas_set = 1;
  return 0;
}

/* Overwrite position, velocity and angular velocity of particles
   q0 .. q0+count-1 (0-based) between steps, the way a reference user hook
   (USR2_DES, des_time_march.f:453) edits DES_POS_NEW / DES_VEL_NEW /
   OMEGA_NEW.  Arrays are [c*count + k]; any may be NULL to leave that field
   alone. */
int orc_user_set_particles(orc_sys* s, int64_t q0, int64_t count, const double* pos,
                           const double* vel, const double* omega) {
  if (q0 < 0 || count < 0 || q0 + count > s->pip) return MFIXDEM_ERR_ARG;
  const int64_t L = s->ld;
  for (int64_t k = 0; k < count; k++)
    for (int c = 0; c < 3; c++) {
      if (pos) s->pos[c * L + q0 + k] = pos[c * count + k];
      if (vel) s->vel[c * L + q0 + k] = vel[c * count + k];
      if (omega) s->omega[c * L + q0 + k] = omega[c * count + k];
    }
  return MFIXDEM_OK;
}

int orc_get_gas(const orc_sys* s, double* ep_g, double* rop_s) {
  for (int c = 0; c < s->ncell; c++) {
    if (ep_g) ep_g[c] = s->ep_g[c];
    if (rop_s)
      for (int m = 0; m < s->p.nphases; m++) rop_s[(size_t)m * s->ncell + c] = s->rop_s[(size_t)m * s->ncell + c];
  }
  return 0;
}

/* --------------------------------------------------------- DESGRID_PIC */
/* desgrid_mod.f:877-995, serial fill order (ascending particle index). */
int orc_desgrid_pic(orc_sys* s) {
  int64_t L = s->ld;
  memset(s->dg_cnt, 0, sizeof(int) * s->dg_ncell);
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] == NONEXISTENT) continue;
    int li = imin_i(s->dg_iend2, imax_i(s->dg_istart2, iofpos(s, s->pos[q])));
    int lj = imin_i(s->dg_jend2, imax_i(s->dg_jstart2, jofpos(s, s->pos[L + q])));
    int lk = imin_i(s->dg_kend2, imax_i(s->dg_kstart2, kofpos(s, s->pos[2 * L + q])));
    s->dg_pijk[q] = dg_funijk(s, li, lj, lk);
    s->dg_cnt[s->dg_pijk[q] - 1]++;
  }
  s->dg_off[0] = 0;
  for (int c = 0; c < s->dg_ncell; c++) s->dg_off[c + 1] = s->dg_off[c] + s->dg_cnt[c];
  int* cur = malloc(sizeof(int) * s->dg_ncell);
  memcpy(cur, s->dg_off, sizeof(int) * s->dg_ncell);
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] == NONEXISTENT) continue;
    s->dg_list[cur[s->dg_pijk[q] - 1]++] = (int)q;
  }
  free(cur);
  return 0;
}

static void delete_particle(orc_sys* s, int64_t q) {
  /* DELETE_PARTICLE, mass_outflow_dem.f:201-280 */
  int64_t L = s->ld;
  for (int c = 0; c < 3; c++) {
    s->pos[c * L + q] = 0.0;
    s->vel[c * L + q] = 0.0;
    s->omega[c * L + q] = 0.0;
    s->pos_old[c * L + q] = 0.0;
    s->vel_old[c * L + q] = 0.0;
    s->omega_old[c * L + q] = 0.0;
    s->fc[c * L + q] = 0.0;
    s->tow[c * L + q] = 0.0;
  }
  s->radius[q] = 0.0;
  s->pmass[q] = 3.4028234663852886e38;
  s->pvol[q] = 3.4028234663852886e38;
  s->ro_sol[q] = 0.0;
  s->omoi[q] = 0.0;
  for (int k = 0; k < COLLISION_ARRAY_MAX; k++) s->wall_fid[k * L + q] = -1;
  s->state[q] = NONEXISTENT;
  s->pip--;
}

/* ------------------------------------------------- DESGRID_NEIGH_BUILD */
static void add_pair(orc_sys* s, int64_t ii, int jj) {
  /* add_pair / NEIGHBOR_GROW, des_allocate_mod.f:555-668 */
  int64_t pos1 = s->neighbor_index[ii];
  if (pos1 > s->cap) {
    int64_t ncap = s->cap;
    while (ncap < pos1) ncap *= 2;
    s->neighbors = realloc(s->neighbors, sizeof(int) * ncap);
    s->neighbors_old = realloc(s->neighbors_old, sizeof(int) * ncap);
    s->pft = realloc(s->pft, sizeof(double) * 3 * ncap);
    s->pft_old = realloc(s->pft_old, sizeof(double) * 3 * ncap);
    for (int64_t c = s->cap; c < ncap; c++) {
      s->neighbors[c] = 0;
      s->neighbors_old[c] = 0;
      for (int d = 0; d < 3; d++) s->pft[3 * c + d] = s->pft_old[3 * c + d] = 0.0;
    }
    s->cap = ncap;
  }
  s->neighbors[pos1 - 1] = jj;
  s->neighbor_index[ii] = (int)(pos1 + 1);
}

/* NSQUARE (nsquare.f:15-142) without periodic walls: every existing L
   against every existing LL > L, DIST < R_LM**2 (strict), rows in
   ascending LL.  Slots after the early exit (PC >= PIP) keep the 0 that
   NEIGHBOUR wrote (neighbour.f:60), which the force loop reads as empty
   rows; here they get the explicit empty-row index. */
static int nsquare_build(orc_sys* s) {
  int64_t L = s->ld;
  const double frlm = s->p.factor_rlm;
  int64_t pc = 1, l = 0;
  for (; l < s->max_pip; l++) {
    s->neighbor_index[l] = (l > 0) ? s->neighbor_index[l - 1] : 1;
    if (pc >= s->pip) break;
    if (s->state[l] == NONEXISTENT) continue;
    int64_t pnpc = s->pip - pc;
    for (int64_t ll = l + 1; ll < s->max_pip; ll++) {
      if (pnpc <= 0) break;
      if (s->state[ll] == NONEXISTENT) continue;
      double r_lm = s->radius[l] + s->radius[ll];
      r_lm = frlm * r_lm;
      double d0 = s->pos[ll] - s->pos[l], d1 = s->pos[L + ll] - s->pos[L + l],
             d2 = s->pos[2 * L + ll] - s->pos[2 * L + l];
      double dist = d0 * d0 + d1 * d1 + d2 * d2;
      if (dist < r_lm * r_lm) add_pair(s, l, (int)(ll + 1));
      pnpc--;
    }
    pc++;
  }
  for (int64_t q = (l < s->max_pip ? l + 1 : l); q < s->max_pip; q++)
    s->neighbor_index[q] = s->neighbor_index[q - 1];
  return 0;
}

/* candidate pairs of one row particle, in the reference's stencil order
   (desgrid_mod.f:1106-1214); returns the number written into out[] */
static int row_candidates(orc_sys* s, int64_t q, int* out, int* rogue) {
  int64_t L = s->ld;
  int lijk = s->dg_pijk[q];
  int lic, ljc, lkc;
  dg_ijk_of(s, lijk, &lic, &ljc, &lkc);
  if (lic < s->dg_istart1 || lic > s->dg_iend1 || ljc < s->dg_jstart1 ||
      ljc > s->dg_jend1 || lkc < s->dg_kstart1 || lkc > s->dg_kend1) {
    *rogue = 1;
    return 0;
  }
  int il = 1, iu = 1, jl = 1, ju = 1, kl = 1, ku = 1;
  double px = s->pos[q], py = s->pos[L + q], pz = s->pos[2 * L + q];
  double frlm = s->p.factor_rlm, mr = s->p.max_radius;
  int coarse_x = 1.0 > 2.0 * mr * frlm * s->dg_dxinv;
  int coarse_y = 1.0 > 2.0 * mr * frlm * s->dg_dyinv;
  int coarse_z = 1.0 > 2.0 * mr * frlm * s->dg_dzinv;
  if (coarse_x) {
    double t = (px - s->dg_xstart) * s->dg_dxinv;
    if (t - floor(t) >= 0.5) il = 0; else iu = 0;
  }
  if (coarse_y) {
    double t = (py - s->dg_ystart) * s->dg_dyinv;
    if (t - floor(t) >= 0.5) jl = 0; else ju = 0;
  }
  if (coarse_z) {
    double t = (pz - s->dg_zstart) * s->dg_dzinv;
    if (t - floor(t) >= 0.5) kl = 0; else ku = 0;
  }
  int n = 0;
  for (int lk = lkc - kl; lk <= lkc + ku; lk++)
    for (int lj = ljc - jl; lj <= ljc + ju; lj++)
      for (int li = lic - il; li <= lic + iu; li++) {
        int c = dg_funijk(s, li, lj, lk) - 1;
        for (int t = s->dg_off[c]; t < s->dg_off[c + 1]; t++) {
          int nb = s->dg_list[t];
          if (nb == q) continue;
          if (nb < q && s->state[nb] == NORMAL_PARTICLE) continue;
          if (s->state[nb] == NONEXISTENT) continue;
          double sr = frlm * (s->radius[q] + s->radius[nb]);
          double d0 = px - s->pos[nb], d1 = py - s->pos[L + nb], d2 = pz - s->pos[2 * L + nb];
          double ds = d0 * d0 + d1 * d1 + d2 * d2;
          if (ds > sr * sr) continue;
          out[n++] = nb + 1;
        }
      }
  return n;
}

static int desgrid_neigh_build(orc_sys* s) {
  int* buf = malloc(sizeof(int) * (s->max_pip + 1));
  int do_exit = 0;
  /* NOTE: the reference never clears do_exit inside the loop
     (desgrid_mod.f:1096-1104): after the first rogue particle every later
     row is left empty.  Restated as is. */
#ifdef _OPENMP
  int nth = omp_get_max_threads();
  if (nth > 1) {
    /* row search in parallel chunks, merged in ascending particle order as
       the reference's ordered merge (desgrid_mod.f:1219-1238) does */
    int64_t np = s->max_pip;
    int* cnt = calloc(np, sizeof(int));
    int* rog = calloc(np, sizeof(int));
    int** rows = calloc(np, sizeof(int*));
#pragma omp parallel
    {
      int* lb = malloc(sizeof(int) * (s->max_pip + 1));
#pragma omp for schedule(static)
      for (int64_t q = 0; q < np; q++) {
        if (s->state[q] != NORMAL_PARTICLE) continue;
        int rg = 0;
        int n = row_candidates(s, q, lb, &rg);
        rog[q] = rg;
        cnt[q] = n;
        if (n) {
          rows[q] = malloc(sizeof(int) * n);
          memcpy(rows[q], lb, sizeof(int) * n);
        }
      }
      free(lb);
    }
    for (int64_t q = 0; q < np; q++) {
      if (s->neighbor_index[q] == 0) s->neighbor_index[q] = (q == 0) ? 1 : s->neighbor_index[q - 1];
      if (s->state[q] != NORMAL_PARTICLE) { free(rows[q]); continue; }
      if (rog[q]) {
        do_exit = 1;
        s->st.rogue_deleted++;
        if (s->p.remove_rogue) delete_particle(s, q);
      }
      if (!do_exit)
        for (int t = 0; t < cnt[q]; t++) add_pair(s, q, rows[q][t]);
      free(rows[q]);
    }
    free(cnt); free(rog); free(rows); free(buf);
    return (do_exit && !s->p.remove_rogue) ? MFIXDEM_ERR_ROGUE : 0;
  }
#endif
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->neighbor_index[q] == 0) s->neighbor_index[q] = (q == 0) ? 1 : s->neighbor_index[q - 1];
    if (s->state[q] != NORMAL_PARTICLE) continue;
    int rg = 0;
    int n = row_candidates(s, q, buf, &rg);
    if (rg) {
      do_exit = 1;
      s->st.rogue_deleted++;
      if (s->p.remove_rogue) delete_particle(s, q);
    }
    if (do_exit) continue;
    for (int t = 0; t < n; t++) add_pair(s, q, buf[t]);
  }
  free(buf);
  return (do_exit && !s->p.remove_rogue) ? MFIXDEM_ERR_ROGUE : 0;
}

/* NEIGHBOUR, neighbour.f:21-126 (DES_NEIGHBOR_SEARCH = 4). */
int orc_neighbour(orc_sys* s) {
  int64_t np = s->max_pip;
  memcpy(s->neighbor_index_old, s->neighbor_index, sizeof(int) * np);
  memcpy(s->neighbors_old, s->neighbors, sizeof(int) * s->cap);
  memcpy(s->pft_old, s->pft, sizeof(double) * 3 * s->cap);
  memset(s->neighbor_index, 0, sizeof(int) * np);
  /* neighbour.f:60-66: NSQUARE for DES_NEIGHBOR_SEARCH = 1, else the grid */
  int rc = (s->p.des_neighbor_search == MFIXDEM_NSEARCH_NSQUARE) ? nsquare_build(s)
                                                                 : desgrid_neigh_build(s);
  /* the arrays may have grown; old copies were sized before */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t ll = 0; ll < np; ll++) {
    int cs = (ll > 0) ? s->neighbor_index[ll - 1] : 1;
    int ce = s->neighbor_index[ll];
    int cso = (ll > 0) ? s->neighbor_index_old[ll - 1] : 1;
    int ceo = s->neighbor_index_old[ll];
    for (int cc = cs; cc < ce; cc++) {
      int found = 0;
      for (int co = cso; co < ceo; co++) {
        if (s->neighbors[cc - 1] == s->neighbors_old[co - 1]) {
          for (int d = 0; d < 3; d++) s->pft[3 * (cc - 1) + d] = s->pft_old[3 * (co - 1) + d];
          found = 1;
          break;
        }
      }
      if (!found)
        for (int d = 0; d < 3; d++) s->pft[3 * (cc - 1) + d] = 0.0;
    }
  }
  s->st.rebuilds++;
  return rc;
}

int64_t orc_nsquare_pairs(const orc_sys* s, int32_t* pairs, int64_t capacity) {
  /* NSQUARE criterion (nsquare.f:59-60,131-133) without periodic images */
  int64_t L = s->ld, n = 0;
  double frlm = s->p.factor_rlm;
  for (int64_t i = 0; i < s->max_pip; i++) {
    if (s->state[i] != NORMAL_PARTICLE) continue;
    for (int64_t j = i + 1; j < s->max_pip; j++) {
      if (s->state[j] == NONEXISTENT) continue;
      double sr = frlm * (s->radius[i] + s->radius[j]);
      double d0 = s->pos[i] - s->pos[j], d1 = s->pos[L + i] - s->pos[L + j],
             d2 = s->pos[2 * L + i] - s->pos[2 * L + j];
      double ds = d0 * d0 + d1 * d1 + d2 * d2;
      if (ds > sr * sr) continue;
      if (n < capacity) {
        pairs[2 * n] = (int32_t)(i + 1);
        pairs[2 * n + 1] = (int32_t)(j + 1);
      }
      n++;
    }
  }
  return n;
}

/* ----------------------------------------------------- wall contact */
/* ClosestPtPointTriangle, stl_functions_des_mod.f:25-160 */
static void closest_pt_tri(const orc_sys* s, int nf, const double pp[3], double cp[3],
                           int* pe, int* edge) {
  const double(*v)[3] = s->vert[nf];
  double ab[3], ac[3], ap[3], bp[3], cpv[3];
  *pe = 0;
  *edge = 0;
  for (int i = 0; i < 3; i++) {
    ab[i] = v[1][i] - v[0][i];
    ac[i] = v[2][i] - v[0][i];
    ap[i] = pp[i] - v[0][i];
  }
  double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 < 0.0 && d2 < 0.0) {
    for (int i = 0; i < 3; i++) cp[i] = v[0][i];
    *pe = 1; *edge = 1;
    if (s->ignore_edge[nf][1]) cp[0] = UNDEF;
    return;
  }
  for (int i = 0; i < 3; i++) bp[i] = pp[i] - v[1][i];
  double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 > 0.0 && d4 < d3) {
    for (int i = 0; i < 3; i++) cp[i] = v[1][i];
    *pe = 1; *edge = 1;
    if (s->ignore_edge[nf][1]) cp[0] = UNDEF;
    return;
  }
  double vc = d1 * d4 - d3 * d2;
  if (vc < 0.0 && d1 > 0.0 && d3 < 0.0) {
    double vv = d1 / (d1 - d3);
    for (int i = 0; i < 3; i++) cp[i] = v[0][i] + vv * ab[i];
    *pe = 1; *edge = 1;
    if (s->ignore_edge[nf][1]) cp[0] = UNDEF;
    return;
  }
  for (int i = 0; i < 3; i++) cpv[i] = pp[i] - v[2][i];
  double d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
  if (d6 > 0.0 && d5 < d6) {
    for (int i = 0; i < 3; i++) cp[i] = v[2][i];
    *pe = 1; *edge = 2;
    if (s->ignore_edge[nf][2]) cp[0] = UNDEF;
    return;
  }
  double vb = d5 * d2 - d1 * d6;
  if (vb < 0.0 && d2 > 0.0 && d6 < 0.0) {
    double w = d2 / (d2 - d6);
    for (int i = 0; i < 3; i++) cp[i] = v[0][i] + w * ac[i];
    *pe = 1; *edge = 2;
    if (s->ignore_edge[nf][2]) cp[0] = UNDEF;
    return;
  }
  double va = d3 * d6 - d5 * d4;
  if (va < 0.0 && (d4 - d3) > 0.0 && (d5 - d6) > 0.0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; i++) cp[i] = v[1][i] + w * (v[2][i] - v[1][i]);
    *pe = 1; *edge = 3;
    if (s->ignore_edge[nf][3]) cp[0] = UNDEF;
    return;
  }
  double denom = 1.0 / (va + vb + vc);
  double vv = vb * denom, w = vc * denom;
  for (int i = 0; i < 3; i++) cp[i] = v[0][i] + ab[i] * vv + ac[i] * w;
}

static void remove_collision(orc_sys* s, int64_t q, int nf) {
  int64_t L = s->ld;
  for (int cc = 0; cc < COLLISION_ARRAY_MAX; cc++)
    if (s->wall_fid[cc * L + q] == nf) {
      s->wall_fid[cc * L + q] = -1;
      return;
    }
}

static void get_collision(orc_sys* s, int64_t q, int nf, double out[3]) {
  int64_t L = s->ld;
  int free_index = -1;
  for (int cc = 0; cc < COLLISION_ARRAY_MAX; cc++) {
    if (nf == s->wall_fid[cc * L + q]) {
      for (int d = 0; d < 3; d++) out[d] = s->wall_pft[(cc * 3 + d) * L + q];
      return;
    } else if (s->wall_fid[cc * L + q] == -1) {
      free_index = cc;
    }
  }
  if (free_index == -1) {
    int cell = s->dg_pijk[q] - 1;
    for (int cc = 0; cc < COLLISION_ARRAY_MAX && free_index == -1; cc++) {
      int listed = 0;
      for (int lc = 0; lc < s->fc_count[cell]; lc++)
        if (s->wall_fid[cc * L + q] == s->fc_id[s->fc_off[cell] + lc]) listed = 1;
      if (!listed) free_index = cc;
    }
  }
  if (free_index == -1) {
    /* the reference doubles COLLISION_ARRAY_MAX here (:691-696); this
       restatement keeps 8 slots and reports the run */
    s->wall_slot_overflow = 1;
    for (int d = 0; d < 3; d++) out[d] = 0.0;
    return;
  }
  s->wall_fid[free_index * L + q] = nf;
  for (int d = 0; d < 3; d++) {
    s->wall_pft[(free_index * 3 + d) * L + q] = 0.0;
    out[d] = 0.0;
  }
}

static void update_collision(orc_sys* s, int64_t q, int nf, const double pft[3]) {
  int64_t L = s->ld;
  for (int cc = 0; cc < COLLISION_ARRAY_MAX; cc++)
    if (nf == s->wall_fid[cc * L + q]) {
      for (int d = 0; d < 3; d++) s->wall_pft[(cc * 3 + d) * L + q] = pft[d];
      return;
    }
}

/* CALC_DEM_FORCE_WITH_WALL_STL, calc_collision_wall_mod.f:54-794 (sphere,
   no cohesion, no rolling friction, static walls) */
static void wall_force_particle(orc_sys* s, int64_t q) {
  int64_t L = s->ld;
  if (s->state[q] != NORMAL_PARTICLE) return;
  int cell = s->dg_pijk[q] - 1;
  if (s->fc_count[cell] < 1) {
    for (int cc = 0; cc < COLLISION_ARRAY_MAX; cc++) {
      s->wall_fid[cc * L + q] = -1;
      for (int d = 0; d < 3; d++) s->wall_pft[(cc * 3 + d) * L + q] = 0.0;
    }
    return;
  }
  double r = s->radius[q];
  double radsq = r * r;
  double pos[3] = {s->pos[q], s->pos[L + q], s->pos[2 * L + q]};
  double pmax[3], pmin[3];
  for (int c = 0; c < 3; c++) {
    pmax[c] = pos[c] + r;
    pmin[c] = pos[c] - r;
  }
  int ph = s->phase[q];
  for (int cnt = 0; cnt < s->fc_count[cell]; cnt++) {
    int lq = s->fc_off[cell] + cnt;
    int axis = s->fc_dir[lq];
    int nf = s->fc_id[lq];
    if (s->fc_min[lq] > pmax[axis]) { remove_collision(s, q, nf); continue; }
    if (s->fc_max[lq] < pmin[axis]) { remove_collision(s, q, nf); continue; }
    double vv[3] = {s->vert[nf][0][0] - pos[0], s->vert[nf][0][1] - pos[1], s->vert[nf][0][2] - pos[2]};
    double line_t = dot3(vv, s->norm[nf]);
    if (line_t <= -1.0001 * r) { remove_collision(s, q, nf); continue; }
    double cp[3];
    int pe, edge;
    closest_pt_tri(s, nf, pos, cp, &pe, &edge);
    double dist[3] = {cp[0] - pos[0], cp[1] - pos[1], cp[2] - pos[2]};
    double distsq = dot3(dist, dist);
    if (distsq >= radsq - SMALL_NUMBER) { remove_collision(s, q, nf); continue; }
    if (pe && s->ignore_edge[nf][edge]) continue;
    double distmod = sqrt(distsq);
    double normal[3] = {dist[0], dist[1], dist[2]};
    if (distsq > 0.0) {
      double sq = sqrt(distsq);
      for (int c = 0; c < 3; c++) normal[c] = dist[c] / sq;
    }
    double overlap_n = r - distmod;
    double om[3] = {s->omega[q], s->omega[L + q], s->omega[2 * L + q]};
    double cr[3], vrot[3];
    cross3(om, normal, cr);
    for (int c = 0; c < 3; c++) vrot[c] = distmod * cr[c];
    /* CFRELVEL_WALL (:806-847): the wall's tangential and normal velocity */
    double vrel[3];
    for (int c = 0; c < 3; c++) {
      vrel[c] = s->vel[c * L + q] + vrot[c];
      vrel[c] = vrel[c] - (s->wall_vt[nf][c] + s->wall_vn[nf][c]);
    }
    double vrn = dot3(vrel, normal);
    double vrt[3];
    for (int c = 0; c < 3; c++) vrt[c] = vrel[c] - vrn * normal[c];
    double kn, kt, etan, etat;
    if (s->p.coll_model == MFIXDEM_HERTZ) {
      double sqo = 0.0;
      if (overlap_n > 0.0) sqo = sqrt(overlap_n);
      kn = s->d.hert_kwn[ph] * sqo * sqrt(r);
      kt = s->d.hert_kwt[ph] * sqo * sqrt(r);
      etan = s->d.des_etan_wall[ph] * sqrt(sqo) * sqrt(s->d.hert_kwn[ph] * sqrt(r) * s->pmass[q]);
      etat = s->d.des_etat_wall[ph] * sqrt(sqo) * sqrt(s->d.hert_kwt[ph] * sqrt(r) * s->pmass[q]);
    } else {
      kn = s->p.kn_w;
      kt = s->kt_w;
      etan = s->d.des_etan_wall[ph] * sqrt(s->pmass[q]);
      etat = s->d.des_etat_wall[ph] * sqrt(s->pmass[q]);
    }
    double fnorm[3], ot[3], ftan[3], hist[3];
    for (int c = 0; c < 3; c++) fnorm[c] = -(kn * overlap_n * normal[c] + etan * vrn * normal[c]);
    get_collision(s, q, nf, hist);
    for (int c = 0; c < 3; c++) ot[c] = s->d.dtsolid * vrt[c] + hist[c];
    double mag = sqrt(dot3(ot, ot));
    if (mag > 0.0) {
      for (int c = 0; c < 3; c++) ftan[c] = -kt * ot[c] - etat * vrt[c];
      double ftmd = sqrt(dot3(ftan, ftan));
      double fnmd = s->p.mew_w * sqrt(dot3(fnorm, fnorm));
      if (ftmd > fnmd) {
        double tg[3];
        for (int c = 0; c < 3; c++) tg[c] = ot[c] / mag;
        for (int c = 0; c < 3; c++) ftan[c] = -fnmd * tg[c];
        for (int c = 0; c < 3; c++) ot[c] = (fnmd / kt) * tg[c];
      }
    } else {
      ftan[0] = ftan[1] = ftan[2] = 0.0;
    }
    update_collision(s, q, nf, ot);
    for (int c = 0; c < 3; c++) s->fc[c * L + q] = s->fc[c * L + q] + fnorm[c] + ftan[c];
    double tw[3];
    cross3(normal, ftan, tw);
    for (int c = 0; c < 3; c++) s->tow[c * L + q] = s->tow[c * L + q] + distmod * tw[c];
  }
}

/* ----------------------------------------------------- CALC_FORCE_DEM */
#ifdef _OPENMP
#define OMP_ATOMIC _Pragma("omp atomic")
#else
#define OMP_ATOMIC
#endif

/* calc_force_dem.f:18-544 (sphere path) with CFRELVEL (cfrelvel.f:21-81) */
int orc_calc_force_dem(orc_sys* s) {
  int64_t L = s->ld;
  int err = 0;
  int64_t contacts = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t q = 0; q < s->max_pip; q++) wall_force_particle(s, q);
  double dt = s->d.dtsolid;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : contacts) reduction(| : err)
#endif
  for (int64_t ll = 0; ll < s->max_pip; ll++) {
    if (s->state[ll] == NONEXISTENT) continue;
    double pos[3] = {s->pos[ll], s->pos[L + ll], s->pos[2 * L + ll]};
    double rad = s->radius[ll];
    int cs = (ll > 0) ? s->neighbor_index[ll - 1] : 1;
    int ce = s->neighbor_index[ll];
    for (int cc = cs; cc < ce; cc++) {
      int64_t i = s->neighbors[cc - 1] - 1;
      if (s->state[i] == NONEXISTENT) continue;
      double r_lm = rad + s->radius[i];
      double dist[3] = {s->pos[i] - pos[0], s->pos[L + i] - pos[1], s->pos[2 * L + i] - pos[2]};
      double dist_mag = dot3(dist, dist);
      double* pft = &s->pft[3 * (cc - 1)];
      double rs = r_lm - SMALL_NUMBER;
      if (dist_mag > rs * rs) {
        pft[0] = pft[1] = pft[2] = 0.0;
        continue;
      }
      if (dist_mag == 0.0) {
        err = 1;
        continue;
      }
      contacts++;
      dist_mag = sqrt(dist_mag);
      double normal[3];
      for (int c = 0; c < 3; c++) normal[c] = dist[c] / dist_mag;
      double overlap_n = r_lm - dist_mag;
      /* CFRELVEL */
      double vrel[3];
      for (int c = 0; c < 3; c++) vrel[c] = s->vel[c * L + ll] - s->vel[c * L + i];
      double rL = s->radius[ll], rI = s->radius[i];
      double dcl = (dist_mag * dist_mag + rL * rL - rI * rI) / (2.0 * dist_mag);
      double dci = dist_mag - dcl;
      double osum[3], vrot[3];
      for (int c = 0; c < 3; c++) osum[c] = s->omega[c * L + ll] * dcl + s->omega[c * L + i] * dci;
      cross3(osum, normal, vrot);
      for (int c = 0; c < 3; c++) vrel[c] = vrel[c] + vrot[c];
      double vrn = dot3(vrel, normal);
      double vslip[3];
      for (int c = 0; c < 3; c++) vslip[c] = vrel[c] - vrn * normal[c];
      int phl = s->phase[ll], phi = s->phase[i];
      double rad_eff = (rad * s->radius[i]) / (rad + s->radius[i]);
      double m_eff = (s->pmass[ll] * s->pmass[i]) / (s->pmass[ll] + s->pmass[i]);
      double kn, kt, etan, etat;
      if (s->p.coll_model == MFIXDEM_HERTZ) {
        double sqo = sqrt(overlap_n);
        kn = s->d.hert_kn[phl][phi] * sqo * sqrt(rad_eff);
        kt = s->d.hert_kt[phl][phi] * sqo * sqrt(rad_eff);
        sqo = sqrt(sqo);
        etan = s->d.des_etan[phl][phi] * sqo * sqrt(s->d.hert_kn[phl][phi] * sqrt(rad_eff) * m_eff);
        etat = s->d.des_etat[phl][phi] * sqo * sqrt(s->d.hert_kt[phl][phi] * sqrt(rad_eff) * m_eff);
      } else {
        kn = s->p.kn;
        kt = s->kt;
        etan = s->d.des_etan[phl][phi] * sqrt(m_eff);
        etat = s->d.des_etat[phl][phi] * sqrt(m_eff);
      }
      double fn[3], ot[3], ft[3];
      for (int c = 0; c < 3; c++) fn[c] = -(kn * overlap_n * normal[c] + etan * vrn * normal[c]);
      for (int c = 0; c < 3; c++) ot[c] = dt * vslip[c] + pft[c];
      double mag = sqrt(dot3(ot, ot));
      if (mag > 0.0) {
        for (int c = 0; c < 3; c++) ft[c] = -kt * ot[c] - etat * vslip[c];
        double ftmd = sqrt(dot3(ft, ft));
        double fnmd = s->p.mew * sqrt(dot3(fn, fn));
        if (ftmd > fnmd) {
          double tg[3];
          for (int c = 0; c < 3; c++) tg[c] = ot[c] / mag;
          for (int c = 0; c < 3; c++) ft[c] = -fnmd * tg[c];
          for (int c = 0; c < 3; c++) ot[c] = (fnmd / kt) * tg[c];
        }
      } else {
        ft[0] = ft[1] = ft[2] = 0.0;
      }
      for (int c = 0; c < 3; c++) pft[c] = ot[c];
      double dist_cl = dist_mag / 2.0 + (s->radius[ll] * s->radius[ll] - s->radius[i] * s->radius[i]) / (2.0 * dist_mag);
      double dist_ci = dist_mag - dist_cl;
      double tf[3];
      cross3(normal, ft, tf);
      double fct[3];
      for (int c = 0; c < 3; c++) fct[c] = 0.0 + fn[c] + ft[c];
      for (int c = 0; c < 3; c++) {
        OMP_ATOMIC
        s->fc[c * L + ll] += fct[c];
      }
      for (int c = 0; c < 3; c++) {
        OMP_ATOMIC
        s->fc[c * L + i] -= fct[c];
      }
      for (int c = 0; c < 3; c++) {
        double t1 = dist_cl * tf[c];
        OMP_ATOMIC
        s->tow[c * L + ll] += t1;
      }
      for (int c = 0; c < 3; c++) {
        double t2 = dist_ci * tf[c];
        OMP_ATOMIC
        s->tow[c * L + i] += t2;
      }
    }
  }
  s->st.n_contacts = contacts;
  if (err) {
    s->st.zero_distance++;
    return MFIXDEM_ERR_ZERO_DISTANCE;
  }
  return 0;
}

/* ----------------------------------------------------------- drag */
double orc_drag_fgp(int32_t drag_type, double epg, double mu, double rog, double vrel,
                    double dpm, double pvol) {
  double ropg = rog * epg;
  double dga;
  if (drag_type == MFIXDEM_DRAG_WEN_YU) {
    /* DRAG_WEN_YU, drag_gs.f:703-758 */
    double re = (mu > 0.0) ? dpm * vrel * ropg / mu : LARGE_NUMBER;
    if (re == 0.0) {
      dga = 0.0;
    } else {
      double cd;
      if (re <= 1000.0) cd = (24.0 / (re + SMALL_NUMBER)) * (1.0 + 0.15 * pow(re, 0.687));
      else cd = 0.44;
      dga = 0.75 * cd * vrel * ropg * pow(epg, -2.65) / dpm;
    }
  } else {
    /* DRAG_GIDASPOW, drag_gs.f:536-598; C_DS_SN drag_mod.f:48-56 */
    double re = (mu > 0.0) ? dpm * vrel * ropg / mu : LARGE_NUMBER;
    if (re == 0.0) {
      dga = 0.0;
    } else if (epg <= 0.8) {
      dga = 150.0 * (1.0 - epg) * mu / (epg * (dpm * dpm)) + 1.75 * rog * vrel / dpm;
    } else {
      double cd;
      if (re <= 1000.0) cd = 24.0 * (1.0 + 0.15 * pow(re, 0.687)) / (re + SMALL_NUMBER);
      else cd = 0.44;
      dga = 0.75 * cd * vrel * ropg * pow(epg, -2.65) / dpm;
    }
  }
  /* DES_DRAG_GP Model A, UR_F_GS = 1 (drag_gp_des.f:249-257) */
  return dga * pvol;
}

/* node velocities of a fluid cell's 8 corners, DRAG_GS_DES0 :140-180 */
static void cell_nodes(const orc_sys* s, int i, int j, int k, double gst[2][2][2][3],
                       double vst[2][2][2][3]) {
  const double af = 0.25;
  int iw = i - 1, js = j - 1, kb = k - 1;
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++) {
        int I = iw + ii, J = js + jj, K = kb + kk;
        gst[kk][jj][ii][0] = s->xe[I];
        gst[kk][jj][ii][1] = s->yn[J];
        gst[kk][jj][ii][2] = s->zt[K];
        if (!vst) continue;
        int cur = fluid_ijk(s, I, J, K) - 1, ipjk = fluid_ijk(s, I + 1, J, K) - 1;
        int ijpk = fluid_ijk(s, I, J + 1, K) - 1, ipjpk = fluid_ijk(s, I + 1, J + 1, K) - 1;
        int ijpkp = fluid_ijk(s, I, J + 1, K + 1) - 1, ipjkp = fluid_ijk(s, I + 1, J, K + 1) - 1;
        int ijkp = fluid_ijk(s, I, J, K + 1) - 1;
        double vx = af * (s->u_g[cur] + s->u_g[ijpk]);
        double vy = af * (s->v_g[cur] + s->v_g[ipjk]);
        vx = vx + af * (s->u_g[ijkp] + s->u_g[ijpkp]);
        vy = vy + af * (s->v_g[ijkp] + s->v_g[ipjkp]);
        double vz = af * (s->w_g[cur] + s->w_g[ijpk] + s->w_g[ipjk] + s->w_g[ipjpk]);
        vst[kk][jj][ii][0] = vx;
        vst[kk][jj][ii][1] = vy;
        vst[kk][jj][ii][2] = vz;
      }
}

static void trilinear_weights(const double gst[2][2][2][3], const double pos[3], double w[2][2][2]) {
  /* DRAG_INTERPOLATION / DRAG_WEIGHTFACTOR, drag_gs_des0.f:507-566,
     comp_mean_fields0.f:470-530 */
  double dxx = gst[0][0][1][0] - gst[0][0][0][0];
  double dyy = gst[0][1][0][1] - gst[0][0][0][1];
  double dzz = gst[1][0][0][2] - gst[0][0][0][2];
  double z0 = pos[0] - gst[0][0][0][0], z1 = pos[1] - gst[0][0][0][1];
  z0 = z0 / dxx;
  z1 = z1 / dyy;
  double z2 = pos[2] - gst[0][0][0][2];
  z2 = z2 / dzz;
  double xv[2] = {1.0 - z0, z0}, yv[2] = {1.0 - z1, z1}, zv[2] = {1.0 - z2, z2};
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++) w[kk][jj][ii] = xv[ii] * yv[jj] * zv[kk];
}

/* CALC_DRAG_DES -> DRAG_GS_DES0 (GARG_2012), drag_gs_des0.f:22-219 */
/* DRAG_GS_DES1 (drag_gs_des1.f:50-160), Model A/B as the GARG path. */
static void drag_des1(orc_sys* s) {
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
  if (p->des_explicitly_coupled) {
    /* CALC_DRAG_DES, calc_drag_des.f:34-43: the drag formed at DES_TIME_INIT */
    for (int64_t np = 0; np < s->max_pip; np++) {
      if (s->state[np] != NORMAL_PARTICLE) continue;
      for (int d = 0; d < 3; d++) s->fc[d * L + np] = s->fc[d * L + np] + s->drag_fc[d * L + np];
    }
    return;
  }
  int nf = filter_size(s);
  for (int64_t np = 0; np < s->max_pip; np++) {
    if (s->state[np] != NORMAL_PARTICLE) continue;
    int ijk0 = s->pijk[3 * L + np] - 1;
    double lepg = 0.0, velfp[3] = {0.0, 0.0, 0.0}, lpf[3] = {0.0, 0.0, 0.0};
    if (p->des_interp_on) {
      for (int lc = 0; lc < nf; lc++) {
        int ijk = s->filter_cell[lc * L + np];
        double w = s->filter_weight[lc * L + np];
        lepg = lepg + s->ep_g[ijk] * w;
        for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + s->ugc[3 * ijk + d] * w;
        for (int d = 0; d < 3; d++) lpf[d] = lpf[d] + s->p_force[3 * ijk + d] * w;
      }
    } else {
      lepg = s->ep_g[ijk0];
      for (int d = 0; d < 3; d++) velfp[d] = s->ugc[3 * ijk0 + d];
      for (int d = 0; d < 3; d++) lpf[d] = s->p_force[3 * ijk0 + d];
    }
    if (lepg < 1.0e-6) lepg = s->ep_g[ijk0];
    if (!s->fluid[ijk0]) continue; /* DRAG_FC = 0: FC unchanged */
    double vel[3] = {s->vel[np], s->vel[L + np], s->vel[2 * L + np]};
    double vslp[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
    double vrel = sqrt(dot3(vslp, vslp));
    double fgp = orc_drag_fgp(p->drag_type, lepg, p->mu_g0, p->ro_g0, vrel,
                              2.0 * s->radius[np], s->pvol[np]);
    for (int d = 0; d < 3; d++) {
      double dfc = fgp * (velfp[d] - vel[d]);
      s->fc[d * L + np] = s->fc[d * L + np] + dfc;
    }
    if (!p->model_b)
      for (int d = 0; d < 3; d++) s->fc[d * L + np] = s->fc[d * L + np] + lpf[d] * s->pvol[np];
  }
}

int orc_calc_drag_des(orc_sys* s) {
  if (!s->p.coupled) return 0;
  if (s->p.interp_scheme != MFIXDEM_INTERP_GARG_2012) {
    drag_des1(s);
    return 0;
  }
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk] || s->pinc[ijk] == 0) continue;
    int z = ijk;
    int j = z % s->nj + 1;
    z /= s->nj;
    int i = z % s->ni + 1;
    int k = z / s->ni + 1;
    double gst[2][2][2][3], vst[2][2][2][3];
    cell_nodes(s, i, j, k, gst, vst);
    for (int t = s->pic_off[ijk]; t < s->pic_off[ijk + 1]; t++) {
      int np = s->pic_list[t];
      if (s->state[np] == NONEXISTENT) continue;
      double pos[3] = {s->pos[np], s->pos[L + np], s->pos[2 * L + np]};
      double w[2][2][2];
      trilinear_weights(gst, pos, w);
      double velfp[3] = {0.0, 0.0, 0.0};
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++)
            for (int c = 0; c < 3; c++) velfp[c] = velfp[c] + vst[kk][jj][ii][c] * w[kk][jj][ii];
      double vel[3] = {s->vel[np], s->vel[L + np], s->vel[2 * L + np]};
      double vslp[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
      double vrel = sqrt(dot3(vslp, vslp));
      double fgp = orc_drag_fgp(p->drag_type, s->ep_g[ijk], p->mu_g0, p->ro_g0, vrel,
                                2.0 * s->radius[np], s->pvol[np]);
      for (int c = 0; c < 3; c++) {
        double dfc = fgp * (velfp[c] - vel[c]);
        s->fc[c * L + np] = s->fc[c * L + np] + dfc;
      }
      if (!p->model_b)
        for (int c = 0; c < 3; c++)
          s->fc[c * L + np] = s->fc[c * L + np] + s->p_force[3 * ijk + c] * s->pvol[np];
    }
  }
  return 0;
}

/* ------------------------------------------------------- integrator */
/* CFUPDATEOLD, cfupdateold.f:15-46 */
static void cfupdateold(orc_sys* s) {
  int64_t L = s->ld;
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] == NONEXISTENT) continue;
    for (int c = 0; c < 3; c++) {
      s->pos_old[c * L + q] = s->pos[c * L + q];
      s->vel_old[c * L + q] = s->vel[c * L + q];
      s->omega_old[c * L + q] = s->omega[c * L + q];
    }
  }
}

/* CFNEWVALUES, cfnewvalues.f:641-864 (EULER :700-738, ADAMS_BASHFORTH
   :673-682 and :742-811) */
int orc_cfnewvalues(orc_sys* s) {
  int64_t L = s->ld;
  double dt = s->d.dtsolid;
  const double* g = s->p.grav;
  if (s->p.integrator == MFIXDEM_ADAMS_BASHFORTH) {
    if (s->first_pass) {
      for (int64_t q = 0; q < s->max_pip; q++) {
        if (s->state[q] == NONEXISTENT) continue;
        for (int c = 0; c < 3; c++) {
          s->acc_old[c * L + q] = s->fc[c * L + q] / s->pmass[q] + g[c];
          s->rot_acc_old[c * L + q] = s->tow[c * L + q];
        }
      }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t q = 0; q < s->max_pip; q++) {
      for (int c = 0; c < 3; c++) {
        int64_t k = c * L + q;
        if (s->state[q] == NORMAL_PARTICLE) {
          s->fc[k] = s->fc[k] / s->pmass[q] + g[c];
          s->vel[k] = s->vel_old[k] + 0.5 * (3.0 * s->fc[k] - s->acc_old[k]) * dt;
        }
        s->acc_old[k] = s->fc[k];
        if (s->state[q] < 101) s->pos[k] = s->pos_old[k] + 0.5 * (s->vel_old[k] + s->vel[k]) * dt;
        s->fc[k] = 0.0;
      }
      for (int c = 0; c < 3; c++) {
        int64_t k = c * L + q;
        if (s->state[q] == NORMAL_PARTICLE) {
          s->omega[k] = s->omega_old[k] + 0.5 * (3.0 * s->tow[k] * s->omoi[q] - s->rot_acc_old[k]) * dt;
          s->rot_acc_old[k] = s->tow[k] * s->omoi[q];
        }
        s->tow[k] = 0.0;
      }
    }
  } else {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t q = 0; q < s->max_pip; q++) {
      for (int c = 0; c < 3; c++) {
        int64_t k = c * L + q;
        if (s->state[q] == NORMAL_PARTICLE) s->vel[k] = s->vel[k] + dt * (s->fc[k] / s->pmass[q] + g[c]);
        if (s->state[q] < 101) s->pos[k] = s->pos[k] + s->vel[k] * dt;
        s->fc[k] = 0.0;
      }
      for (int c = 0; c < 3; c++) {
        int64_t k = c * L + q;
        if (s->state[q] == NORMAL_PARTICLE) s->omega[k] = s->omega[k] + s->tow[k] * s->omoi[q] * dt;
        s->tow[k] = 0.0;
      }
    }
  }
  s->first_pass = 0;
  return 0;
}

/* ------------------------------------------------- PARTICLES_IN_CELL */
/* particles_in_cell.f:20-234 with PIC_SEARCH :347-370; the incremental
   +/-1 search lands on the same unique face interval as the scan. */
static int locate(const double* faces, int nmax2, double x) {
  for (int idx = 1; idx <= nmax2; idx++)
    if (x >= faces[idx - 1] && x < faces[idx]) return idx;
  return nmax2 + 1;
}

int orc_particles_in_cell(orc_sys* s) {
  int64_t L = s->ld;
  memset(s->pinc, 0, sizeof(int) * s->ncell);
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] == NONEXISTENT) continue;
    int i = locate(s->xe, s->ni, s->pos[q]);
    int j = locate(s->yn, s->nj, s->pos[L + q]);
    int k = locate(s->zt, s->nk, s->pos[2 * L + q]);
    if (i <= s->ni && j <= s->nj && k <= s->nk) {
      int ijk = fluid_ijk(s, i, j, k);
      s->pijk[q] = i;
      s->pijk[L + q] = j;
      s->pijk[2 * L + q] = k;
      s->pijk[3 * L + q] = ijk;
      s->pinc[ijk - 1]++;
    } else {
      delete_particle(s, q);
      s->st.out_of_domain++;
    }
  }
  s->pic_off[0] = 0;
  for (int c = 0; c < s->ncell; c++) s->pic_off[c + 1] = s->pic_off[c] + s->pinc[c];
  int* cur = malloc(sizeof(int) * s->ncell);
  memcpy(cur, s->pic_off, sizeof(int) * s->ncell);
  for (int64_t q = 0; q < s->max_pip; q++) {
    if (s->state[q] == NONEXISTENT) continue;
    int ijk = s->pijk[3 * L + q] - 1;
    s->pic_list[cur[ijk]++] = (int)q;
  }
  free(cur);
  return 0;
}

/* --------------------------------------------------- COMP_MEAN_FIELDS */
/* COMP_MEAN_FIELDS0 (GARG), comp_mean_fields0.f:6-466, then CALC_EPG_DES
   calc_epg_des.f:19-205 */
/* CALC_EPG_DES (calc_epg_des.f:100-145) */
static int calc_epg_des(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  int nph = p->nphases;
  size_t C = (size_t)s->ncell;
  double packed = 1.0;
  if (p->des_epg_clip != UNDEF) packed = 1.0 - p->des_epg_clip;
  int64_t ier = 0;
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk]) continue;
    double sum_eps = 0.0;
    for (int m = 0; m < nph; m++) sum_eps = sum_eps + s->rop_s[m * C + ijk] / p->ro_s0[m];
    s->ep_g[ijk] = 1.0 - (packed < sum_eps ? packed : sum_eps);
    if (s->ep_g[ijk] <= 0.0 || s->ep_g[ijk] > 1.0) ier++;
  }
  s->st.epg_errors = ier;
  return ier ? MFIXDEM_ERR_EPG : 0;
}

/* COMP_MEAN_FIELDS1 (comp_mean_fields1.f:7-116): filter-weighted solids
   volume per cell, ROP_S = RO_S*SOLVOLINC/VOL in fluid cells. */
static int comp_mean_fields1(orc_sys* s) {
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
  int nph = p->nphases, nf = filter_size(s);
  size_t C = (size_t)s->ncell;
  memset(s->solvolinc, 0, sizeof(double) * C * nph);
  memset(s->rop_s, 0, sizeof(double) * C * nph);
  for (int64_t np = 0; np < s->max_pip; np++) {
    if (s->state[np] != NORMAL_PARTICLE) continue;
    double vol_wt = s->pvol[np];
    int m = s->phase[np];
    if (nf == 0) {
      int ijk = s->pijk[3 * L + np] - 1;
      s->solvolinc[m * C + ijk] = s->solvolinc[m * C + ijk] + vol_wt;
    }
    for (int lc = 0; lc < nf; lc++) {
      int ijk = s->filter_cell[lc * L + np];
      double vw = vol_wt * s->filter_weight[lc * L + np];
      s->solvolinc[m * C + ijk] = s->solvolinc[m * C + ijk] + vw;
    }
  }
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk]) continue;
    for (int m = 0; m < nph; m++)
      s->rop_s[m * C + ijk] = p->ro_s0[m] * s->solvolinc[m * C + ijk] / s->vol[ijk];
  }
  return calc_epg_des(s);
}

int orc_comp_mean_fields(orc_sys* s) {
  if (s->p.interp_scheme != MFIXDEM_INTERP_GARG_2012) return comp_mean_fields1(s);
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
  int nph = p->nphases;
  size_t C = (size_t)s->ncell;
  memset(s->rop_s, 0, sizeof(double) * C * nph);
  memset(s->rops_node, 0, sizeof(double) * C * nph);
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk] || s->pinc[ijk] == 0) continue;
    int z = ijk;
    int j = z % s->nj + 1;
    z /= s->nj;
    int i = z % s->ni + 1;
    int k = z / s->ni + 1;
    double gst[2][2][2][3];
    cell_nodes(s, i, j, k, gst, NULL);
    for (int t = s->pic_off[ijk]; t < s->pic_off[ijk + 1]; t++) {
      int np = s->pic_list[t];
      double pos[3] = {s->pos[np], s->pos[L + np], s->pos[2 * L + np]};
      double w[2][2][2];
      trilinear_weights(gst, pos, w);
      int m = s->phase[np];
      double temp2 = s->pmass[np];
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++) {
            int cur = fluid_ijk(s, i - 1 + ii, j - 1 + jj, k - 1 + kk) - 1;
            double temp1 = w[kk][jj][ii] * temp2;
            s->rops_node[m * C + cur] = s->rops_node[m * C + cur] + temp1;
          }
    }
  }
  for (int k = 1; k <= p->kmax + 1; k++)
    for (int j = 1; j <= p->jmax + 1; j++)
      for (int i = 1; i <= p->imax + 1; i++) {
        int ijk = fluid_ijk(s, i, j, k) - 1;
        if (s->vol_surr[ijk] == 0.0) continue;
        for (int m = 0; m < nph; m++) {
          double dens = s->rops_node[m * C + ijk] / s->vol_surr[ijk];
          for (int kk = k; kk <= k + 1; kk++)
            for (int jj = j; jj <= j + 1; jj++)
              for (int ii = i; ii <= i + 1; ii++) {
                int ijk2 = fluid_ijk(s, ii, jj, kk) - 1;
                s->rop_s[m * C + ijk2] = s->rop_s[m * C + ijk2] + dens * s->vol[ijk2];
              }
        }
      }
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk]) continue;
    for (int m = 0; m < nph; m++)
      if (s->rop_s[m * C + ijk] > 0.0) s->rop_s[m * C + ijk] = s->rop_s[m * C + ijk] / s->vol[ijk];
  }
  return calc_epg_des(s);
}

/* DES_VOL_NODE, compute_volume_of_nodes (cfassign.f:100-230): the volume
   around each grid node (i, j, k) in [1, imax+1] x [1, jmax+1] x [1, kmax+1]
   from the fluid cells of its 2x2x2 block, in the reference's sum order
   (the upper K layer is added as ijkp, ipjpkp, ijpkp, ipjkp), each cell's
   volume dx(I)*dy(J)*dz(K) (:154-157). */
static void des_vol_node(const orc_sys* s, double* vn) {
  const mfixdem_params* p = &s->p;
  const double af = 0.125;
  memset(vn, 0, sizeof(double) * (size_t)s->ncell);
  for (int k = 1; k <= p->kmax + 1; k++)
    for (int j = 1; j <= p->jmax + 1; j++)
      for (int i = 1; i <= p->imax + 1; i++) {
#define VOLF(I, J, K) \
  (fluid_at(s, (I), (J), (K)) ? s->cdx[(I)-1] * s->cdy[(J)-1] * s->cdz[(K)-1] : 0.0)
        double a = af * (VOLF(i, j, k) + VOLF(i + 1, j, k) + VOLF(i, j + 1, k) +
                         VOLF(i + 1, j + 1, k));
        a = a + af * (VOLF(i, j, k + 1) + VOLF(i + 1, j + 1, k + 1) + VOLF(i, j + 1, k + 1) +
                      VOLF(i + 1, j, k + 1));
#undef VOLF
        vn[fluid_ijk(s, i, j, k) - 1] = a;
      }
}

/* DRAG_GS_GAS0 (drag_gs_des0.f:235-497): the gas side of the GARG_2012
   drag.  For every particle of every fluid cell (cell order, then PIC
   order) the drag coefficient F_GP of DES_DRAG_GP with the interpolated
   gas velocity, and its trilinear node weights; each node gets
   F_GP*w/DES_VOL_NODE in DRAG_AM and F_GP*v_p*w/DES_VOL_NODE in DRAG_BM
   (WTP = 1 for DEM).  F_GDS of a fluid cell is 1/8 of the DRAG_AM of its 8
   corner nodes (:455-494), 0 elsewhere.  Outputs: drag_am[ncell] and
   f_gds[ncell] in IJK order, drag_bm[3*ncell] as planes (d*ncell + ijk). */
static int drag_gs_gas0(orc_sys* s, double* f_gds, double* drag_am, double* drag_bm) {
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
  size_t C = (size_t)s->ncell;
  double* vn = malloc(sizeof(double) * C);
  if (!vn) return MFIXDEM_ERR_OOM;
  des_vol_node(s, vn);
  memset(drag_am, 0, sizeof(double) * C);
  memset(drag_bm, 0, sizeof(double) * 3 * C);
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk] || s->pinc[ijk] == 0) continue;
    int z = ijk;
    int j = z % s->nj + 1;
    z /= s->nj;
    int i = z % s->ni + 1;
    int k = z / s->ni + 1;
    double gst[2][2][2][3], vst[2][2][2][3];
    cell_nodes(s, i, j, k, gst, vst);
    for (int t = s->pic_off[ijk]; t < s->pic_off[ijk + 1]; t++) {
      int np = s->pic_list[t];
      if (s->state[np] == NONEXISTENT) continue;
      double pos[3] = {s->pos[np], s->pos[L + np], s->pos[2 * L + np]};
      double w[2][2][2];
      trilinear_weights(gst, pos, w);
      double velfp[3] = {0.0, 0.0, 0.0};
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++)
            for (int c = 0; c < 3; c++) velfp[c] = velfp[c] + vst[kk][jj][ii][c] * w[kk][jj][ii];
      double vel[3] = {s->vel[np], s->vel[L + np], s->vel[2 * L + np]};
      double vslp[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
      double vrel = sqrt(dot3(vslp, vslp));
      double fgp = orc_drag_fgp(p->drag_type, s->ep_g[ijk], p->mu_g0, p->ro_g0, vrel,
                                2.0 * s->radius[np], s->pvol[np]);
      for (int kk = 0; kk < 2; kk++)
        for (int jj = 0; jj < 2; jj++)
          for (int ii = 0; ii < 2; ii++) {
            int cur = fluid_ijk(s, i - 1 + ii, j - 1 + jj, k - 1 + kk) - 1;
            double ovol = 1.0 / vn[cur];
            drag_am[cur] = drag_am[cur] + fgp * w[kk][jj][ii] * ovol;
            for (int d = 0; d < 3; d++)
              drag_bm[d * C + cur] = drag_bm[d * C + cur] + fgp * vel[d] * w[kk][jj][ii] * ovol;
          }
    }
  }
  for (int ijk = 0; ijk < s->ncell; ijk++) {
    if (!s->fluid[ijk]) {
      f_gds[ijk] = 0.0;
      continue;
    }
    int z = ijk;
    int j = z % s->nj + 1;
    z /= s->nj;
    int i = z % s->ni + 1;
    int k = z / s->ni + 1;
#define AM(I, J, K) drag_am[fluid_ijk(s, (I), (J), (K)) - 1]
    double f = 0.125 * (AM(i, j, k) + AM(i, j - 1, k) + AM(i - 1, j - 1, k) + AM(i - 1, j, k));
    f = f + 0.125 * (AM(i, j, k - 1) + AM(i, j - 1, k - 1) + AM(i - 1, j - 1, k - 1) +
                     AM(i - 1, j, k - 1));
#undef AM
    f_gds[ijk] = f;
  }
  free(vn);
  return 0;
}

/* DRAG_GS_GAS1 (drag_gs_des1.f:179-301), the gas side of the drag for the
   filter schemes: every normal particle in a fluid cell (ascending id)
   interpolates EP_G and the cell-centre gas velocity with its filter
   weights as DRAG_GS_DES1 does, takes F_GP from DES_DRAG_GP, and adds
   F_GP*w/VOL to F_GDS and F_GP*v_p*w/VOL to DRAG_BM of each filter cell
   (its own cell with weight 1 when DES_INTERP_ON is off).  There is no
   DRAG_AM in this scheme; DES_COLLECT_gDATA is a halo exchange with no
   partner on one process. */
static int drag_gs_gas1(orc_sys* s, double* f_gds, double* drag_bm) {
  int64_t L = s->ld;
  const mfixdem_params* p = &s->p;
  size_t C = (size_t)s->ncell;
  memset(f_gds, 0, sizeof(double) * C);
  memset(drag_bm, 0, sizeof(double) * 3 * C);
  if (p->des_interp_on) orc_calc_interp_weights(s);
  int nf = filter_size(s);
  for (int64_t np = 0; np < s->max_pip; np++) {
    if (s->state[np] != NORMAL_PARTICLE) continue;
    int ijk0 = s->pijk[3 * L + np] - 1;
    if (!s->fluid[ijk0]) continue;
    double lepg = 0.0, velfp[3] = {0.0, 0.0, 0.0};
    if (p->des_interp_on) {
      for (int lc = 0; lc < nf; lc++) {
        int ijk = s->filter_cell[lc * L + np];
        double w = s->filter_weight[lc * L + np];
        lepg = lepg + s->ep_g[ijk] * w;
        for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + s->ugc[3 * ijk + d] * w;
      }
    } else {
      lepg = s->ep_g[ijk0];
      for (int d = 0; d < 3; d++) velfp[d] = s->ugc[3 * ijk0 + d];
    }
    if (lepg < 1.0e-6) lepg = s->ep_g[ijk0];
    double vel[3] = {s->vel[np], s->vel[L + np], s->vel[2 * L + np]};
    double vslp[3] = {velfp[0] - vel[0], velfp[1] - vel[1], velfp[2] - vel[2]};
    double vrel = sqrt(dot3(vslp, vslp));
    double lforce = orc_drag_fgp(p->drag_type, lepg, p->mu_g0, p->ro_g0, vrel,
                                 2.0 * s->radius[np], s->pvol[np]);
    s->f_gp[np] = lforce; /* F_GP(NP), set by DES_DRAG_GP */
    double lbm[3] = {lforce * vel[0], lforce * vel[1], lforce * vel[2]};
    int nl = p->des_interp_on ? nf : 1;
    for (int lc = 0; lc < nl; lc++) {
      int ijk = p->des_interp_on ? s->filter_cell[lc * L + np] : ijk0;
      double w = p->des_interp_on ? s->filter_weight[lc * L + np] / s->vol[ijk]
                                  : 1.0 / s->vol[ijk];
      for (int d = 0; d < 3; d++) drag_bm[d * C + ijk] = drag_bm[d * C + ijk] + lbm[d] * w;
      f_gds[ijk] = f_gds[ijk] + lforce * w;
    }
  }
  return 0;
}

/* CALC_DES_2FLUID's drag call (calc_des_2fluid.f:70-73): DRAG_GS_GAS0 for
   GARG_2012, DRAG_GS_GAS1 for the other schemes (drag_am all 0 there). */
int orc_drag_gs_gas(orc_sys* s, double* f_gds, double* drag_am, double* drag_bm) {
  if (!s->p.coupled) return MFIXDEM_ERR_UNSUPPORTED;
  if (s->p.interp_scheme == MFIXDEM_INTERP_GARG_2012)
    return drag_gs_gas0(s, f_gds, drag_am, drag_bm);
  memset(drag_am, 0, sizeof(double) * (size_t)s->ncell);
  return drag_gs_gas1(s, f_gds, drag_bm);
}

/* DES_TIME_INIT's coupled part with DES_EXPLICITLY_COUPLED
   (des_time_march.f:155-161, 271): DRAG_GS_DES1's explicit branch
   (drag_gs_des1.f:96-137) DRAG_FC = F_GP*(VELFP - DES_VEL_NEW) with the
   F_GP of the last DRAG_GS_GAS1 and VELFP from the current gas field (the
   filter weights of the last CALC_INTERP_WEIGHTS; the particles have not
   moved since), then CALC_PG_GRAD (calc_pg_grad.f:74-105) adds
   P_FORCE*PVOL, P_FORCE interpolated with the same weights. */
int orc_des_time_init(orc_sys* s) {
  if (!s->inited) return MFIXDEM_ERR_NOT_READY;
  const mfixdem_params* p = &s->p;
  if (!p->coupled || !p->des_explicitly_coupled) return 0;
  int64_t L = s->ld;
  int nf = filter_size(s);
  for (int64_t np = 0; np < s->max_pip; np++) {
    if (s->state[np] != NORMAL_PARTICLE) continue;
    int ijk0 = s->pijk[3 * L + np] - 1;
    if (!s->fluid[ijk0]) {
      for (int d = 0; d < 3; d++) s->drag_fc[d * L + np] = 0.0;
      s->f_gp[np] = 0.0;
      continue;
    }
    double velfp[3] = {0.0, 0.0, 0.0}, lpf[3] = {0.0, 0.0, 0.0};
    if (p->des_interp_on) {
      for (int lc = 0; lc < nf; lc++) {
        int ijk = s->filter_cell[lc * L + np];
        double w = s->filter_weight[lc * L + np];
        for (int d = 0; d < 3; d++) velfp[d] = velfp[d] + s->ugc[3 * ijk + d] * w;
      }
      for (int lc = 0; lc < nf; lc++) {
        int ijk = s->filter_cell[lc * L + np];
        double w = s->filter_weight[lc * L + np];
        for (int d = 0; d < 3; d++) lpf[d] = lpf[d] + s->p_force[3 * ijk + d] * w;
      }
    } else {
      for (int d = 0; d < 3; d++) velfp[d] = s->ugc[3 * ijk0 + d];
      for (int d = 0; d < 3; d++) lpf[d] = s->p_force[3 * ijk0 + d];
    }
    double vel[3] = {s->vel[np], s->vel[L + np], s->vel[2 * L + np]};
    for (int d = 0; d < 3; d++) s->drag_fc[d * L + np] = s->f_gp[np] * (velfp[d] - vel[d]);
    for (int d = 0; d < 3; d++)
      s->drag_fc[d * L + np] = s->drag_fc[d * L + np] + lpf[d] * s->pvol[np];
  }
  return 0;
}

/* PARTICLE_STATE of slots 0..n-1 before set-up: 0 makes the slot a hole
   (NONEXISTENT, discretelement_mod.f:68-78; its data cleared as
   DELETE_PARTICLE clears it), 1 leaves it NORMAL; entering / exiting
   particles (2, 3) and ghosts are not supported. */
int orc_set_particle_state(orc_sys* s, const int8_t* state, int64_t n) {
  if (!s || !state || n < 0 || n > s->max_pip || s->inited) return MFIXDEM_ERR_ARG;
  for (int64_t q = 0; q < n; q++)
    if (state[q] != NONEXISTENT && state[q] != NORMAL_PARTICLE) return MFIXDEM_ERR_UNSUPPORTED;
  for (int64_t q = 0; q < n; q++)
    if (state[q] == NONEXISTENT && s->state[q] != NONEXISTENT) delete_particle(s, q);
  return 0;
}

/* ------------------------------------------- DEM mass inlets / outlets */
/* The seeded stream standing in for RANDOM_NUMBER: splitmix64 (Steele,
   Lea and Flood, 2014), 53 bits to [0, 1). */
static double orc_random(orc_sys* s) {
  uint64_t z = (s->rng += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* CALC_CELL_INTERSECT (calc_cell.f:129-180): the interior cell 2..n+1
   whose faces bracket loc (both ends inclusive, the lower cell first);
   c[i-1] = size of cell i. */
static int cell_intersect(double rmin, double loc, const double* c, int n) {
  double start = rmin, end = rmin;
  for (int lc = 2; lc <= n + 1; lc++) {
    end = start + c[lc - 1];
    if (start <= loc && loc <= end) return lc;
    start = end;
  }
  if (loc < rmin) return 2;
  if (loc > end) return n + 1;
  return -1;
}

static void bc_axes(const mfixdem_dem_bc* b, int* ax, int* pa, int* qa, int* low) {
  switch (b->plane) {
    case MFIXDEM_PLANE_N: case MFIXDEM_PLANE_S: *ax = 1; *pa = 0; *qa = 2; break;
    case MFIXDEM_PLANE_E: case MFIXDEM_PLANE_W: *ax = 0; *pa = 1; *qa = 2; break;
    default: *ax = 2; *pa = 0; *qa = 1; break;
  }
  *low = b->plane == MFIXDEM_PLANE_N || b->plane == MFIXDEM_PLANE_E || b->plane == MFIXDEM_PLANE_T;
}

static int ofpos_ax(const orc_sys* s, int a, double x) {
  return a == 0 ? iofpos(s, x) : (a == 1 ? jofpos(s, x) : kofpos(s, x));
}
/* IofBCPOS (desgrid_mod.f:428-432): NINT((x - start) dxinv) + start1 - 1 */
static int ofbcpos_ax(const orc_sys* s, int a, double x) {
  const double st = a == 0 ? s->dg_xstart : (a == 1 ? s->dg_ystart : s->dg_zstart);
  const double inv = a == 0 ? s->dg_dxinv : (a == 1 ? s->dg_dyinv : s->dg_dzinv);
  const int s1 = a == 0 ? s->dg_istart1 : (a == 1 ? s->dg_jstart1 : s->dg_kstart1);
  return (int)lround((x - st) * inv) + s1 - 1;
}

static void free_bcs(orc_sys* s) {
  for (int b = 0; b < s->nbc; b++) {
    free(s->bcs[b].P); free(s->bcs[b].Q); free(s->bcs[b].W); free(s->bcs[b].H);
    free(s->bcs[b].poly_layout);
  }
  free(s->bcs);
  s->bcs = NULL;
  s->nbc = 0;
}

/* SET_BC_DEM_MI (set_bc_dem_mi.f:18-377) with LAYOUT_MI_DEM
   (layout_mi_dem.f:88-444 and its E/W, T/B twins), SET_DEM_BCMI_IJK
   (set_bc_dem_mi.f:391-553) and SET_BC_DEM_MO (set_bc_dem_mo.f:32-177) for
   boundaries on faces of the box (no STL facets across an inlet). */
static int setup_bcs(orc_sys* s) {
  const mfixdem_params* p = &s->p;
  const double dt = s->d.dtsolid;
  const double* csz[3] = {s->cdx, s->cdy, s->cdz};
  const double* faces[3] = {s->xe, s->yn, s->zt};
  const int nax[3] = {p->imax, p->jmax, p->kmax};
  const double amin[3] = {p->x_min, p->y_min, p->z_min};
  const int lo1[3] = {s->dg_istart2, s->dg_jstart2, s->dg_kstart2};
  const int hi1[3] = {s->dg_iend2, s->dg_jend2, s->dg_kend2};
  const int dmin1[3] = {s->dg_istart1, s->dg_jstart1, s->dg_kstart1};
  const int dmax1[3] = {s->dg_iend1, s->dg_jend1, s->dg_kend1};
  double min_radius = 1e300;
  for (int m = 0; m < p->nphases; m++)
    if (p->d_p0[m] > 0.0 && 0.5 * p->d_p0[m] < min_radius) min_radius = 0.5 * p->d_p0[m];
  for (int bi = 0; bi < s->nbc; bi++) {
    orc_bc* B = &s->bcs[bi];
    const mfixdem_dem_bc* b = &B->bc;
    bc_axes(b, &B->ax, &B->pa, &B->qa, &B->low);
    const int ax = B->ax, pa = B->pa, qa = B->qa;
    const double blo[3] = {b->x_w, b->y_s, b->z_b}, bhi[3] = {b->x_e, b->y_n, b->z_t};
    const double coord = blo[ax];
    B->face = B->low ? faces[ax][1] : faces[ax][nax[ax] + 1];
    for (int c = 0; c < 3; c++) {
      B->vel[0][c] = 0.0;
    }
    if (b->type == MFIXDEM_BC_MASS_INFLOW) {
      /* DEM_BCMI_IJK */
      for (int a = 0; a < 3; a++) {
        B->clo[a] = imax_i(ofbcpos_ax(s, a, blo[a]) - 1, dmin1[a]);
        B->chi[a] = imin_i(ofbcpos_ax(s, a, bhi[a]) + 1, dmax1[a]);
      }
      if (B->low) {
        B->clo[ax] = ofpos_ax(s, ax, coord + min_radius);
        B->chi[ax] = ofpos_ax(s, ax, coord + p->max_radius);
      } else {
        B->clo[ax] = ofpos_ax(s, ax, coord - p->max_radius);
        B->chi[ax] = ofpos_ax(s, ax, coord - min_radius);
      }
      /* phases at the inlet, the largest diameter */
      int plist[MP], pcnt = 0;
      double max_dia = 0.0;
      for (int m = 0; m < p->nphases; m++) {
        if (!(b->ep_s[m] > 0.0)) continue;
        plist[pcnt++] = m;
        if (p->d_p0[m] > max_dia) max_dia = p->d_p0[m];
      }
      if (pcnt == 0) return MFIXDEM_ERR_ARG;
      B->polydisperse = pcnt > 1;
      /* LAYOUT_MI_DEM: windows of the inlet rectangle */
      const double plen = bhi[pa] - blo[pa], qlen = bhi[qa] - blo[qa];
      const int wmax = (int)floor((float)(plen / max_dia));
      const int hmax = (int)floor((float)(qlen / max_dia));
      if (wmax < 1 || hmax < 1) return MFIXDEM_ERR_ARG;
      const int J = cell_intersect(amin[ax], coord, csz[ax], nax[ax]);
      const double shift_side = B->low ? -1.0 : 1.0;
      B->offset = coord + max_dia * shift_side;
      B->L = J + (int)shift_side;
      B->window = fmin(plen / wmax, qlen / hmax);
      const double window = B->window;
      double* mesh_p = malloc(sizeof(double) * wmax);
      double* mesh_q = malloc(sizeof(double) * hmax);
      int* mesh_w = malloc(sizeof(int) * wmax);
      int* mesh_h = malloc(sizeof(int) * hmax);
      double shift = 0.5 * (plen - wmax * window);
      const double p0 = blo[pa] + shift;
      for (int lc = 1; lc <= wmax; lc++) {
        mesh_p[lc - 1] = p0 + (double)(lc - 1) * window;
        mesh_w[lc - 1] = cell_intersect(amin[pa], mesh_p[lc - 1] + 0.5 * window, csz[pa], nax[pa]);
      }
      shift = 0.5 * (qlen - hmax * window);
      const double q0 = blo[qa] + shift;
      for (int lc = 1; lc <= hmax; lc++) {
        mesh_q[lc - 1] = q0 + (double)(lc - 1) * window;
        mesh_h[lc - 1] = cell_intersect(amin[qa], mesh_q[lc - 1] + 0.5 * window, csz[qa], nax[qa]);
      }
      /* FULL_MAP: windows whose corner cells next to the inlet are fluid */
      unsigned char* full = calloc((size_t)wmax * hmax, 1);
      int occupants = 0;
      for (int h = 0; h < hmax; h++)
        for (int w = 0; w < wmax; w++) {
          int ok = 1;
          const double cp[2] = {mesh_p[w], mesh_p[w] + window};
          const double cq[2] = {mesh_q[h], mesh_q[h] + window};
          for (int a = 0; a < 2 && ok; a++)
            for (int c = 0; c < 2 && ok; c++) {
              int idx[3];
              idx[ax] = J;
              idx[pa] = cell_intersect(amin[pa], cp[a], csz[pa], nax[pa]);
              idx[qa] = cell_intersect(amin[qa], cq[c], csz[qa], nax[qa]);
              if (idx[0] < 1 || idx[1] < 1 || idx[2] < 1 ||
                  !s->fluid[fluid_ijk(s, idx[0], idx[1], idx[2]) - 1])
                ok = 0;
            }
          full[h * wmax + w] = (unsigned char)ok;
          occupants += ok;
        }
      if (occupants == 0) {
        free(mesh_p); free(mesh_q); free(mesh_w); free(mesh_h); free(full);
        return MFIXDEM_ERR_ARG;
      }
      B->occupants = occupants;
      /* RAND_MAP: a random order of the windows, by rejection */
      int* rand_map = calloc((size_t)occupants + 1, sizeof(int));
      int ll = 1;
      while (rand_map[occupants] == 0) {
        const double tmp = orc_random(s);
        const int ti = (int)ceil((float)(tmp * (double)occupants));
        for (int lc = 1; lc <= ll; lc++) {
          if (ti == rand_map[lc]) break;
          if (lc == ll) {
            rand_map[lc] = ti;
            ll++;
            break;
          }
        }
      }
      B->vacancy = 1;
      B->P = malloc(sizeof(double) * (occupants + 1));
      B->Q = malloc(sizeof(double) * (occupants + 1));
      B->W = malloc(sizeof(int) * (occupants + 1));
      B->H = malloc(sizeof(int) * (occupants + 1));
      int lc = 0;
      for (int h = 0; h < hmax; h++)
        for (int w = 0; w < wmax; w++) {
          if (!full[h * wmax + w]) continue;
          const int L2 = rand_map[++lc];
          B->W[L2] = mesh_w[w];
          B->H[L2] = mesh_h[h];
          B->P[L2] = mesh_p[w];
          B->Q[L2] = mesh_q[h];
        }
      free(rand_map); free(mesh_p); free(mesh_q); free(mesh_w); free(mesh_h); free(full);
      /* seeding rate (set_bc_dem_mi.f:104-175) */
      const double area = plen * qlen;
      double vel_tmp[MP] = {0}, npm[MP] = {0}, max_vel = 0.0;
      const double* bcv[3] = {b->u_s, b->v_s, b->w_s};
      for (int k = 0; k < pcnt; k++) {
        const int m = plist[k];
        vel_tmp[m] = fabs(bcv[ax][m]);
        if (fabs(vel_tmp[m]) > max_vel) max_vel = fabs(vel_tmp[m]);
        const double d = p->d_p0[m];
        const double p_vol = (PI_D * (d * d * d)) / 6.0;
        if (b->massflow_s[m] == 0.0) npm[m] = 0.0;
        else if (b->massflow_s[m] != UNDEF) npm[m] = b->massflow_s[m] / (p->ro_s0[m] * p_vol);
        else npm[m] = vel_tmp[m] * area * b->ep_s[m] / p_vol;
      }
      double nppsec = 0.0;
      for (int m = 0; m < MP; m++) nppsec += npm[m];
      const double nppdt = nppsec * dt;
      if (nppdt == 0.0) {
        B->pi_count = 0;
        B->pi_factor = 987654321.0;   /* UNDEFINED_I */
      } else if (nppdt < 1.0) {
        B->pi_count = 1;
        B->pi_factor = (double)(float)(1.0 / nppdt);
      } else {
        B->pi_count = (int)ceil((float)nppdt);
        B->pi_factor = 1.0;
      }
      if (B->pi_count > 0) {
        const double minipv = max_dia / (dt * B->pi_factor *
                                         floor((float)occupants / (float)B->pi_count));
        if (max_vel < minipv) return MFIXDEM_ERR_ARG;   /* Error 1110 */
      }
      /* one speed, normal to the inlet, for every phase (:184-205) */
      for (int m = 0; m < MP; m++) {
        B->vel[m][0] = b->u_s[m];
        B->vel[m][1] = b->v_s[m];
        B->vel[m][2] = b->w_s[m];
      }
      for (int k = 0; k < pcnt; k++) {
        const int m = plist[k];
        if (b->massflow_s[m] == 0.0) B->vel[m][ax] = 0.0;
        else B->vel[m][ax] = B->low ? fabs(max_vel) : -fabs(max_vel);
      }
      /* DEM_BC_POLY_LAYOUT (:240-255) */
      B->poly_layout = malloc(sizeof(int) * NUMFRAC_LIMIT);
      if (B->polydisperse) {
        int bot = 1;
        for (int k = 0; k < pcnt - 1; k++) {
          const int m = plist[k];
          const double scaled = (double)NUMFRAC_LIMIT * (npm[m] / nppsec);
          const int top = (int)floor(scaled) + (bot - 1);
          for (int r = bot; r <= top; r++) B->poly_layout[r - 1] = m;
          bot = top + 1;
        }
        for (int r = bot; r <= NUMFRAC_LIMIT; r++) B->poly_layout[r - 1] = plist[pcnt - 1];
      } else {
        for (int r = 0; r < NUMFRAC_LIMIT; r++) B->poly_layout[r] = plist[0];
      }
      B->mi_time = 0.0 + B->pi_factor * dt;
    } else {
      /* DEM_BCMO_IJK */
      for (int a = 0; a < 3; a++) {
        B->clo[a] = ofbcpos_ax(s, a, blo[a]);
        B->chi[a] = ofbcpos_ax(s, a, bhi[a]) + 1;
      }
    }
    for (int a = 0; a < 3; a++) {   /* dg_is_ON_myPE_plus1layers */
      B->clo[a] = imax_i(B->clo[a], lo1[a]);
      B->chi[a] = imin_i(B->chi[a], hi1[a]);
    }
  }
  return 0;
}

/* distance of the particle centre from the boundary plane, into the domain */
static double bc_dist(const orc_sys* s, const orc_bc* B, int64_t q) {
  const double x = s->pos[B->ax * s->ld + q];
  return B->low ? x - B->face : B->face - x;
}

#define FOR_BC_PARTICLES(s, B, q)                                               \
  for (int bk = (B)->clo[2]; bk <= (B)->chi[2]; bk++)                           \
    for (int bj = (B)->clo[1]; bj <= (B)->chi[1]; bj++)                         \
      for (int bi_ = (B)->clo[0]; bi_ <= (B)->chi[0]; bi_++)                    \
        for (int lp = (s)->dg_off[dg_funijk(s, bi_, bj, bk) - 1],               \
                 le = (s)->dg_off[dg_funijk(s, bi_, bj, bk)], q = 0;            \
             lp < le && ((q = (s)->dg_list[lp]), 1); lp++)

/* SET_NEW_PARTICLE_PROPS (mass_inflow_dem.f:419-642) for a sphere. */
static void new_particle(orc_sys* s, const orc_bc* B, int m, int64_t q, const double pos[3],
                         const int ijkp[3]) {
  const int64_t L = s->ld;
  const double r = s->p.d_p0[m] * 0.5;
  s->state[q] = ENTERING_PARTICLE;
  for (int c = 0; c < 3; c++) {
    s->pos[c * L + q] = pos[c];
    s->vel[c * L + q] = B->vel[m][c];
    s->omega[c * L + q] = 0.0;
    s->pijk[c * L + q] = ijkp[c];
  }
  s->radius[q] = r;
  s->ro_sol[q] = s->p.ro_s0[m];
  s->pijk[3 * L + q] = fluid_ijk(s, ijkp[0], ijkp[1], ijkp[2]);
  s->phase[q] = m;
  const int li = imin_i(s->dg_iend2, imax_i(s->dg_istart2, iofpos(s, pos[0])));
  const int lj = imin_i(s->dg_jend2, imax_i(s->dg_jstart2, jofpos(s, pos[1])));
  const int lk = imin_i(s->dg_kend2, imax_i(s->dg_kstart2, kofpos(s, pos[2])));
  s->dg_pijk[q] = dg_funijk(s, li, lj, lk);
  s->pvol[q] = (4.0 / 3.0) * PI_D * (r * r * r);
  s->pmass[q] = s->pvol[q] * s->ro_sol[q];
  s->omoi[q] = 2.5 / (s->pmass[q] * (r * r));
  for (int k = 0; k < COLLISION_ARRAY_MAX; k++) s->wall_fid[k * L + q] = -1;
  if (s->p.des_explicitly_coupled) {
    s->f_gp[q] = 0.0;
    for (int c = 0; c < 3; c++) s->drag_fc[c * L + q] = 0.0;
  }
  s->pip++;
}

/* MASS_INFLOW_DEM (mass_inflow_dem.f:44-293) with SEED_NEW_PARTICLE
   (:306-408) */
static int mass_inflow(orc_sys* s) {
  const double dt = s->d.dtsolid;
  for (int bi = 0; bi < s->nbc; bi++) {
    orc_bc* B = &s->bcs[bi];
    if (B->bc.type != MFIXDEM_BC_MASS_INFLOW) continue;
    FOR_BC_PARTICLES(s, B, q) {
      if (s->state[q] == EXITING_PARTICLE) continue;
      if (bc_dist(s, B, q) > s->radius[q] && s->state[q] == ENTERING_PARTICLE)
        s->state[q] = NORMAL_PARTICLE;
    }
    double rnd[3];
    for (int c = 0; c < 3; c++) rnd[c] = orc_random(s);
    if (B->mi_time > s->s_time) continue;
    if (s->s_time < B->bc.mi_start_time || s->s_time > B->bc.mi_end_time) continue;
    int64_t ls = 0;
    for (int ip = 0; ip < B->pi_count; ip++) {
      const int vac = B->vacancy;
      B->vacancy = vac % B->occupants + 1;
      int m = B->poly_layout[0];
      if (B->polydisperse) {
        const int ri = (int)ceil((double)NUMFRAC_LIMIT * rnd[0]);
        m = B->poly_layout[(ri < 1 ? 1 : ri) - 1];
      }
      const double d = s->p.d_p0[m];
      const double r1 = 0.5 * d + (B->window - d) * rnd[0];
      const double r2 = 0.5 * d + (B->window - d) * rnd[1];
      double pos[3];
      int ijkp[3];
      pos[B->pa] = B->P[vac] + r1;
      pos[B->qa] = B->Q[vac] + r2;
      pos[B->ax] = B->offset;
      ijkp[B->pa] = B->W[vac];
      ijkp[B->qa] = B->H[vac];
      ijkp[B->ax] = B->L;
      while (ls < s->max_pip && s->state[ls] != NONEXISTENT) ls++;
      if (ls >= s->max_pip) return MFIXDEM_ERR_CAPACITY;   /* the reference grows MAX_PIP */
      new_particle(s, B, m, ls, pos, ijkp);
    }
    B->mi_time = B->mi_time + B->pi_factor * dt;
  }
  return 0;
}

/* MASS_OUTFLOW_DEM (mass_outflow_dem.f:13-188) */
static void mass_outflow(orc_sys* s, int* force_nsearch) {
  const int64_t L = s->ld;
  for (int bi = 0; bi < s->nbc; bi++) {
    orc_bc* B = &s->bcs[bi];
    if (B->bc.type == MFIXDEM_BC_MASS_INFLOW) continue;
    const int freeze_vel = B->bc.type != MFIXDEM_BC_MASS_OUTFLOW;
    const double sgn = B->low ? -1.0 : 1.0;
    FOR_BC_PARTICLES(s, B, q) {
      if (s->state[q] == NONEXISTENT || s->state[q] == ENTERING_PARTICLE) continue;
      const double dist = bc_dist(s, B, q);
      if (dist > s->radius[q]) {
        s->state[q] = NORMAL_PARTICLE;
      } else if (dist > 0.0) {
        if (freeze_vel) {
          if (s->vel[B->ax * L + q] * sgn > 0.0) {
            for (int c = 0; c < 3; c++)
              s->vel[c * L + q] = s->vel[c * L + q] * (c == B->ax ? 1.0 : 0.0);
            s->state[q] = EXITING_PARTICLE;
          }
        } else {
          const int m = s->phase[q];
          s->vel[q] = B->bc.u_s[m];
          s->vel[L + q] = B->bc.v_s[m];
          s->vel[2 * L + q] = B->bc.w_s[m];
          s->state[q] = EXITING_PARTICLE;
        }
      } else {
        delete_particle(s, q);
        *force_nsearch = 1;
      }
    }
  }
}

/* UPDATE_EXIT_DES_VEL (mass_outflow_dem.f:444-514), Euler */
static void update_exit_des_vel(orc_sys* s) {
  const int64_t L = s->ld;
  const double dt = s->d.dtsolid;
  const double* g = s->p.grav;
  for (int bi = 0; bi < s->nbc; bi++) {
    orc_bc* B = &s->bcs[bi];
    if (B->bc.type != MFIXDEM_BC_OUTFLOW) continue;
    FOR_BC_PARTICLES(s, B, q) {
      if (s->state[q] != EXITING_PARTICLE) continue;
      for (int c = 0; c < 3; c++) {
        const int64_t k = c * L + q;
        s->vel[k] = s->vel[k] + dt * (s->fc[k] / s->pmass[q] + g[c]);
        s->vel[k] = s->vel[k] * (c == B->ax ? 1.0 : 0.0);
      }
    }
  }
}

int orc_set_dem_bcs(orc_sys* s, int32_t n, const mfixdem_dem_bc* bcs, uint64_t seed) {
  if (!s || n < 0 || (n > 0 && !bcs) || s->inited) return MFIXDEM_ERR_ARG;
  if (n > 0 && s->p.integrator != MFIXDEM_EULER) return MFIXDEM_ERR_UNSUPPORTED;
  for (int b = 0; b < n; b++) {
    if (bcs[b].type < MFIXDEM_BC_MASS_INFLOW || bcs[b].type > MFIXDEM_BC_OUTFLOW)
      return MFIXDEM_ERR_ARG;
    if (bcs[b].plane < MFIXDEM_PLANE_W || bcs[b].plane > MFIXDEM_PLANE_T) return MFIXDEM_ERR_ARG;
  }
  free_bcs(s);
  s->nbc = n;
  s->bcs = calloc((size_t)(n > 0 ? n : 1), sizeof(orc_bc));
  for (int b = 0; b < n; b++) s->bcs[b].bc = bcs[b];
  s->rng = seed;
  return setup_bcs(s);
}

/* ------------------------------------------------------- driver */
int orc_desgrid_and_neighbour(orc_sys* s) {
  orc_desgrid_pic(s);
  return orc_neighbour(s);
}

/* MAKE_ARRAYS_DES, make_arrays_des.f:343-432 */
int orc_init(orc_sys* s) {
  int rc = 0;
  orc_desgrid_pic(s);
  if (s->p.coupled) {
    if (!s->gas_set) return MFIXDEM_ERR_NOT_READY;
    orc_particles_in_cell(s);
  }
  memset(s->neighbor_index, 0, sizeof(int) * s->max_pip);
  rc = orc_neighbour(s);
  if (rc) return rc;
  if (s->p.coupled) {
    orc_calc_interp_weights(s);
    rc = orc_comp_mean_fields(s);
  }
  s->inited = 1;
  s->st.steps = 0;
  return rc;
}

/* DES_TIME_STEP(NN), des_time_march.f:293-460 */
int orc_time_step(orc_sys* s, int32_t nn) {
  if (!s->inited) return MFIXDEM_ERR_NOT_READY;
  int64_t L = s->ld;
  int rc = orc_calc_force_dem(s);
  if (rc) return rc;
  if (s->wall_slot_overflow) return MFIXDEM_ERR_CAPACITY;
  if (s->p.coupled) orc_calc_drag_des(s);
  if (s->p.integrator == MFIXDEM_ADAMS_BASHFORTH) cfupdateold(s);
  memcpy(s->fc_last, s->fc, sizeof(double) * 3 * L);
  memcpy(s->tow_last, s->tow, sizeof(double) * 3 * L);
  if (s->nbc) update_exit_des_vel(s);
  orc_cfnewvalues(s);
  int do_nsearch = (nn == 1) || (nn % s->p.neighbor_search_n == 0);
  if (s->nbc) {
    rc = mass_inflow(s);
    if (rc) return rc;
    mass_outflow(s, &do_nsearch);
  }
  if (do_nsearch) {
    orc_desgrid_pic(s);
    rc = orc_neighbour(s);
    if (rc) return rc;
  }
  /* explicit coupling rebins and deposits once per fluid step, in
     CALC_DES_2FLUID, not here (des_time_march.f:419-425) */
  if (s->p.coupled && !s->p.des_explicitly_coupled) {
    orc_particles_in_cell(s);
    orc_calc_interp_weights(s);
    rc = orc_comp_mean_fields(s);
  }
  s->s_time = s->s_time + s->d.dtsolid;
  s->st.steps++;
  return rc;
}

int orc_run(orc_sys* s, int32_t nn_first, int32_t nsteps) {
  for (int32_t t = 0; t < nsteps; t++) {
    int rc = orc_time_step(s, nn_first + t);
    if (rc) return rc;
  }
  return 0;
}

int orc_get_forces(const orc_sys* s, double* fc, double* tow, int64_t ld) {
  int64_t L = s->ld, n = s->max_pip < ld ? s->max_pip : ld;
  for (int64_t q = 0; q < n; q++)
    for (int c = 0; c < 3; c++) {
      if (fc) fc[c * ld + q] = s->fc_last[c * L + q];
      if (tow) tow[c * ld + q] = s->tow_last[c * L + q];
    }
  return 0;
}

int orc_get_cells(const orc_sys* s, int32_t* dg_pijk, int32_t* pijk, int64_t ld) {
  int64_t L = s->ld, n = s->max_pip < ld ? s->max_pip : ld;
  for (int64_t q = 0; q < n; q++) {
    if (dg_pijk) dg_pijk[q] = s->dg_pijk[q];
    if (pijk)
      for (int c = 0; c < 4; c++) pijk[c * ld + q] = s->pijk[c * L + q];
  }
  return 0;
}

int orc_get_pairs(const orc_sys* s, int32_t* neighbor_index, int32_t* neighbors, double* pft,
                  int64_t capacity, int64_t* npairs) {
  int64_t np = s->max_pip;
  int64_t total = s->neighbor_index[np - 1] - 1;
  if (total < 0) total = 0;
  if (npairs) *npairs = total;
  if (capacity == 0) return 0;
  if (capacity < total) return MFIXDEM_ERR_CAPACITY;
  if (neighbor_index) memcpy(neighbor_index, s->neighbor_index, sizeof(int) * np);
  if (neighbors) memcpy(neighbors, s->neighbors, sizeof(int) * total);
  if (pft) memcpy(pft, s->pft, sizeof(double) * 3 * total);
  return 0;
}

int orc_get_status(const orc_sys* s, mfixdem_status* out) {
  *out = s->st;
  int64_t n = 0;
  for (int64_t q = 0; q < s->max_pip; q++) n += (s->state[q] != NONEXISTENT);
  out->n_particles = n;
  int64_t tot = s->neighbor_index[s->max_pip - 1] - 1;
  out->n_pairs = tot > 0 ? tot : 0;
  return 0;
}

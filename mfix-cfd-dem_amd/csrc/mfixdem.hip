// mfixdem.hip -- C ABI of the gfx950 DEM hot path (include/mfixdem.h).
//
// One handle = one DEM system resident on the current HIP device, driven on
// one stream.  The time step is a fused particle kernel (walls, pairs, drag,
// integrator, fluid-cell locate) plus, when gas-coupled, the deposition
// kernels; every NEIGHBOR_SEARCH_N steps the particles are re-sorted and the
// neighbour list rebuilt.  Runs of steps between rebuilds are replayed from
// captured hipGraphs.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <array>
#include <mutex>
#include <vector>

#include "../../include/mfixdem.h"
#include "device_physics.hpp"
#include "host_setup.hpp"
#include "kernels.hpp"
#include "dem_bcs.hpp"
#include "params.hpp"

using namespace mfixdem;

#define HIPOK(x)                                   \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) {                        \
      std::fprintf(stderr, "mfixdem: %s failed: %s (%s:%d)\n", #x, \
                   hipGetErrorString(e_), __FILE__, __LINE__);     \
      return MFIXDEM_ERR_HIP;                      \
    }                                              \
  } while (0)

#define CHECKRC(x)            \
  do {                        \
    int32_t r_ = (x);         \
    if (r_ != MFIXDEM_OK) return r_; \
  } while (0)

enum { kTimeStep = MFIXDEM_TIME_STEP, kTimeRebuild = MFIXDEM_TIME_REBUILD,
       kTimeGas = MFIXDEM_TIME_GAS, kTimeSlots = MFIXDEM_TIME_GROUPS };

struct mfixdem_handle {
  Setup S;
  DevParams P{};
  GridDev G{};
  // ensembles (nsys > 1): per-system set-ups and device tables
  int nsys = 1;
  BatchDev B{};
  int key_bits = 0;                       // significant bits of the rebuild sort key
  std::vector<Setup> sys_setup;           // per system (ensembles)
  std::vector<GridDev> sys_grid;          // per-system grids (host copy of d_gtab)
  std::vector<int> sys_off;               // first slot of each system (+ total)
  std::vector<int> sys_n;                 // particles set per system
  std::vector<int> rogue_init;            // {INT_MAX, 0} per system
  uint16_t* d_sys = nullptr;
  DevParams* d_ptab = nullptr;
  GridDev* d_gtab = nullptr;
  int *d_cell_off = nullptr, *d_dg_off = nullptr, *d_rogue = nullptr;
  // host staging of an ensemble's particles until mfixdem_init
  std::vector<double> st_pos, st_vel, st_om, st_r, st_rho;
  std::vector<int> st_phase;
  hipStream_t stream = nullptr;
  int device = 0;              // HIP device the handle was created on
  double* scratch_ptr = nullptr;  // that device's address of `scratch`
  int n = 0, stride = 0;
  int cur = 0, pcur = 0, tw = 0, lt = 0;
  // particles (cur / tw twins)
  double *pos[2]{}, *vel[2]{}, *omega[2]{};
  double* xr_pack = nullptr;  // (x, y, z, radius) per particle, written before each row build
  double *radius[2]{}, *pmass[2]{}, *omoi[2]{}, *pvol[2]{};
  double *acc[2]{}, *racc[2]{};
  int *gid[2]{}, *fcell[2]{}, *dgc[2]{};
  uint8_t *state[2]{}, *phase[2]{}, *wmask[2]{};
  int* wfid[2]{};       // wall-collision slots: facet id per slot
  double* wpft[2]{};    // and history per slot
  double *fc = nullptr, *tow = nullptr, *fc_id = nullptr, *tow_id = nullptr;
  // rebuild
  uint64_t* keys[2]{};
  int* vals[2]{};
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int *fstart = nullptr, *fend = nullptr;
  int n_fine = 0;
  // one system: counting sort over the search cells (Morton rank of each
  // cell, counts with one more cell for nonexistent particles, their scan)
  int *fine_rank = nullptr, *fine_cnt = nullptr, *fine_off = nullptr;
  int *own_cnt = nullptr, *own_tmp = nullptr, *in_cnt = nullptr, *in_cur = nullptr,
      *tot_cnt = nullptr;
  uint64_t* own_key = nullptr;
  int maxrow = 48;
  // tuning (mfixdem_set_tuning): search-cell size / cut-off, row kernel
  double cell_scale = 1.0;
  int row_kernel = MFIXDEM_ROWS_THREAD;
  int *nbr_start[2]{}, *nbr_split[2]{}, *pair_base[2]{};
  int *nbr[2]{}, *nbr_pid[2]{};
  int64_t nbr_cap[2]{0, 0};
  double* pft[2]{};
  uint8_t* pfl[2]{};            // per pair: 1 when pft holds the history, 0 when it is zero
  int64_t pft_cap[2]{0, 0};
  int64_t n_pairs = 0;
  int* scan_bsum = nullptr;
  long long* scan_bsum64 = nullptr;
  bool counts_zeroed = false;  // own_cnt, in_cnt, in_cur cleared by the last binning
  bool xr_packed = false;      // the row walk's candidate data written by the last binning
  bool flags_set = false;      // the rebuild flags hold their start values (k_carry sets them)
  // A binning of one system with the thread-row search moves the particle
  // data into the search-cell order only every kPermuteEvery-th time (and
  // the first time); otherwise the slots stay and bin_idx lists them in that
  // order.  bin_inv: old slot -> new slot of the last binning (null: the
  // slots stayed), for the history carry.
  const int* bin_idx = nullptr;
  const int* bin_inv = nullptr;
  int bins_since_permute = 0;
  bool order_set = false;
  int* flags = nullptr;
  // gas
  double* ugc = nullptr;  // cell-centre gas velocity (filter schemes)
  double* gas_stage = nullptr;  // 3*ncell: staged U_G/V_G/W_G, P_G or P_FORCE uploads
  double* csz = nullptr;  // cell sizes DX(1..ni), DY(1..nj), DZ(1..nk)
  double *xe = nullptr, *yn = nullptr, *zt = nullptr, *vnode = nullptr, *p_force = nullptr,
         *ep_g = nullptr, *rop_s = nullptr, *rops_node = nullptr, *vol = nullptr,
         *vol_surr = nullptr, *ro_s = nullptr;
  uint8_t* fluid = nullptr;
  int *pic_cnt = nullptr, *pic_off = nullptr, *pic_cur = nullptr, *pic = nullptr;
  // gas side of the drag (DRAG_GS_GAS0), allocated on first use: the F_GP
  // record per particle slot, DES_VOL_NODE, DRAG_AM, DRAG_BM (3 planes), F_GDS
  double *gd_fgp = nullptr, *gd_vol_node = nullptr, *gd_am = nullptr, *gd_bm = nullptr,
         *gd_fgds = nullptr;
  bool gas_drag_ready = false;
  unsigned long long* rops_fix[2]{};  // fixed-point node sums, double-buffered
  // removals queued by the fused step: two counters (used by turns with
  // the node-sum buffers) and the slot list
  int* rm_cnt = nullptr;
  int* rm_list = nullptr;
  // explicit coupling (global-id order): F_GP of the gas side, DRAG_FC
  double *f_gp = nullptr, *drag_fc = nullptr;
  int fb = 0;
  double dep_inv[MFIXDEM_MAX_PHASES]{};
  // walls: facets and the facet lists of the DES cells (upload_facets)
  FacetDev* fac = nullptr;
  int *dg_foff = nullptr, *dg_fids = nullptr;
  uint8_t* cell_fmask = nullptr;
  double* stage = nullptr;  // export staging, 3*n doubles
  // DEM mass inlets / outlets (mfixdem_set_dem_bcs): host schedule and
  // layout, the device table, this step's seeds, the slot of each global id
  // (rewritten at every binning) and the particles deleted in a step
  DemBoundaries dbc;
  BcDev* d_bcs = nullptr;
  SeedDev* d_seeds = nullptr;
  int *slot_of_gid = nullptr, *del_cnt = nullptr, *del_list = nullptr;
  // status
  mfixdem_status st{};
  bool inited = false, gas_set = false, first_pass = true, binned = false;
  int debug = 0;
  // graphs: [cur parity][deposition buffer]
  hipGraphExec_t graph[2][2]{};
  // executable graphs replaced while launches of them may still be queued on
  // the stream; destroyed once the stream has drained (reap_graphs)
  std::vector<hipGraphExec_t> retired;
  // kernel timing
  // kernel timing: a pool of events recorded around each timed launch and
  // read back in one go, so launches queue back to back while timed
  std::vector<hipEvent_t> ev;
  int ev_next = 0, ev_open = -1;
  std::vector<std::array<int, 3>> t_pending;   // {slot, begin event, end event}
  double t_ms[kTimeSlots]{};
  int64_t t_calls[kTimeSlots]{};
};

// Device scratch of mfixdem_drag_gs_gas_noalloc, one copy per GPU, held
// until the process ends: the per-particle F_GP records (kGasRec doubles per
// particle slot) followed by DES_VOL_NODE (one double per cell).  Host code
// never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in); it uses the device address resolved by bind_scratch.
constexpr int64_t kScratchDoubles = 10LL << 20;  // 80 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

inline int nblk(int64_t n) { return (int)((n + kBS - 1) / kBS); }
inline int field_blocks(int64_t ncell) {
  return (int)((ncell + kFieldCells - 1) / kFieldCells);
}

// Stable LSD radix sort of (key, slot) pairs on key bits [0, bits).  The
// library default switches to a merge sort up to 2^20 items; for the few
// key bits used here the onesweep radix passes are several times faster,
// so the merge-sort limit is set to 0.
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;
hipError_t sort_pairs(void* tmp, size_t& bytes, const uint64_t* kin, uint64_t* kout,
                      const int* vin, int* vout, int n, int bits, hipStream_t s) {
  return rocprim::radix_sort_pairs<SortConfig>(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0u,
                                               (unsigned)bits, s);
}

template <typename T>
int32_t dalloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) count = 1;
  if (hipMalloc((void**)p, sizeof(T) * (size_t)count) != hipSuccess) return MFIXDEM_ERR_OOM;
  return MFIXDEM_OK;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int32_t fill_i32(hipStream_t s, int* a, int64_t n, int v) {
  if (n <= 0) return MFIXDEM_OK;
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(kBS), 0, s, a, (int)n, v);
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

// exclusive scan of in[0..n) -> out[0..n]; out[n] = total
int32_t scan_i32(mfixdem_handle* h, const int* in, int* out, int64_t n) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(std::max(nb, 1)), dim3(kBS), 0, h->stream, in, out,
                     (int)n, h->scan_bsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBS), 0, h->stream, h->scan_bsum, std::max(nb, 1));
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(n + 1)), dim3(kBS), 0, h->stream, out, (int)n,
                     h->scan_bsum, std::max(nb, 1));
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

// Pair ids and row starts of the neighbour lists in one scan of the own and
// incoming counts (k_scan_rows_tiles).
int32_t scan_rows(mfixdem_handle* h, const int* own, const int* in, int* pair_base,
                  int* row_start, int64_t n) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_rows_tiles, dim3(std::max(nb, 1)), dim3(kBS), 0, h->stream, own, in,
                     pair_base, row_start, (int)n, h->scan_bsum64);
  hipLaunchKernelGGL(k_scan_sums64, dim3(1), dim3(kBS), 0, h->stream, h->scan_bsum64,
                     std::max(nb, 1));
  hipLaunchKernelGGL(k_scan_rows_add, dim3(nblk(n + 1)), dim3(kBS), 0, h->stream, pair_base,
                     row_start, (int)n, h->scan_bsum64, std::max(nb, 1), h->flags);
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

StepPtrs step_ptrs(mfixdem_handle* h) {
  StepPtrs a{};
  const int c = h->cur, t = h->tw, l = h->lt, p = h->pcur;
  a.pos = h->pos[c]; a.vel = h->vel[c]; a.omega = h->omega[c];
  a.pos_o = h->pos[1 - c]; a.vel_o = h->vel[1 - c]; a.omega_o = h->omega[1 - c];
  a.radius = h->radius[t]; a.pmass = h->pmass[t]; a.omoi = h->omoi[t]; a.pvol = h->pvol[t];
  a.gid = h->gid[t]; a.state = h->state[t]; a.phase = h->phase[t]; a.dgc = h->dgc[t];
  a.cell_fmask = h->cell_fmask;
  a.fac = h->fac; a.dg_foff = h->dg_foff; a.dg_fids = h->dg_fids;
  a.wmask = h->wmask[t]; a.wfid = h->wfid[t]; a.wpft = h->wpft[t];
  a.nbr_start = h->nbr_start[l]; a.nbr_split = h->nbr_split[l];
  a.nbr = h->nbr[l]; a.nbr_pid = h->nbr_pid[l];
  a.pft_in = h->pft[p]; a.pft_out = h->pft[1 - p];
  a.pfl_in = h->pfl[p]; a.pfl_out = h->pfl[1 - p];
  a.fc = h->fc; a.tow = h->tow; a.fc_id = h->fc_id; a.tow_id = h->tow_id;
  a.acc_old = h->acc[t]; a.rot_acc_old = h->racc[t];
  a.fcell = h->fcell[t];
  a.xe = h->xe; a.yn = h->yn; a.zt = h->zt; a.csz = h->csz;
  a.vnode = h->vnode; a.p_force = h->p_force; a.ep_g = h->ep_g; a.fluid = h->fluid;
  a.ugc = h->ugc;
  a.flags = h->flags;
  a.dep_fix = h->rops_fix[h->fb];
  a.rm_cnt = h->rm_cnt ? h->rm_cnt + h->fb : nullptr;
  a.rm_list = h->rm_list;
  a.f_gp = h->f_gp;
  a.drag_fc = h->drag_fc;
  a.bcs = h->d_bcs;
  return a;
}

// The captured graphs no longer match the handle's buffers: they are
// retired, not destroyed, because launches of them may still be queued on
// the handle's stream; sync_stream destroys them once the stream has drained.
void invalidate_graphs(mfixdem_handle* h) {
  for (auto& row : h->graph)
    for (auto& g : row)
      if (g) {
        h->retired.push_back(g);
        g = nullptr;
      }
}
void reap_graphs(mfixdem_handle* h) {
  for (hipGraphExec_t g : h->retired) (void)hipGraphExecDestroy(g);
  h->retired.clear();
}
// Waits for the handle's stream; every wait goes through here.
hipError_t sync_stream(mfixdem_handle* h) {
  const hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) reap_graphs(h);
  return e;
}

// New particle state: deposition restarts from buffer 0 with both
// fixed-point node-sum buffers zero (the last step left its sums in one of
// them; only the other is cleared by the cell-field kernel).
int32_t reset_deposit(mfixdem_handle* h) {
  h->fb = 0;
  const size_t nb = sizeof(unsigned long long) * (size_t)h->S.ncell * h->S.p.nphases;
  for (int b = 0; b < 2; b++)
    if (h->rops_fix[b]) HIPOK(hipMemsetAsync(h->rops_fix[b], 0, nb, h->stream));
  if (h->rm_cnt) HIPOK(hipMemsetAsync(h->rm_cnt, 0, 2 * sizeof(int), h->stream));
  return MFIXDEM_OK;
}

// the reference-order deposition exists for the GARG_2012 node sums
bool ref_deposition(const mfixdem_handle* h) {
  return (h->debug & MFIXDEM_DEBUG_REFERENCE_DEPOSITION) != 0 &&
         h->S.p.interp_scheme == MFIXDEM_INTERP_GARG_2012;
}
// The fused step kernel deposits (GARG_2012 node sums) for a coupled
// single system; the filter schemes and the reference-order deposition
// keep their own kernels.
bool fused_dep(const mfixdem_handle* h) {
  return h->nsys == 1 && h->S.p.coupled && h->rm_cnt && !h->S.p.des_explicitly_coupled &&
         h->S.p.interp_scheme == MFIXDEM_INTERP_GARG_2012 && !ref_deposition(h);
}

bool timing(const mfixdem_handle* h) { return (h->debug & MFIXDEM_DEBUG_TIME_KERNELS) != 0; }
constexpr int kTimeEvents = 2048;

// Waits for the recorded timing events and adds their intervals to t_ms.
void t_flush(mfixdem_handle* h) {
  if (!h->t_pending.empty()) {
    (void)hipEventSynchronize(h->ev[h->t_pending.back()[2]]);
    for (const auto& t : h->t_pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, h->ev[t[1]], h->ev[t[2]]) == hipSuccess) {
        h->t_ms[t[0]] += ms;
        h->t_calls[t[0]] += 1;
      }
    }
    h->t_pending.clear();
  }
  h->ev_next = 0;
  h->ev_open = -1;
}
void t_begin(mfixdem_handle* h) {
  if (!timing(h)) return;
  if (h->ev.empty()) {
    h->ev.assign(kTimeEvents, nullptr);
    // timing-only events: no system-scope fence (cache writeback and
    // invalidate) when they are recorded
    for (auto& e : h->ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) e = nullptr;
  }
  if (h->ev_next + 2 > kTimeEvents) t_flush(h);
  h->ev_open = h->ev_next++;
  if (h->ev[h->ev_open]) (void)hipEventRecord(h->ev[h->ev_open], h->stream);
}
void t_end(mfixdem_handle* h, int slot) {
  if (!timing(h) || h->ev_open < 0) return;
  const int b = h->ev_next++;
  if (h->ev[h->ev_open] && h->ev[b]) {
    (void)hipEventRecord(h->ev[b], h->stream);
    h->t_pending.push_back({slot, h->ev_open, b});
  }
  h->ev_open = -1;
}

template <int MODE>
int32_t launch_step(mfixdem_handle* h) {
  DevParams P = h->P;
  P.first_pass = h->first_pass ? 1 : 0;
  P.store_forces = (h->debug & MFIXDEM_DEBUG_STORE_FORCES) ? 1 : 0;
  P.fuse_dep = (MODE == kModeFused && fused_dep(h)) ? 1 : 0;
  StepPtrs a = step_ptrs(h);
  const dim3 grid(nblk(h->n)), blk(kBS);
  const bool mixed = h->S.p.precision == MFIXDEM_MIXED;
  const BatchDev B = h->B;
  if (h->nsys > 1) {  // ensembles: pure DEM, per-system parameter tables
    if (mixed) hipLaunchKernelGGL((k_step<true, MODE, false, true>), grid, blk, 0, h->stream, P, a, B);
    else hipLaunchKernelGGL((k_step<false, MODE, false, true>), grid, blk, 0, h->stream, P, a, B);
  } else if (P.nfs == 0) {  // more facets than wall slots: facet table and slot search
    const bool filt = h->S.p.interp_scheme != MFIXDEM_INTERP_GARG_2012;
    if (filt && mixed)
      hipLaunchKernelGGL((k_step<true, MODE, true, false, true>), grid, blk, 0, h->stream, P, a, B);
    else if (filt)
      hipLaunchKernelGGL((k_step<false, MODE, true, false, true>), grid, blk, 0, h->stream, P, a, B);
    else if (mixed)
      hipLaunchKernelGGL((k_step<true, MODE, false, false, true>), grid, blk, 0, h->stream, P, a, B);
    else
      hipLaunchKernelGGL((k_step<false, MODE, false, false, true>), grid, blk, 0, h->stream, P, a, B);
  } else if (h->S.p.interp_scheme != MFIXDEM_INTERP_GARG_2012) {
    if (mixed) hipLaunchKernelGGL((k_step<true, MODE, true>), grid, blk, 0, h->stream, P, a, B);
    else hipLaunchKernelGGL((k_step<false, MODE, true>), grid, blk, 0, h->stream, P, a, B);
  } else {
    if (mixed) hipLaunchKernelGGL((k_step<true, MODE, false>), grid, blk, 0, h->stream, P, a, B);
    else hipLaunchKernelGGL((k_step<false, MODE, false>), grid, blk, 0, h->stream, P, a, B);
  }
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

int32_t ensure_list_capacity(mfixdem_handle* h, int nl, int64_t entries) {
  if (h->nbr_cap[nl] >= entries) return MFIXDEM_OK;
  const int64_t cap = std::max<int64_t>(entries + entries / 4, 1024);
  dfree(h->nbr[nl]);
  dfree(h->nbr_pid[nl]);
  CHECKRC(dalloc(&h->nbr[nl], cap));
  CHECKRC(dalloc(&h->nbr_pid[nl], cap));
  h->nbr_cap[nl] = cap;
  return MFIXDEM_OK;
}

// history buffer b must hold `pairs` entries; its contents are not kept
int32_t ensure_pft_capacity(mfixdem_handle* h, int b, int64_t pairs) {
  if (h->pft_cap[b] >= pairs && h->pft[b]) return MFIXDEM_OK;
  const int64_t cap = std::max<int64_t>(pairs + pairs / 4, 1024);
  dfree(h->pft[b]);
  dfree(h->pfl[b]);
  CHECKRC(dalloc(&h->pft[b], 3 * cap));
  CHECKRC(dalloc(&h->pfl[b], cap));
  h->pft_cap[b] = cap;
  return MFIXDEM_OK;
}

// The particle data moves into the search-cell order every kPermuteEvery-th
// binning: particles move little between binnings (at most a fraction of a
// search cell in NEIGHBOR_SEARCH_N steps), so the slots stay close to that
// order for many binnings.
#ifndef MFIXDEM_PERMUTE_EVERY
#define MFIXDEM_PERMUTE_EVERY 16
#endif
constexpr int kPermuteEvery = MFIXDEM_PERMUTE_EVERY;

// DESGRID_PIC(.TRUE.) (desgrid_mod.f:877): DES cells, search-grid sort,
// permutation of every per-particle array into the new order.
int32_t do_bin(mfixdem_handle* h) {
  const int n = h->n, S = h->stride;
  int init[kFlagCount] = {INT_MAX, 0, 0, 0, 0, 0, 0};
  // only the rebuild flags are reset here
  const bool thread_rows = h->S.p.des_neighbor_search != MFIXDEM_NSEARCH_NSQUARE &&
                           (h->row_kernel == MFIXDEM_ROWS_THREAD || h->P.nbc);
  const bool permute = h->nsys > 1 || !thread_rows || !h->order_set ||
                       h->bins_since_permute + 1 >= kPermuteEvery;
  if (!h->flags_set) {
    HIPOK(hipMemcpyAsync(h->flags + kFlagRogueGid, init, 2 * sizeof(int),
                         hipMemcpyHostToDevice, h->stream));
    HIPOK(hipMemcpyAsync(h->flags + kFlagOwnOverflow, init + kFlagOwnOverflow, sizeof(int),
                         hipMemcpyHostToDevice, h->stream));
  }
  h->flags_set = false;
  const bool batch = h->nsys > 1;
  if (batch)
    HIPOK(hipMemcpyAsync(h->d_rogue, h->rogue_init.data(), sizeof(int) * 2 * h->nsys,
                         hipMemcpyHostToDevice, h->stream));
  const int t = h->tw, c = h->cur;
  if (batch) {  // ensembles: key sort (system id above the Morton code)
    hipLaunchKernelGGL((k_bin<true>), dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S, h->G,
                       h->pos[c], h->state[t], h->gid[t], h->dgc[t], h->keys[0], h->vals[0],
                       h->flags, h->B);
    HIPOK(hipGetLastError());
    size_t tb = h->sort_tmp_bytes;
    HIPOK(sort_pairs(h->sort_tmp, tb, h->keys[0], h->keys[1], h->vals[0], h->vals[1], n,
                     h->key_bits, h->stream));
    hipLaunchKernelGGL(k_invert, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->vals[1],
                       h->vals[0]);
    HIPOK(hipGetLastError());
  } else {  // one system: counting sort over the search cells
    const int nf = h->n_fine;
    hipLaunchKernelGGL(k_bin_count, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S, h->G,
                       h->pos[c], h->state[t], h->gid[t], h->dgc[t], h->fine_rank, h->fine_cnt,
                       nf, h->keys[0], h->flags);
    HIPOK(hipGetLastError());
    CHECKRC(scan_i32(h, h->fine_cnt, h->fine_off, (int64_t)nf + 1));
    BinPack bp{};
    bp.G = h->G;
    bp.S = S;
    bp.pos = h->pos[c];
    bp.radius = h->radius[t];
    bp.state = h->state[t];
    bp.xr = h->xr_pack;
    bp.xf = reinterpret_cast<float4*>(h->stage);
    hipLaunchKernelGGL(k_bin_place, dim3(nblk(std::max(n, nf))), dim3(kBS), 0, h->stream, n, nf,
                       h->keys[0], h->fine_rank, h->fine_off, h->fine_cnt, h->vals[1],
                       h->vals[0], h->fstart, h->fend, h->own_cnt, h->in_cnt, h->in_cur, bp);
    HIPOK(hipGetLastError());
    h->counts_zeroed = true;
    h->xr_packed = true;
  }
  PermPtrs p{};
  p.perm = h->vals[1];
  p.pos = h->pos[c]; p.vel = h->vel[c]; p.omega = h->omega[c];
  p.pos_o = h->pos[1 - c]; p.vel_o = h->vel[1 - c]; p.omega_o = h->omega[1 - c];
  p.radius = h->radius[t]; p.pmass = h->pmass[t]; p.omoi = h->omoi[t]; p.pvol = h->pvol[t];
  p.acc = h->acc[t]; p.racc = h->racc[t];
  p.radius_o = h->radius[1 - t]; p.pmass_o = h->pmass[1 - t]; p.omoi_o = h->omoi[1 - t];
  p.pvol_o = h->pvol[1 - t]; p.acc_o = h->acc[1 - t]; p.racc_o = h->racc[1 - t];
  p.gid = h->gid[t]; p.fcell = h->fcell[t]; p.dgc = h->dgc[t];
  p.gid_o = h->gid[1 - t]; p.fcell_o = h->fcell[1 - t]; p.dgc_o = h->dgc[1 - t];
  p.state = h->state[t]; p.phase = h->phase[t]; p.wmask = h->wmask[t];
  p.state_o = h->state[1 - t]; p.phase_o = h->phase[1 - t]; p.wmask_o = h->wmask[1 - t];
  p.wpft = h->wpft[t]; p.wpft_o = h->wpft[1 - t];
  p.wfid = h->wfid[t]; p.wfid_o = h->wfid[1 - t];
  p.ab2 = h->P.ab2;
  if (permute) {
    hipLaunchKernelGGL(k_permute, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S, p);
    HIPOK(hipGetLastError());
    h->cur = 1 - c;
    h->tw = 1 - t;
    h->bin_idx = nullptr;
    h->bin_inv = h->vals[0];
    h->bins_since_permute = 0;
    h->order_set = true;
  } else {
    h->bin_idx = h->vals[1];
    h->bin_inv = nullptr;
    h->bins_since_permute++;
  }
  if (batch) {
    CHECKRC(fill_i32(h->stream, h->fstart, h->n_fine, 0));
    CHECKRC(fill_i32(h->stream, h->fend, h->n_fine, 0));
    hipLaunchKernelGGL((k_cell_ranges<true>), dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->G,
                       h->keys[1], h->state[h->tw], h->fstart, h->fend, h->B);
    HIPOK(hipGetLastError());
  }
  if (h->slot_of_gid) {  // inlets place their seeds by global id
    hipLaunchKernelGGL(k_slot_of_gid, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->gid[h->tw],
                       h->slot_of_gid);
    HIPOK(hipGetLastError());
  }
  h->binned = true;
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

template <bool BATCH>
void launch_rows_wave_t(mfixdem_handle* h, const BuildPtrs& b, int64_t threads) {
  const dim3 g(nblk(threads)), blk(kBS);
  const int S = h->stride;
  switch (h->row_kernel) {
    case MFIXDEM_ROWS_WAVE128:
      hipLaunchKernelGGL((k_own_rows_wave<BATCH, 128>), g, blk, 0, h->stream, S, h->G, b,
                         h->n_fine, h->B);
      break;
    case MFIXDEM_ROWS_WAVE256:
      hipLaunchKernelGGL((k_own_rows_wave<BATCH, 256>), g, blk, 0, h->stream, S, h->G, b,
                         h->n_fine, h->B);
      break;
    default:
      hipLaunchKernelGGL((k_own_rows_wave<BATCH, 192>), g, blk, 0, h->stream, S, h->G, b,
                         h->n_fine, h->B);
  }
}
void launch_rows_wave(mfixdem_handle* h, const BuildPtrs& b, int64_t threads) {
  if (h->nsys > 1) launch_rows_wave_t<true>(h, b, threads);
  else launch_rows_wave_t<false>(h, b, threads);
}

// NEIGHBOUR (neighbour.f:21-126) on the binned order.
// The reference's fatal device-side conditions, as counted in the flags:
// a coincident pair (calc_force_dem.f:267-271) and EP_G out of (0, 1] or a
// deposition sum out of the fixed-point range (calc_epg_des.f:146-203).
// Sticky: the counts are cumulative, so every later check reports them.
int32_t fatal_flags(const int* fl) {
  if (fl[kFlagZeroDist] > 0) return MFIXDEM_ERR_ZERO_DISTANCE;
  if (fl[kFlagEpg] != 0) return MFIXDEM_ERR_EPG;
  if (fl[kFlagWallSlots] != 0) return MFIXDEM_ERR_CAPACITY;  // > 8 live wall contacts
  return MFIXDEM_OK;
}

int32_t do_neighbour(mfixdem_handle* h) {
  if (!h->binned) CHECKRC(do_bin(h));
  const int n = h->n, S = h->stride, t = h->tw, c = h->cur;
  BuildPtrs b{};
  b.pos = h->pos[c]; b.radius = h->radius[t]; b.gid = h->gid[t]; b.dgc = h->dgc[t];
  b.xr = h->xr_pack;
  // the single-precision copy for the walk's prefilter lives in the export
  // stage (3 doubles per slot), which no call uses during a rebuild
  b.xf = reinterpret_cast<float4*>(h->stage);
  b.state = h->state[t]; b.fstart = h->fstart; b.fend = h->fend;
  b.idx = h->bin_idx;
  b.own_cnt = h->own_cnt; b.flags = h->flags; b.in_cnt = h->in_cnt;
  int fl[kFlagCount];
  // (the wavefront row kernels take normal particles only: boundaries use
  // the thread kernel, whose owner rule knows entering and exiting ones)
  bool wave_rows = h->row_kernel != MFIXDEM_ROWS_THREAD && !h->P.nbc;
  // neighbour.f:60-66: NSQUARE for DES_NEIGHBOR_SEARCH = 1 (no rogue handling
  // there: that lives in DESGRID_NEIGH_BUILD)
  const bool nsq = h->S.p.des_neighbor_search == MFIXDEM_NSEARCH_NSQUARE;
  const int nl = 1 - h->lt;
  int npairs = 0;
  // placement of the built rows (k_place, the history carry, the incoming
  // order); guarded when launched before the host has read the flags
  auto place = [&](const PlaceGuard& pg) -> int32_t {
    CarryPtrs cp{};
    cp.flags = h->flags;
    cp.old_start = h->nbr_start[h->lt]; cp.old_split = h->nbr_split[h->lt];
    cp.old_nbr = h->nbr[h->lt]; cp.old_pair_base = h->pair_base[h->lt];
    cp.pft_old = h->pft[h->pcur];
    cp.pft_new = h->pft[1 - h->pcur];
    cp.pfl_old = h->pfl[h->pcur];
    cp.pfl_new = h->pfl[1 - h->pcur];
    hipLaunchKernelGGL(k_place, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->own_cnt,
                       h->own_tmp, nsq || wave_rows ? nullptr : h->own_key, h->pair_base[nl],
                       h->nbr_start[nl], h->in_cnt, h->nbr_split[nl], h->in_cur, h->nbr[nl],
                       h->nbr_pid[nl], cp.pfl_new, pg);
    hipLaunchKernelGGL(k_carry, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->bin_inv, cp,
                       h->pair_base[nl], h->nbr_split[nl], h->nbr_start[nl], h->nbr[nl], pg);
    hipLaunchKernelGGL(k_sort_in, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->gid[t],
                       h->nbr_start[nl], h->nbr_split[nl], h->nbr[nl], h->nbr_pid[nl], pg);
    HIPOK(hipGetLastError());
    return MFIXDEM_OK;
  };
  PlaceGuard spec{};  // the guard of a placement launched ahead of the wait
  bool placed = false;
  bool zeroed = h->counts_zeroed;       // row counters cleared by the binning
  const bool cur_zeroed = zeroed;
  const bool packed = h->xr_packed;     // and the candidates packed by it
  h->counts_zeroed = false;
  h->xr_packed = false;
  for (;;) {
    b.own_tmp = h->own_tmp; b.own_key = h->own_key; b.maxrow = h->maxrow; b.ld = n;
    if (nsq) {
      hipLaunchKernelGGL(k_own_rows_nsquare, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S,
                         h->S.p.factor_rlm, b);
    } else if (wave_rows) {
      // particles outside every cell range (state 0) get no row
      CHECKRC(fill_i32(h->stream, h->own_cnt, n, 0));
      const int64_t threads = (int64_t)h->n_fine * 64;
      launch_rows_wave(h, b, threads);
    } else {
      // half-stencil rows appended through atomic slots (each entry also
      // counted in its partner's incoming row), put in order by k_place
      if (!zeroed) {
        CHECKRC(fill_i32(h->stream, h->own_cnt, n, 0));
        CHECKRC(fill_i32(h->stream, h->in_cnt, n, 0));
      }
      if (!packed)
        hipLaunchKernelGGL(k_pack_xr, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S, h->G,
                           b.idx, b.pos, b.radius, b.state, h->xr_pack,
                           reinterpret_cast<float4*>(h->stage));
      const dim3 rg(nblk((int64_t)n * kRowLanes));  // kRowLanes threads per particle
      if (h->nsys > 1)
        hipLaunchKernelGGL((k_own_rows<true>), rg, dim3(kBS), 0, h->stream, n, S, h->G, b, h->B);
      else
        hipLaunchKernelGGL((k_own_rows<false>), rg, dim3(kBS), 0, h->stream, n, S, h->G, b, h->B);
    }
    HIPOK(hipGetLastError());
    zeroed = false;
    // own rows -> pair ids (and, with the thread rows, whose incoming counts
    // are known already, the row starts), read back with the flags in the
    // same wait (an overflowing pass is built again and its pair ids with it)
    if (nsq || wave_rows) {
      CHECKRC(scan_i32(h, h->own_cnt, h->pair_base[nl], n));
      HIPOK(hipMemcpyAsync(h->flags + kFlagPairs, h->pair_base[nl] + n, sizeof(int),
                           hipMemcpyDeviceToDevice, h->stream));
    } else {
      CHECKRC(scan_rows(h, h->own_cnt, h->in_cnt, h->pair_base[nl], h->nbr_start[nl], n));
    }
    HIPOK(hipMemcpyAsync(fl, h->flags, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
    // thread rows have their row starts already: place them ahead of the
    // wait, guarded, while the host waits for the flags
    placed = false;
    if (!nsq && !wave_rows && cur_zeroed && h->nbr[nl] && h->pft[1 - h->pcur]) {
      spec = PlaceGuard{h->flags, h->maxrow, h->nbr_cap[nl], h->pft_cap[1 - h->pcur]};
      CHECKRC(place(spec));
      placed = true;
    }
    HIPOK(sync_stream(h));
    npairs = fl[kFlagPairs];
    if (fl[kFlagOwnOverflow] <= h->maxrow) break;
    int z = 0;
    HIPOK(hipMemcpyAsync(h->flags + kFlagOwnOverflow, &z, sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
    if (fl[kFlagOwnOverflow] >= (1 << 20)) {
      // a row longer than the wavefront kernel holds: thread-per-particle rows
      wave_rows = false;
      continue;
    }
    const int mr = fl[kFlagOwnOverflow] + 16;
    dfree(h->own_tmp);
    dfree(h->own_key);
    CHECKRC(dalloc(&h->own_tmp, (int64_t)n * mr));
    CHECKRC(dalloc(&h->own_key, (int64_t)n * mr));
    h->maxrow = mr;
  }
  CHECKRC(fatal_flags(fl));
  if (!nsq && fl[kFlagRogueCount] > 0 && !h->S.p.remove_rogue) return MFIXDEM_ERR_ROGUE;
  const bool fits = placed && 2 * (int64_t)npairs <= spec.cap_entries &&
                    (int64_t)npairs <= spec.cap_pairs;
  if (!fits) {
    CHECKRC(ensure_list_capacity(h, nl, 2 * (int64_t)npairs));
    CHECKRC(ensure_pft_capacity(h, 1 - h->pcur, npairs));
    if (nsq || wave_rows) {  // rows built sorted: count the incoming entries here
      CHECKRC(fill_i32(h->stream, h->in_cnt, n, 0));
      hipLaunchKernelGGL(k_count_in, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->own_cnt,
                         h->own_tmp, h->maxrow, h->in_cnt);
      hipLaunchKernelGGL(k_add_counts, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->in_cnt,
                         h->own_cnt, h->tot_cnt);
      HIPOK(hipGetLastError());
      CHECKRC(scan_i32(h, h->tot_cnt, h->nbr_start[nl], n));
    }
    if (!cur_zeroed) CHECKRC(fill_i32(h->stream, h->in_cur, n, 0));
    CHECKRC(place(PlaceGuard{}));
  }
  h->flags_set = true;
  HIPOK(hipGetLastError());
  h->lt = nl;
  h->pcur = 1 - h->pcur;
  h->n_pairs = npairs;
  // the old history buffer becomes the next step's output; k_place reads it,
  // so a buffer too small is replaced only once the carry is done
  if (h->pft_cap[1 - h->pcur] < npairs || !h->pft[1 - h->pcur]) {
    HIPOK(sync_stream(h));
    CHECKRC(ensure_pft_capacity(h, 1 - h->pcur, npairs));
  }
  if (!nsq && fl[kFlagRogueCount] > 0) {
    if (h->nsys > 1)
      hipLaunchKernelGGL((k_delete_rogue<true>), dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S,
                         h->G, h->dgc[t], h->state[t], h->wmask[t], h->pos[c], h->vel[c],
                         h->omega[c], h->radius[t], h->B);
    else
      hipLaunchKernelGGL((k_delete_rogue<false>), dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S,
                         h->G, h->dgc[t], h->state[t], h->wmask[t], h->pos[c], h->vel[c],
                         h->omega[c], h->radius[t], h->B);
    HIPOK(hipGetLastError());
    h->st.rogue_deleted += fl[kFlagRogueCount];
  }
  h->binned = false;
  h->st.rebuilds++;
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

// PARTICLES_IN_CELL lists (particles_in_cell.f:157-229) from fcell
int32_t do_pic_lists(mfixdem_handle* h) {
  const int n = h->n, S = h->stride, t = h->tw, c = h->cur, nc = h->S.ncell;
  CHECKRC(fill_i32(h->stream, h->pic_cnt, nc, 0));
  hipLaunchKernelGGL(k_pic_count, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, S, h->S.ni, h->S.nj,
                     h->fcell[t],
                     h->state[t], h->wmask[t], h->pos[c], h->vel[c], h->omega[c], h->radius[t],
                     h->pic_cnt, h->flags);
  HIPOK(hipGetLastError());
  CHECKRC(scan_i32(h, h->pic_cnt, h->pic_off, nc));
  CHECKRC(fill_i32(h->stream, h->pic_cur, nc, 0));
  hipLaunchKernelGGL(k_pic_scatter, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->S.ni, h->S.nj,
                     h->fcell[t],
                     h->state[t], h->pic_off, h->pic_cur, h->pic);
  hipLaunchKernelGGL(k_pic_sort, dim3(nblk(nc)), dim3(kBS), 0, h->stream, nc, h->pic_off,
                     h->gid[t], h->pic);
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

FieldPtrs field_ptrs(mfixdem_handle* h) {
  FieldPtrs f{};
  f.rops_node = h->rops_node;
  f.fix = h->rops_fix[h->fb];
  f.fix_clear = h->rops_fix[1 - h->fb];
  f.vol = h->vol; f.vol_surr = h->vol_surr; f.ro_s = h->ro_s; f.fluid = h->fluid;
  f.rop_s = h->rop_s; f.ep_g = h->ep_g; f.flags = h->flags;
  if (fused_dep(h)) {
    f.rm_cnt = h->rm_cnt + h->fb;
    f.rm_cnt_clear = h->rm_cnt + (1 - h->fb);
    f.rm_list = h->rm_list;
    f.state = h->state[h->tw];
    f.radius = h->radius[h->tw];
  }
  return f;
}

// node sums -> ROP_S, EP_G (comp_mean_fields0.f:298-370, calc_epg_des.f)
int32_t do_cell_fields(mfixdem_handle* h, bool fixed) {
  const mfixdem_params& p = h->S.p;
  const double packed = (p.des_epg_clip != kUndef) ? 1.0 - p.des_epg_clip : 1.0;
  FieldPtrs f = field_ptrs(h);
  const double* iv = h->dep_inv;
  if (p.interp_scheme != MFIXDEM_INTERP_GARG_2012) {
    hipLaunchKernelGGL(k_cell_fields_des1, dim3(nblk(h->S.ncell)), dim3(kBS), 0, h->stream,
                       h->S.ncell, p.nphases, packed, iv[0], iv[1], iv[2], iv[3], f);
    HIPOK(hipGetLastError());
    h->fb = 1 - h->fb;
    return MFIXDEM_OK;
  }
  if (fixed) {
    hipLaunchKernelGGL(k_cell_fields<true>, dim3(field_blocks(h->S.ncell)), dim3(kBS), 0,
                       h->stream, p.imax, p.jmax, p.kmax, h->S.ni, h->S.nj, h->S.nk, p.nphases,
                       packed, iv[0], iv[1], iv[2], iv[3], f);
    h->fb = 1 - h->fb;
  } else {
    hipLaunchKernelGGL(k_cell_fields<false>, dim3(field_blocks(h->S.ncell)), dim3(kBS), 0,
                       h->stream, p.imax, p.jmax, p.kmax, h->S.ni, h->S.nj, h->S.nk, p.nphases,
                       packed, iv[0], iv[1], iv[2], iv[3], f);
  }
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

int32_t delete_outside(mfixdem_handle* h) {
  const int t = h->tw, c = h->cur;
  hipLaunchKernelGGL(k_delete_outside, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->n, h->stride,
                     h->fcell[t], h->state[t], h->wmask[t], h->pos[c], h->vel[c], h->omega[c],
                     h->radius[t], h->flags);
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

// reference-order deposition: PIC lists sorted by particle id, node gather
int32_t ref_mean_fields(mfixdem_handle* h) {
  const mfixdem_params& p = h->S.p;
  CHECKRC(do_pic_lists(h));
  const int nnode = (p.imax + 1) * (p.jmax + 1) * (p.kmax + 1);
  hipLaunchKernelGGL(k_deposit_nodes, dim3(nblk(nnode)), dim3(kBS), 0, h->stream, p.imax, p.jmax,
                     p.kmax, h->S.ni, h->S.nj, p.nphases, h->stride, h->xe, h->yn, h->zt,
                     h->fluid, h->pic_off, h->pic, h->pos[h->cur], h->pmass[h->tw],
                     h->phase[h->tw], h->rops_node, h->S.ncell);
  HIPOK(hipGetLastError());
  return do_cell_fields(h, false);
}

// deposition of the current positions into the fixed-point accumulators
int32_t deposit_current(mfixdem_handle* h) {
  DevParams P = h->P;
  const int t = h->tw, c = h->cur;
  const Removal rm{h->state[t], h->wmask[t], h->pos[c], h->vel[c], h->omega[c], h->radius[t],
                   h->flags};
  if (h->S.p.interp_scheme != MFIXDEM_INTERP_GARG_2012) {
    hipLaunchKernelGGL(k_deposit_filter, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, P,
                       nullptr, h->pos[h->cur], h->pvol[h->tw], h->phase[h->tw],
                       h->state[h->tw], h->fcell[h->tw], h->xe, h->yn, h->zt, h->csz,
                       h->fluid, h->rops_fix[h->fb], rm);
    HIPOK(hipGetLastError());
    return do_cell_fields(h, true);
  }
  hipLaunchKernelGGL(k_deposit_list, dim3(nblk(((int64_t)h->n + kDepPPT - 1) / kDepPPT)),
                     dim3(kBS), 0, h->stream, P, nullptr, h->pos[h->cur], h->pmass[h->tw],
                     h->phase[h->tw], h->state[h->tw], h->fcell[h->tw], h->fluid, h->xe, h->yn,
                     h->zt, h->rops_fix[h->fb], rm);
  HIPOK(hipGetLastError());
  return do_cell_fields(h, true);
}

// the gas-side tail of a step after the fused kernel (and any rebuild):
// PARTICLES_IN_CELL deletions (inside the deposition kernels, which visit
// every particle), deposition, ROP_S / EP_G
int32_t gas_tail(mfixdem_handle* h) {
  if (ref_deposition(h)) return ref_mean_fields(h);  // deletion inside do_pic_lists
  if (fused_dep(h)) return do_cell_fields(h, true);  // the step kernel deposited
  return deposit_current(h);
}

// PARTICLES_IN_CELL, weights and mean fields after every step
// (des_time_march.f:419-425): coupled runs, unless explicitly coupled
bool rebin_each_step(const mfixdem_handle* h) {
  return h->S.p.coupled && !h->S.p.des_explicitly_coupled;
}

// the part of a step that is identical between rebuilds (graph-capturable)
int32_t plain_step_body(mfixdem_handle* h) {
  CHECKRC(launch_step<kModeFused>(h));
  h->cur = 1 - h->cur;
  h->pcur = 1 - h->pcur;
  if (rebin_each_step(h)) CHECKRC(gas_tail(h));
  return MFIXDEM_OK;
}

// The free global ids, from the particle states (at init, and after rogue
// particles have been removed).
int32_t refresh_free_gids(mfixdem_handle* h) {
  const int n = h->n;
  hipLaunchKernelGGL(k_scatter_gid_u8, dim3(nblk(n)), dim3(kBS), 0, h->stream, n,
                     h->gid[h->tw], h->state[h->tw], (int8_t*)h->stage);
  HIPOK(hipGetLastError());
  std::vector<int8_t> st(n);
  HIPOK(hipMemcpyAsync(st.data(), h->stage, n, hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  h->dbc.free_gids.clear();
  for (int g = 0; g < n; g++)
    if (st[g] == kNonexistent) h->dbc.free_gids.insert(g);
  return MFIXDEM_OK;
}

// DES_TIME_STEP with DEM boundaries (des_time_march.f:354-380): the step
// kernel (with UPDATE_EXIT_DES_VEL), then MASS_INFLOW_DEM and
// MASS_OUTFLOW_DEM, and the neighbour search, which a deleted particle
// forces.  Launched directly (the seeds differ from step to step), with one
// wait for the count of deleted particles.
int32_t bc_step(mfixdem_handle* h, int32_t nn) {
  DemBoundaries& D = h->dbc;
  bool rebuild = (nn == 1) || (nn % h->S.p.neighbor_search_n == 0);
  const int n = h->n;
  t_begin(h);
  CHECKRC(launch_step<kModeFused>(h));
  t_end(h, kTimeStep);
  h->cur = 1 - h->cur;
  h->pcur = 1 - h->pcur;
  const int c = h->cur, t = h->tw;
  CHECKRC(fill_i32(h->stream, h->del_cnt, 1, 0));
  BcPtrs b{};
  b.pos = h->pos[c]; b.vel = h->vel[c]; b.omega = h->omega[c];
  b.radius = h->radius[t]; b.pmass = h->pmass[t]; b.omoi = h->omoi[t]; b.pvol = h->pvol[t];
  b.gid = h->gid[t]; b.dgc = h->dgc[t]; b.state = h->state[t]; b.wmask = h->wmask[t];
  b.phase = h->phase[t];
  b.bcs = h->d_bcs; b.nbc = (int)D.b.size(); b.n = n; b.S = h->stride;
  b.del_cnt = h->del_cnt; b.del_list = h->del_list;
  hipLaunchKernelGGL(k_dem_bcs, dim3(nblk(n)), dim3(kBS), 0, h->stream, b);
  HIPOK(hipGetLastError());
  CHECKRC(plan_seeds(h->S, D));
  const int ns = (int)D.seeds.size();
  if (ns > 0) {
    HIPOK(hipMemcpyAsync(h->d_seeds, D.seeds.data(), sizeof(SeedDev) * ns,
                         hipMemcpyHostToDevice, h->stream));
    SeedPtrs sp{};
    sp.pos = h->pos[c]; sp.vel = h->vel[c]; sp.omega = h->omega[c];
    sp.radius = h->radius[t]; sp.pmass = h->pmass[t]; sp.omoi = h->omoi[t];
    sp.pvol = h->pvol[t]; sp.dgc = h->dgc[t]; sp.fcell = h->fcell[t];
    sp.state = h->state[t]; sp.phase = h->phase[t]; sp.wmask = h->wmask[t];
    sp.slot_of_gid = h->slot_of_gid; sp.S = h->stride;
    hipLaunchKernelGGL(k_seed, dim3(nblk(ns)), dim3(kBS), 0, h->stream, ns, h->d_seeds, sp);
    HIPOK(hipGetLastError());
  }
  int nd = 0;
  HIPOK(hipMemcpyAsync(&nd, h->del_cnt, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));  // also frees the seed records for the next step
  if (nd > 0) {
    std::vector<int> gone(nd);
    HIPOK(hipMemcpyAsync(gone.data(), h->del_list, sizeof(int) * nd, hipMemcpyDeviceToHost,
                         h->stream));
    HIPOK(sync_stream(h));
    for (int g : gone) D.free_gids.insert(g);
    rebuild = true;  // FORCE_NSEARCH (mass_outflow_dem.f:178)
  }
  if (rebuild) {
    const int64_t rogue0 = h->st.rogue_deleted;
    t_begin(h);
    CHECKRC(do_bin(h));
    CHECKRC(do_neighbour(h));
    t_end(h, kTimeRebuild);
    if (h->st.rogue_deleted != rogue0) CHECKRC(refresh_free_gids(h));
  }
  D.s_time = D.s_time + h->S.d.dtsolid;
  h->first_pass = false;
  h->st.steps++;
  return MFIXDEM_OK;
}

int32_t one_step(mfixdem_handle* h, int32_t nn) {
  if (h->P.nbc) return bc_step(h, nn);
  const bool rebuild = (nn == 1) || (nn % h->S.p.neighbor_search_n == 0);
  const bool use_graph = !rebuild && !h->first_pass && !timing(h) &&
                         !(h->debug & (MFIXDEM_DEBUG_STORE_FORCES | MFIXDEM_DEBUG_NO_GRAPH));
  if (use_graph) {
    hipGraphExec_t& g = h->graph[h->cur][h->fb];
    if (!g) {
      const int c0 = h->cur, p0 = h->pcur, f0 = h->fb;
      HIPOK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      int32_t rc = plain_step_body(h);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(h->stream, &graph);
      h->cur = c0;
      h->pcur = p0;
      h->fb = f0;
      if (rc != MFIXDEM_OK) return rc;
      HIPOK(e);
      HIPOK(hipGraphInstantiate(&g, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
    }
    HIPOK(hipGraphLaunch(g, h->stream));
    h->cur = 1 - h->cur;
    h->pcur = 1 - h->pcur;
    if (h->S.p.coupled) h->fb = 1 - h->fb;
  } else {
    t_begin(h);
    CHECKRC(launch_step<kModeFused>(h));
    t_end(h, kTimeStep);
    h->cur = 1 - h->cur;
    h->pcur = 1 - h->pcur;
    if (rebuild) {
      t_begin(h);
      if (fused_dep(h)) {  // the rebuild reorders the slots the removal queue names
        hipLaunchKernelGGL(k_apply_removals, dim3(1), dim3(kBS), 0, h->stream, h->rm_cnt + h->fb,
                           h->rm_list, h->state[h->tw], h->radius[h->tw]);
        HIPOK(hipGetLastError());
      }
      CHECKRC(do_bin(h));
      CHECKRC(do_neighbour(h));
      t_end(h, kTimeRebuild);
    }
    if (rebin_each_step(h)) {
      t_begin(h);
      CHECKRC(gas_tail(h));
      t_end(h, kTimeGas);
    }
  }
  h->first_pass = false;
  h->st.steps++;
  return MFIXDEM_OK;
}

int32_t scatter_d3(mfixdem_handle* h, const double* src, double* host, int64_t ld) {
  const int n = h->n;
  hipLaunchKernelGGL(k_scatter_gid_d3, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->stride, n,
                     h->gid[h->tw], h->state[h->tw], src, h->stage);
  HIPOK(hipGetLastError());
  std::vector<double> tmp((size_t)3 * n);
  HIPOK(hipMemcpyAsync(tmp.data(), h->stage, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                       h->stream));
  HIPOK(sync_stream(h));
  for (int c = 0; c < 3; c++) std::memcpy(host + c * ld, tmp.data() + (size_t)c * n, sizeof(double) * n);
  return MFIXDEM_OK;
}

template <typename T>
int32_t download(mfixdem_handle* h, std::vector<T>& v, const T* d, int64_t count) {
  v.resize((size_t)std::max<int64_t>(count, 0));
  if (count > 0)
    HIPOK(hipMemcpyAsync(v.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  return MFIXDEM_OK;
}

// The walls of one or more set-ups (an ensemble's systems in order) to the
// device, the systems' DES cells concatenated as the ensemble's DES-cell
// offsets number them: for a system of at most kWallSlots facets a facet
// mask per DES cell (its own facet ids; the facets travel in the kernel
// arguments), otherwise the facet table (system s's ids offset by the
// facets of the systems before it) and one facet list per DES cell.
int32_t upload_facets(mfixdem_handle* h, const std::vector<const Setup*>& sys) {
  std::vector<FacetDev> fac, one;
  std::vector<int> foff(1, 0), fids;
  std::vector<uint8_t> fmask;
  for (const Setup* S : sys) {
    const int base = (int)fac.size();
    facets_dev(*S, one);
    fac.insert(fac.end(), one.begin(), one.end());
    const bool small = walls_small(*S);
    for (int c = 0; c < S->dg_ncell; c++) {
      uint8_t m = 0;
      for (int q = S->dg_foff[c]; q < S->dg_foff[c + 1]; q++) {
        fids.push_back(base + S->dg_fids[q]);
        if (small) m |= (uint8_t)(1u << S->dg_fids[q]);
      }
      foff.push_back((int)fids.size());
      fmask.push_back(m);
    }
  }
  dfree(h->fac);
  dfree(h->dg_foff);
  dfree(h->dg_fids);
  dfree(h->cell_fmask);
  int32_t rc;
  if ((rc = dalloc(&h->fac, (int64_t)fac.size())) || (rc = dalloc(&h->dg_foff, (int64_t)foff.size())) ||
      (rc = dalloc(&h->dg_fids, (int64_t)fids.size())) ||
      (rc = dalloc(&h->cell_fmask, (int64_t)fmask.size())))
    return rc;
  if (!fmask.empty())
    HIPOK(hipMemcpy(h->cell_fmask, fmask.data(), fmask.size(), hipMemcpyHostToDevice));
  if (!fac.empty())
    HIPOK(hipMemcpy(h->fac, fac.data(), sizeof(FacetDev) * fac.size(), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->dg_foff, foff.data(), sizeof(int) * foff.size(), hipMemcpyHostToDevice));
  if (!fids.empty())
    HIPOK(hipMemcpy(h->dg_fids, fids.data(), sizeof(int) * fids.size(), hipMemcpyHostToDevice));
  return MFIXDEM_OK;
}
int32_t upload_facets(mfixdem_handle* h, const Setup& S) {
  return upload_facets(h, std::vector<const Setup*>{&S});
}

int32_t upload_setup(mfixdem_handle* h) {
  Setup& S = h->S;
  HIPOK(hipMemcpy(h->xe, S.xe.data(), sizeof(double) * S.xe.size(), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->yn, S.yn.data(), sizeof(double) * S.yn.size(), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->zt, S.zt.data(), sizeof(double) * S.zt.size(), hipMemcpyHostToDevice));
  std::vector<double> csz(S.cdx);
  csz.insert(csz.end(), S.cdy.begin(), S.cdy.end());
  csz.insert(csz.end(), S.cdz.begin(), S.cdz.end());
  HIPOK(hipMemcpy(h->csz, csz.data(), sizeof(double) * csz.size(), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->vol, S.vol.data(), sizeof(double) * S.ncell, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->vol_surr, S.vol_surr.data(), sizeof(double) * S.ncell, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->fluid, S.fluid.data(), S.ncell, hipMemcpyHostToDevice));
  CHECKRC(upload_facets(h, S));
  double ros[MFIXDEM_MAX_PHASES];
  for (int m = 0; m < MFIXDEM_MAX_PHASES; m++) ros[m] = S.p.ro_s0[m];
  HIPOK(hipMemcpy(h->ro_s, ros, sizeof(ros), hipMemcpyHostToDevice));
  std::vector<double> one(S.ncell, 1.0);
  HIPOK(hipMemcpy(h->ep_g, one.data(), sizeof(double) * S.ncell, hipMemcpyHostToDevice));
  return MFIXDEM_OK;
}

// Upload an ensemble's staged particles: per system, ids 0..capacity-1 in
// slot order, PVOL/PMASS/OMOI as make_arrays_des.f:288-291, NORMAL for the
// particles set, NONEXISTENT for the rest of the system's slots.
int32_t upload_ensemble(mfixdem_handle* h) {
  const int64_t N = h->n, S = h->stride;
  std::vector<double> R(S, 0.0), M(S, 3.4028234663852886e38), OI(S, 0.0),
      PV(S, 3.4028234663852886e38);
  std::vector<int> G(S, 0), FC(S, -1), DG(S, 0);
  std::vector<uint8_t> ST(S, 0), PH(S, 0), WM(S, 0);
  for (int s = 0; s < h->nsys; s++) {
    const int off = h->sys_off[s];
    for (int q = off; q < h->sys_off[s + 1]; q++) G[q] = q - off;
    for (int q = off; q < off + h->sys_n[s]; q++) {
      const double r = h->st_r[q];
      R[q] = r;
      PV[q] = (4.0 / 3.0) * kPi * (r * r * r);
      M[q] = PV[q] * h->st_rho[q];
      OI[q] = 2.5 / (M[q] * (r * r));
      PH[q] = (uint8_t)(h->st_phase[q] - 1);
      ST[q] = 1;
    }
  }
  invalidate_graphs(h);
  h->cur = h->pcur = h->tw = h->lt = 0;
  CHECKRC(reset_deposit(h));
  HIPOK(hipMemcpy(h->pos[0], h->st_pos.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->vel[0], h->st_vel.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->omega[0], h->st_om.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->radius[0], R.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->pmass[0], M.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->omoi[0], OI.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->pvol[0], PV.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->gid[0], G.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->fcell[0], FC.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->dgc[0], DG.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->state[0], ST.data(), S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->phase[0], PH.data(), S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->wmask[0], WM.data(), S, hipMemcpyHostToDevice));
  std::vector<int> Z((size_t)N + 1, 0);
  for (int b = 0; b < 2; b++) {
    HIPOK(hipMemcpy(h->nbr_start[b], Z.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->nbr_split[b], Z.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->pair_base[b], Z.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  }
  HIPOK(hipMemsetAsync(h->flags, 0, sizeof(int) * kFlagCount, h->stream));
  h->flags_set = false;
  h->st = mfixdem_status{};
  h->n_pairs = 0;
  h->first_pass = true;
  h->binned = false;
  return MFIXDEM_OK;
}

// Calls f(ptr) for every device buffer a handle owns (the single list that
// destruction and the footprint query share).
template <class F>
void for_each_buffer(mfixdem_handle* h, F f) {
  for (int b = 0; b < 2; b++) {
    f(h->pos[b]); f(h->vel[b]); f(h->omega[b]); f(h->radius[b]);
    f(h->pmass[b]); f(h->omoi[b]); f(h->pvol[b]); f(h->acc[b]);
    f(h->racc[b]); f(h->gid[b]); f(h->fcell[b]); f(h->dgc[b]);
    f(h->state[b]); f(h->phase[b]); f(h->wmask[b]); f(h->wfid[b]); f(h->wpft[b]);
    f(h->keys[b]); f(h->vals[b]); f(h->nbr_start[b]); f(h->nbr_split[b]);
    f(h->pair_base[b]); f(h->nbr[b]); f(h->nbr_pid[b]); f(h->pft[b]); f(h->pfl[b]);
    f(h->rops_fix[b]);
  }
  f(h->fc); f(h->tow); f(h->fc_id); f(h->tow_id); f(h->sort_tmp); f(h->rm_cnt); f(h->rm_list); f(h->f_gp); f(h->drag_fc);
  f(h->xr_pack);
  f(h->fstart); f(h->fend); f(h->fine_rank); f(h->fine_cnt); f(h->fine_off);
  f(h->own_cnt); f(h->own_tmp); f(h->own_key); f(h->in_cnt); f(h->in_cur);
  f(h->tot_cnt); f(h->scan_bsum); f(h->scan_bsum64); f(h->flags); f(h->csz); f(h->xe); f(h->yn);
  f(h->zt); f(h->vnode); f(h->ugc); f(h->gas_stage); f(h->p_force); f(h->ep_g); f(h->rop_s);
  f(h->rops_node); f(h->vol); f(h->vol_surr); f(h->ro_s); f(h->fluid);
  f(h->pic_cnt); f(h->pic_off); f(h->pic_cur); f(h->pic);
  f(h->gd_fgp); f(h->gd_vol_node); f(h->gd_am); f(h->gd_bm); f(h->gd_fgds);
  f(h->fac); f(h->dg_foff); f(h->dg_fids); f(h->cell_fmask); f(h->stage);
  f(h->d_sys); f(h->d_ptab); f(h->d_gtab); f(h->d_cell_off); f(h->d_dg_off); f(h->d_rogue);
  f(h->d_bcs); f(h->d_seeds); f(h->slot_of_gid); f(h->del_cnt); f(h->del_list);
}

void fill_devparams(mfixdem_handle* h) {
  fill_params(h->S, h->n, h->stride, h->P, h->G);
  h->P.nbc = (int)h->dbc.b.size();
}

// Search cells of one system for the counting sort: the rank of each cell
// (x fastest) in the Morton order of its (fx, fy, fz), so that the binned
// particles lie in Morton order; zeroed counts.
int32_t setup_fine_cells(mfixdem_handle* h) {
  dfree(h->fine_rank);
  dfree(h->fine_cnt);
  dfree(h->fine_off);
  if (h->nsys > 1) return MFIXDEM_OK;
  const GridDev& G = h->G;
  const int nf = h->n_fine;
  if ((int64_t)G.fn[0] * G.fn[1] * G.fn[2] != nf) return MFIXDEM_ERR_ARG;
  auto spread = [](uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 10; b++) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, int>> code((size_t)nf);
  for (int z = 0, c = 0; z < G.fn[2]; z++)
    for (int y = 0; y < G.fn[1]; y++)
      for (int x = 0; x < G.fn[0]; x++, c++)
        code[c] = {spread(x) | (spread(y) << 1) | (spread(z) << 2), c};
  std::sort(code.begin(), code.end());
  std::vector<int> rank((size_t)nf);
  for (int k = 0; k < nf; k++) rank[code[k].second] = k;
  CHECKRC(dalloc(&h->fine_rank, nf));
  CHECKRC(dalloc(&h->fine_cnt, (int64_t)nf + 1));
  CHECKRC(dalloc(&h->fine_off, (int64_t)nf + 2));
  HIPOK(hipMemcpy(h->fine_rank, rank.data(), sizeof(int) * nf, hipMemcpyHostToDevice));
  HIPOK(hipMemset(h->fine_cnt, 0, sizeof(int) * ((size_t)nf + 1)));
  return MFIXDEM_OK;
}

// Device buffers, stream and events of a handle whose set-up (S, P, G,
// n, stride, n_fine, key_bits) is filled in.  On failure the caller
// destroys the handle.
int32_t alloc_buffers(mfixdem_handle* h) {
  const int64_t n = h->n, S = h->stride, nc = h->S.ncell;
  int32_t rc = MFIXDEM_OK;
  auto fail = [&](int32_t r) { return r; };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(MFIXDEM_ERR_HIP);
  for (int b = 0; b < 2; b++) {
    if ((rc = dalloc(&h->pos[b], 3 * S)) || (rc = dalloc(&h->vel[b], 3 * S)) ||
        (rc = dalloc(&h->omega[b], 3 * S)) || (rc = dalloc(&h->radius[b], S)) ||
        (rc = dalloc(&h->pmass[b], S)) || (rc = dalloc(&h->omoi[b], S)) ||
        (rc = dalloc(&h->pvol[b], S)) || (rc = dalloc(&h->acc[b], 3 * S)) ||
        (rc = dalloc(&h->racc[b], 3 * S)) || (rc = dalloc(&h->gid[b], S)) ||
        (rc = dalloc(&h->fcell[b], S)) || (rc = dalloc(&h->dgc[b], S)) ||
        (rc = dalloc(&h->state[b], S)) || (rc = dalloc(&h->phase[b], S)) ||
        (rc = dalloc(&h->wmask[b], S)) || (rc = dalloc(&h->wfid[b], kWallSlots * S)) ||
        (rc = dalloc(&h->wpft[b], 3 * kWallSlots * S)) ||
        (rc = dalloc(&h->keys[b], n)) || (rc = dalloc(&h->vals[b], n)) ||
        (rc = dalloc(&h->nbr_start[b], n + 1)) || (rc = dalloc(&h->nbr_split[b], n)) ||
        (rc = dalloc(&h->pair_base[b], n + 1)))
      return fail(rc);
  }
  if ((rc = dalloc(&h->fc, 3 * S)) || (rc = dalloc(&h->tow, 3 * S)) ||
      (rc = dalloc(&h->xr_pack, 4 * S)) ||
      (rc = dalloc(&h->fc_id, 3 * S)) || (rc = dalloc(&h->tow_id, 3 * S)) ||
      (rc = dalloc(&h->fstart, h->n_fine)) || (rc = dalloc(&h->fend, h->n_fine)) ||
      (rc = dalloc(&h->own_cnt, n)) || (rc = dalloc(&h->own_tmp, n * h->maxrow)) ||
      (rc = dalloc(&h->own_key, n * h->maxrow)) || (rc = dalloc(&h->in_cnt, n)) ||
      (rc = dalloc(&h->in_cur, n)) || (rc = dalloc(&h->tot_cnt, n)) ||
      (rc = dalloc(&h->scan_bsum64, n / kScanTile + 2)) ||
      (rc = dalloc(&h->scan_bsum,
                   std::max<int64_t>(std::max<int64_t>(n, nc), (int64_t)h->n_fine + 1) /
                           kScanTile + 2)) ||
      (rc = dalloc(&h->flags, kFlagCount)) || (rc = dalloc(&h->xe, h->S.ni + 1)) ||
      (rc = dalloc(&h->csz, h->S.ni + h->S.nj + h->S.nk)) ||
      (rc = dalloc(&h->yn, h->S.nj + 1)) || (rc = dalloc(&h->zt, h->S.nk + 1)) ||
      (rc = dalloc(&h->vnode, 3 * nc)) || (rc = dalloc(&h->p_force, 3 * nc)) ||
      (rc = dalloc(&h->ugc, 3 * nc)) || (rc = dalloc(&h->gas_stage, 3 * nc)) ||
      (rc = dalloc(&h->ep_g, nc)) || (rc = dalloc(&h->rop_s, nc * h->S.p.nphases)) ||
      (rc = dalloc(&h->rops_node, nc * h->S.p.nphases)) || (rc = dalloc(&h->vol, nc)) ||
      (rc = dalloc(&h->vol_surr, nc)) || (rc = dalloc(&h->ro_s, MFIXDEM_MAX_PHASES)) ||
      (rc = dalloc(&h->fluid, nc)) || (rc = dalloc(&h->pic_cnt, nc)) ||
      (rc = dalloc(&h->pic_off, nc + 1)) || (rc = dalloc(&h->pic_cur, nc)) ||
      (rc = dalloc(&h->pic, n)) ||
      (rc = dalloc(&h->stage, 3 * n)) ||
      (rc = dalloc(&h->rops_fix[0], nc * h->S.p.nphases)) ||
      (rc = dalloc(&h->rops_fix[1], nc * h->S.p.nphases)) || (rc = dalloc(&h->rm_cnt, 2)) ||
      (rc = dalloc(&h->rm_list, n)))
    return fail(rc);
  for (int b = 0; b < 2; b++)
    if (hipMemset(h->rops_fix[b], 0, sizeof(unsigned long long) * nc * h->S.p.nphases) != hipSuccess)
      return fail(MFIXDEM_ERR_HIP);
  if (hipMemset(h->rm_cnt, 0, 2 * sizeof(int)) != hipSuccess) return fail(MFIXDEM_ERR_HIP);
  if (h->S.p.des_explicitly_coupled) {
    if ((rc = dalloc(&h->f_gp, n)) || (rc = dalloc(&h->drag_fc, 3 * n))) return fail(rc);
    if (hipMemset(h->f_gp, 0, sizeof(double) * n) != hipSuccess ||
        hipMemset(h->drag_fc, 0, sizeof(double) * 3 * n) != hipSuccess)
      return fail(MFIXDEM_ERR_HIP);
  }
  if (sort_pairs(nullptr, h->sort_tmp_bytes, h->keys[0], h->keys[1], h->vals[0], h->vals[1],
                 (int)n, h->key_bits, h->stream) != hipSuccess)
    return fail(MFIXDEM_ERR_HIP);
  if (hipMalloc(&h->sort_tmp, std::max<size_t>(h->sort_tmp_bytes, 16)) != hipSuccess)
    return fail(MFIXDEM_ERR_OOM);
  if ((rc = upload_setup(h)) || (rc = setup_fine_cells(h))) return fail(rc);
  if (hipMemset(h->flags, 0, sizeof(int) * kFlagCount) != hipSuccess) return fail(MFIXDEM_ERR_HIP);
  return MFIXDEM_OK;
}



}  // namespace

// ===================================================================== ABI
extern "C" {

int32_t mfixdem_abi_version(void) { return MFIXDEM_ABI_VERSION; }

const char* mfixdem_error_string(int32_t code) {
  switch (code) {
    case MFIXDEM_OK: return "ok";
    case MFIXDEM_ERR_ARG: return "invalid argument";
    case MFIXDEM_ERR_HIP: return "HIP runtime error";
    case MFIXDEM_ERR_OOM: return "device allocation failed";
    case MFIXDEM_ERR_ROGUE: return "normal particle found in a ghost DES cell";
    case MFIXDEM_ERR_EPG: return "unphysical gas volume fraction";
    case MFIXDEM_ERR_ZERO_DISTANCE: return "distance between particles is zero";
    case MFIXDEM_ERR_CAPACITY: return "buffer too small";
    case MFIXDEM_ERR_NOT_READY: return "handle not initialised";
    case MFIXDEM_ERR_UNSUPPORTED: return "unsupported option";
    case MFIXDEM_ERR_SCRATCH: return "input exceeds the fixed device scratch of a no-allocation call";
    default: return "unknown error";
  }
}

int32_t mfixdem_params_default(mfixdem_params* p) {
  if (!p) return MFIXDEM_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->abi_version = MFIXDEM_ABI_VERSION;
  p->coll_model = MFIXDEM_LSD;
  p->integrator = MFIXDEM_EULER;
  p->precision = MFIXDEM_FP64;
  p->nphases = 1;
  p->neighbor_search_n = 25;
  p->drag_type = MFIXDEM_DRAG_GIDASPOW;
  p->interp_scheme = MFIXDEM_INTERP_GARG_2012;
  p->imax = p->jmax = p->kmax = 1;
  p->remove_rogue = 1;
  p->dtsolid = kUndef;
  p->dtsolid_factor = 50.0;
  p->grav[1] = -9.80665;
  p->factor_rlm = (double)1.2f;
  p->kt_fac = (double)0.28571428571429f;
  p->kt_w_fac = (double)0.28571428571429f;
  p->des_etat_fac = kUndef;
  p->des_etat_w_fac = kUndef;
  for (int i = 0; i < MFIXDEM_MAX_PAIRS_IN; i++) p->des_en_input[i] = p->des_et_input[i] = kUndef;
  for (int i = 0; i < MFIXDEM_MAX_PHASES; i++) {
    p->des_en_wall_input[i] = p->des_et_wall_input[i] = kUndef;
    p->e_young[i] = p->v_poisson[i] = p->d_p0[i] = p->ro_s0[i] = kUndef;
  }
  p->ew_young = p->vw_poisson = kUndef;
  p->des_epg_clip = kUndef;
  p->des_interp_width = kUndef;
  p->des_neighbor_search = MFIXDEM_NSEARCH_GRID;
  p->ro_g0 = 1.2;
  p->mu_g0 = 1.8e-5;
  return MFIXDEM_OK;
}

// Device address of `scratch` on each device, resolved once per device with
// that device current (each GPU holds its own copy at its own address).
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;
int32_t bind_scratch(mfixdem_handle* h) {
  int dev = 0;
  HIPOK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HIPOK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return MFIXDEM_ERR_HIP;
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  h->device = dev;
  h->scratch_ptr = g_scratch_dev[dev];
  return MFIXDEM_OK;
}

int32_t mfixdem_create(const mfixdem_params* p, mfixdem_handle** out) {
  if (!p || !out) return MFIXDEM_ERR_ARG;
  *out = nullptr;
  mfixdem_handle* h = new (std::nothrow) mfixdem_handle();
  if (!h) return MFIXDEM_ERR_OOM;
  int32_t rc = build_setup(*p, h->S);
  if (rc != MFIXDEM_OK) {
    delete h;
    return rc;
  }
  h->n = p->max_pip;
  h->stride = (h->n + 63) / 64 * 64;
  fill_devparams(h);
  if (search_grid(*p, h->n, h->G, h->n_fine) != MFIXDEM_OK) {
    delete h;
    return MFIXDEM_ERR_UNSUPPORTED;
  }
  h->key_bits = h->G.mbits;
  if ((rc = alloc_buffers(h)) || (rc = bind_scratch(h))) {
    mfixdem_destroy(h);
    return rc;
  }
  *out = h;
  return MFIXDEM_OK;
}

int32_t mfixdem_destroy(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (h->stream) (void)sync_stream(h);
  invalidate_graphs(h);
  reap_graphs(h);
  for_each_buffer(h, [](auto*& p) { dfree(p); });
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return MFIXDEM_OK;
}

int32_t mfixdem_get_derived(const mfixdem_handle* h, mfixdem_derived* out) {
  if (!h || !out) return MFIXDEM_ERR_ARG;
  *out = h->S.d;
  return MFIXDEM_OK;
}

int32_t mfixdem_set_particles(mfixdem_handle* h, int64_t n, const double* pos, const double* vel,
                              const double* omega, const double* radius, const double* ro_sol,
                              const int32_t* phase, int64_t ld) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || n < 0 || n > h->n || ld < n || !pos || !radius || !ro_sol) return MFIXDEM_ERR_ARG;
  const int64_t N = h->n, S = h->stride;
  std::vector<double> P3((size_t)3 * S, 0.0), V3((size_t)3 * S, 0.0), W3((size_t)3 * S, 0.0);
  std::vector<double> R(S, 0.0), M(S, 3.4028234663852886e38), OI(S, 0.0), PV(S, 3.4028234663852886e38);
  std::vector<int> G(S, 0), FC(S, -1), DG(S, 0);
  std::vector<uint8_t> ST(S, 0), PH(S, 0), WM(S, 0);
  for (int64_t q = 0; q < N; q++) G[q] = (int)q;
  for (int64_t q = 0; q < n; q++) {
    for (int c = 0; c < 3; c++) {
      P3[c * S + q] = pos[c * ld + q];
      V3[c * S + q] = vel ? vel[c * ld + q] : 0.0;
      W3[c * S + q] = omega ? omega[c * ld + q] : 0.0;
    }
    const double r = radius[q];
    R[q] = r;
    // make_arrays_des.f:288-291
    PV[q] = (4.0 / 3.0) * kPi * (r * r * r);
    M[q] = PV[q] * ro_sol[q];
    OI[q] = 2.5 / (M[q] * (r * r));
    const int ph = phase ? phase[q] - 1 : 0;
    if (ph < 0 || ph >= h->S.p.nphases) return MFIXDEM_ERR_ARG;
    PH[q] = (uint8_t)ph;
    ST[q] = 1;
  }
  // Fixed-point scale per phase: the mass a node can collect is bounded by
  // the solids in its 8 cells, <= 8 * VOL * rho_max; allow twice that and
  // keep the int64 sum below 2^62.
  {
    double rho_max[MFIXDEM_MAX_PHASES] = {0, 0, 0, 0};
    for (int64_t q = 0; q < n; q++) rho_max[PH[q]] = std::max(rho_max[PH[q]], ro_sol[q]);
    const double vcell = h->S.vol_max;
    // the filter schemes sum solids volume into cells: bounded by the
    // volume of the 27 cells a weight can come from
    const bool vols = h->S.p.interp_scheme != MFIXDEM_INTERP_GARG_2012;
    for (int m = 0; m < MFIXDEM_MAX_PHASES; m++) {
      const double bound = vols ? 32.0 * vcell : 16.0 * vcell * std::max(rho_max[m], 1.0);
      const int e = (int)std::floor(std::log2(std::ldexp(1.0, 62) / bound));
      h->P.dep_scale[m] = std::ldexp(1.0, e);
      h->dep_inv[m] = std::ldexp(1.0, -e);
    }
  }
  HIPOK(sync_stream(h));  // steps already queued still use the buffers
  invalidate_graphs(h);
  h->cur = h->pcur = h->tw = h->lt = 0;
  CHECKRC(reset_deposit(h));
  HIPOK(hipMemcpy(h->pos[0], P3.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->vel[0], V3.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->omega[0], W3.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->radius[0], R.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->pmass[0], M.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->omoi[0], OI.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->pvol[0], PV.data(), sizeof(double) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->gid[0], G.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->fcell[0], FC.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->dgc[0], DG.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->state[0], ST.data(), S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->phase[0], PH.data(), S, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->wmask[0], WM.data(), S, hipMemcpyHostToDevice));
  std::vector<int> Z((size_t)N + 1, 0);
  for (int b = 0; b < 2; b++) {
    HIPOK(hipMemcpy(h->nbr_start[b], Z.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->nbr_split[b], Z.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->pair_base[b], Z.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
  }
  // on the handle's stream, which every later call on the handle waits for
  HIPOK(hipMemsetAsync(h->flags, 0, sizeof(int) * kFlagCount, h->stream));
  h->flags_set = false;
  HIPOK(hipMemsetAsync(h->fc_id, 0, sizeof(double) * 3 * S, h->stream));
  HIPOK(hipMemsetAsync(h->tow_id, 0, sizeof(double) * 3 * S, h->stream));
  if (h->f_gp) {  // des_init_arrays.f:143
    HIPOK(hipMemsetAsync(h->f_gp, 0, sizeof(double) * N, h->stream));
    HIPOK(hipMemsetAsync(h->drag_fc, 0, sizeof(double) * 3 * N, h->stream));
  }
  h->st = mfixdem_status{};
  h->n_pairs = 0;
  h->inited = false;
  h->first_pass = true;
  h->binned = false;
  h->order_set = false;  // the first binning puts the new particles in order
  h->xr_packed = h->counts_zeroed = false;
  h->bin_idx = h->bin_inv = nullptr;
  h->bins_since_permute = 0;
  return MFIXDEM_OK;
}

int32_t mfixdem_set_particle_state(mfixdem_handle* h, const int8_t* state, int64_t n) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || !state || n < 0 || n > h->n || h->inited) return MFIXDEM_ERR_ARG;
  h->xr_packed = false;
  for (int64_t q = 0; q < n; q++)
    if (state[q] != 0 && state[q] != 1) return MFIXDEM_ERR_UNSUPPORTED;
  if (n == 0) return MFIXDEM_OK;
  HIPOK(sync_stream(h));
  HIPOK(hipMemcpy(h->stage, state, (size_t)n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_set_holes, dim3(nblk(n)), dim3(kBS), 0, h->stream, (int)n,
                     (const int8_t*)h->stage, h->state[0], h->radius[0], h->stride, h->pos[0],
                     h->vel[0], h->omega[0]);
  HIPOK(hipGetLastError());
  HIPOK(sync_stream(h));
  return MFIXDEM_OK;
}

int32_t mfixdem_set_facets(mfixdem_handle* h, int32_t nf, const double* vertex,
                           const double* normal, int32_t default_walls) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || nf < 0 || (nf > 0 && (!vertex || !normal)) || h->inited) return MFIXDEM_ERR_ARG;
  for (int64_t q = 0; q < 3 * (int64_t)nf; q++)
    if (!std::isfinite(normal[q])) return MFIXDEM_ERR_ARG;
  HIPOK(sync_stream(h));
  setup_facets(h->S, nf, vertex, normal, default_walls != 0);
  small_facets(h->S, h->P);
  CHECKRC(upload_facets(h, h->S));
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

int32_t mfixdem_set_wall_motion(mfixdem_handle* h, int32_t first, int32_t count,
                                const double* vel, const double* omega,
                                const double* rot_center) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || !vel || !omega || !rot_center || first < 0 || count < 0 ||
      (int64_t)first + count > (int64_t)h->S.facets.size())
    return MFIXDEM_ERR_ARG;
  HIPOK(sync_stream(h));
  set_wall_motion(h->S, first, count, vel, omega, rot_center);
  small_facets(h->S, h->P);  // facets in the kernel arguments: graphs recorded again
  invalidate_graphs(h);
  std::vector<FacetDev> fac;
  facets_dev(h->S, fac);
  if (!fac.empty())
    HIPOK(hipMemcpy(h->fac, fac.data(), sizeof(FacetDev) * fac.size(), hipMemcpyHostToDevice));
  return MFIXDEM_OK;
}

int32_t mfixdem_set_dem_bcs(mfixdem_handle* h, int32_t n, const mfixdem_dem_bc* bcs,
                            uint64_t seed) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || n < 0 || (n > 0 && !bcs) || h->inited) return MFIXDEM_ERR_ARG;
  const mfixdem_params& p = h->S.p;
  // Euler steps of an uncoupled system on the DES grid search
  if (n > 0 && (p.integrator != MFIXDEM_EULER || p.coupled ||
                p.des_neighbor_search == MFIXDEM_NSEARCH_NSQUARE))
    return MFIXDEM_ERR_UNSUPPORTED;
  for (int k = 0; k < n; k++) {
    if (bcs[k].type < MFIXDEM_BC_MASS_INFLOW || bcs[k].type > MFIXDEM_BC_OUTFLOW)
      return MFIXDEM_ERR_ARG;
    if (bcs[k].plane < MFIXDEM_PLANE_W || bcs[k].plane > MFIXDEM_PLANE_T) return MFIXDEM_ERR_ARG;
  }
  DemBoundaries D;
  D.b.resize(n);
  for (int k = 0; k < n; k++) D.b[k].in = bcs[k];
  D.rng = seed;
  CHECKRC(setup_dem_bcs(h->S, D));
  HIPOK(sync_stream(h));
  for (auto** q : {&h->slot_of_gid, &h->del_cnt, &h->del_list}) dfree(*q);
  dfree(h->d_bcs);
  dfree(h->d_seeds);
  h->dbc = DemBoundaries{};
  h->P.nbc = 0;
  invalidate_graphs(h);
  if (n == 0) return MFIXDEM_OK;
  const int64_t N = h->n;
  CHECKRC(dalloc(&h->d_bcs, n));
  CHECKRC(dalloc(&h->d_seeds, std::max(D.max_seeds, 1)));
  CHECKRC(dalloc(&h->slot_of_gid, N));
  CHECKRC(dalloc(&h->del_cnt, 1));
  CHECKRC(dalloc(&h->del_list, N));
  const std::vector<BcDev> bd = bcs_dev(h->S, D);
  HIPOK(hipMemcpy(h->d_bcs, bd.data(), sizeof(BcDev) * n, hipMemcpyHostToDevice));
  h->dbc = std::move(D);
  h->P.nbc = n;
  return MFIXDEM_OK;
}

int32_t mfixdem_get_particles(mfixdem_handle* h, double* pos, double* vel, double* omega,
                              int8_t* state, int64_t ld) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || ld < h->n) return MFIXDEM_ERR_ARG;
  if (pos) CHECKRC(scatter_d3(h, h->pos[h->cur], pos, ld));
  if (vel) CHECKRC(scatter_d3(h, h->vel[h->cur], vel, ld));
  if (omega) CHECKRC(scatter_d3(h, h->omega[h->cur], omega, ld));
  if (state) {
    hipLaunchKernelGGL(k_scatter_gid_u8, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->n,
                       h->gid[h->tw], h->state[h->tw], (int8_t*)h->stage);
    HIPOK(hipGetLastError());
    HIPOK(hipMemcpyAsync(state, h->stage, h->n, hipMemcpyDeviceToHost, h->stream));
    HIPOK(sync_stream(h));
  }
  return MFIXDEM_OK;
}

int32_t mfixdem_set_gas(mfixdem_handle* h, const double* u_g, const double* v_g,
                        const double* w_g, const double* p_force) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || !u_g || !v_g || !w_g) return MFIXDEM_ERR_ARG;
  const Setup& S = h->S;
  const mfixdem_params& p = S.p;
  const int nc = S.ncell;
  // steps already queued may still read the current fields
  HIPOK(sync_stream(h));
  double* st = h->gas_stage;
  HIPOK(hipMemcpyAsync(st, u_g, sizeof(double) * nc, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(st + nc, v_g, sizeof(double) * nc, hipMemcpyHostToDevice, h->stream));
  HIPOK(hipMemcpyAsync(st + 2 * (size_t)nc, w_g, sizeof(double) * nc, hipMemcpyHostToDevice,
                       h->stream));
  // node velocities (GARG_2012) and cell-centre velocities (filter schemes);
  // the field is frozen until the next call, so they are formed once here
  hipLaunchKernelGGL(k_gas_velocities, dim3(nblk(nc)), dim3(kBS), 0, h->stream, S.ni, S.nj, S.nk,
                     p.imax, p.jmax, p.kmax, h->fluid, st, st + nc, st + 2 * (size_t)nc,
                     h->vnode, h->ugc);
  HIPOK(hipGetLastError());
  // P_FORCE after the velocities have been formed (the stage is reused)
  if (p_force)
    HIPOK(hipMemcpyAsync(st, p_force, sizeof(double) * 3 * nc, hipMemcpyHostToDevice,
                         h->stream));
  hipLaunchKernelGGL(k_copy_or_zero_f64, dim3(nblk(3 * (int64_t)nc)), dim3(kBS), 0, h->stream,
                     h->p_force, p_force ? st : nullptr, 3 * nc);
  HIPOK(hipGetLastError());
  HIPOK(sync_stream(h));  // the caller's arrays are free again on return
  h->gas_set = true;
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

int32_t mfixdem_set_gas_pressure(mfixdem_handle* h, const double* p_g) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || !p_g) return MFIXDEM_ERR_ARG;
  const Setup& S = h->S;
  const int nc = S.ncell;
  HIPOK(sync_stream(h));  // queued steps may still read P_FORCE
  HIPOK(hipMemcpyAsync(h->gas_stage, p_g, sizeof(double) * nc, hipMemcpyHostToDevice,
                       h->stream));
  hipLaunchKernelGGL(k_pg_grad, dim3(nblk(nc)), dim3(kBS), 0, h->stream, S.ni, S.nj, S.nk, h->csz,
                     h->fluid, h->gas_stage, h->p_force);
  HIPOK(hipGetLastError());
  HIPOK(sync_stream(h));
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

int32_t mfixdem_get_p_force(mfixdem_handle* h, double* p_force) {
  if (!h || !p_force) return MFIXDEM_ERR_ARG;
  HIPOK(hipMemcpyAsync(p_force, h->p_force, sizeof(double) * 3 * h->S.ncell,
                       hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  return MFIXDEM_OK;
}

int32_t mfixdem_get_gas(mfixdem_handle* h, double* ep_g, double* rop_s) {
  if (!h) return MFIXDEM_ERR_ARG;
  const int nc = h->S.ncell;
  if (ep_g) HIPOK(hipMemcpyAsync(ep_g, h->ep_g, sizeof(double) * nc, hipMemcpyDeviceToHost, h->stream));
  if (rop_s)
    HIPOK(hipMemcpyAsync(rop_s, h->rop_s, sizeof(double) * nc * h->S.p.nphases,
                         hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  return MFIXDEM_OK;
}

int32_t mfixdem_init(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (h->S.p.coupled && !h->gas_set) return MFIXDEM_ERR_NOT_READY;
  HIPOK(sync_stream(h));
  if (h->nsys > 1) CHECKRC(upload_ensemble(h));
  const int n = h->n;
  HIPOK(hipMemsetAsync(h->flags, 0, sizeof(int) * kFlagCount, h->stream));
  h->flags_set = false;
  if (h->S.p.coupled) {
    hipLaunchKernelGGL(k_locate, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->stride, h->S.ni,
                       h->S.nj, h->S.nk, h->xe, h->yn, h->zt, h->pos[h->cur], h->state[h->tw],
                       h->fcell[h->tw]);
    HIPOK(hipGetLastError());
  }
  CHECKRC(do_bin(h));
  CHECKRC(do_neighbour(h));
  if (h->P.nbc) CHECKRC(refresh_free_gids(h));
  if (h->S.p.coupled) {
    if (ref_deposition(h)) {
      CHECKRC(ref_mean_fields(h));
    } else {
      CHECKRC(deposit_current(h));  // no step has deposited yet
    }
  }
  int fl[kFlagCount];
  HIPOK(hipMemcpyAsync(fl, h->flags, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  h->inited = true;
  h->first_pass = true;
  h->st.steps = 0;
  return fatal_flags(fl);  // EP_G of the initial deposition
}

int32_t mfixdem_step(mfixdem_handle* h, int32_t nn) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (!h->inited) return MFIXDEM_ERR_NOT_READY;
  return one_step(h, nn);
}

int32_t mfixdem_run(mfixdem_handle* h, int32_t nn_first, int32_t nsteps) {
  if (!h || nsteps < 0) return MFIXDEM_ERR_ARG;
  if (!h->inited) return MFIXDEM_ERR_NOT_READY;
  for (int32_t t = 0; t < nsteps; t++) CHECKRC(one_step(h, nn_first + t));
  return MFIXDEM_OK;
}

int32_t mfixdem_synchronize(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (!h->inited) {
    HIPOK(sync_stream(h));
    return MFIXDEM_OK;
  }
  int fl[kFlagCount];
  HIPOK(hipMemcpyAsync(fl, h->flags, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  return fatal_flags(fl);
}

int32_t mfixdem_calc_force_dem(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  invalidate_graphs(h);
  CHECKRC(launch_step<kModeForce>(h));
  h->pcur = 1 - h->pcur;
  return MFIXDEM_OK;
}
int32_t mfixdem_calc_drag_des(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  return launch_step<kModeDrag>(h);
}
int32_t mfixdem_cfnewvalues(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  // inlets and outlets run inside the whole step (bc_step) only
  if (h->P.nbc) return MFIXDEM_ERR_UNSUPPORTED;
  CHECKRC(launch_step<kModeIntegrate>(h));
  h->cur = 1 - h->cur;
  h->first_pass = false;
  return MFIXDEM_OK;
}
int32_t mfixdem_desgrid_pic(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  return do_bin(h);
}
int32_t mfixdem_neighbour(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  return do_neighbour(h);
}
int32_t mfixdem_particles_in_cell(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  h->xr_packed = false;  // particles may be removed: the walk's candidates are packed again
  const int n = h->n;
  hipLaunchKernelGGL(k_locate, dim3(nblk(n)), dim3(kBS), 0, h->stream, n, h->stride, h->S.ni,
                     h->S.nj, h->S.nk, h->xe, h->yn, h->zt, h->pos[h->cur], h->state[h->tw],
                     h->fcell[h->tw]);
  HIPOK(hipGetLastError());
  if (ref_deposition(h)) return do_pic_lists(h);
  return delete_outside(h);
}
// CALC_INTERP_WEIGHTS (calc_interp_weights.f:22-33): the filter weights are
// evaluated inside the drag and deposition kernels from the current
// positions, so there is nothing to store here.
int32_t mfixdem_calc_interp_weights(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  return MFIXDEM_OK;
}

int32_t mfixdem_comp_mean_fields(mfixdem_handle* h) {
  if (!h || !h->inited) return h ? MFIXDEM_ERR_NOT_READY : MFIXDEM_ERR_ARG;
  if (ref_deposition(h)) {
    const mfixdem_params& p = h->S.p;
    const int nnode = (p.imax + 1) * (p.jmax + 1) * (p.kmax + 1);
    hipLaunchKernelGGL(k_deposit_nodes, dim3(nblk(nnode)), dim3(kBS), 0, h->stream, p.imax,
                       p.jmax, p.kmax, h->S.ni, h->S.nj, p.nphases, h->stride, h->xe, h->yn,
                       h->zt, h->fluid, h->pic_off, h->pic, h->pos[h->cur], h->pmass[h->tw],
                       h->phase[h->tw], h->rops_node, h->S.ncell);
    HIPOK(hipGetLastError());
    return do_cell_fields(h, false);
  }
  return deposit_current(h);
}

// DES_VOL_NODE (compute_volume_of_nodes, cfassign.f:100-230) with each
// cell's dx(I)*dy(J)*dz(K), in the reference's sum order, and the buffers
// of the gas-side drag.
int32_t ensure_gas_drag(mfixdem_handle* h) {
  if (h->gd_am) return MFIXDEM_OK;
  const Setup& S = h->S;
  const mfixdem_params& p = S.p;
  const int ni = S.ni, nj = S.nj, nc = S.ncell;
  const double af = 0.125;
  auto cell = [&](int i, int j, int k) { return (j - 1) + (i - 1) * nj + (k - 1) * nj * ni; };
  auto volf = [&](int i, int j, int k) {
    return S.fluid[cell(i, j, k)] ? S.cdx[i - 1] * S.cdy[j - 1] * S.cdz[k - 1] : 0.0;
  };
  std::vector<double> vn((size_t)nc, 0.0);
  for (int k = 1; k <= p.kmax + 1; k++)
    for (int j = 1; j <= p.jmax + 1; j++)
      for (int i = 1; i <= p.imax + 1; i++) {
        double a = af * (volf(i, j, k) + volf(i + 1, j, k) + volf(i, j + 1, k) +
                         volf(i + 1, j + 1, k));
        a = a + af * (volf(i, j, k + 1) + volf(i + 1, j + 1, k + 1) + volf(i, j + 1, k + 1) +
                      volf(i + 1, j, k + 1));
        vn[cell(i, j, k)] = a;
      }
  int32_t rc;
  if ((rc = dalloc(&h->gd_fgp, (int64_t)kGasRec * h->stride)) ||
      (rc = dalloc(&h->gd_vol_node, nc)) ||
      (rc = dalloc(&h->gd_am, nc)) || (rc = dalloc(&h->gd_bm, 3 * (int64_t)nc)) ||
      (rc = dalloc(&h->gd_fgds, nc)))
    return rc;
  HIPOK(hipMemcpy(h->gd_vol_node, vn.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
  return MFIXDEM_OK;
}

int32_t mfixdem_drag_gs_gas(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  const mfixdem_params& p = h->S.p;
  if (h->nsys > 1 || !p.coupled) return MFIXDEM_ERR_UNSUPPORTED;
  if (!h->inited) return MFIXDEM_ERR_NOT_READY;
  CHECKRC(ensure_gas_drag(h));
  const int nc = h->S.ncell;
  if (p.interp_scheme != MFIXDEM_INTERP_GARG_2012) {  // DRAG_GS_GAS1: no DRAG_AM
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nc)), dim3(kBS), 0, h->stream, h->gd_am, nc, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nc)), dim3(kBS), 0, h->stream, h->gd_fgds, nc, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * (int64_t)nc)), dim3(kBS), 0, h->stream,
                       h->gd_bm, 3 * nc, 0.0);
    if (h->n > 0)
      hipLaunchKernelGGL(k_gas_filter, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->P,
                         step_ptrs(h), h->vol, nc, h->gd_fgds, h->gd_bm);
    HIPOK(hipGetLastError());
    h->gas_drag_ready = true;
    return MFIXDEM_OK;
  }
  CHECKRC(do_pic_lists(h));
  if (h->n > 0)
    hipLaunchKernelGGL(k_gas_fgp, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->P, step_ptrs(h),
                       h->gd_fgp);
  hipLaunchKernelGGL(k_gas_nodes, dim3(nblk(nc)), dim3(kBS), 0, h->stream, p.imax, p.jmax, p.kmax,
                     h->S.ni, h->S.nj, h->S.nk, h->fluid, h->pic_off, h->pic, h->gd_fgp,
                     h->gd_vol_node, h->gd_am, h->gd_bm);
  hipLaunchKernelGGL(k_gas_fgds, dim3(nblk(nc)), dim3(kBS), 0, h->stream, h->S.ni, h->S.nj,
                     h->S.nk, h->fluid, h->gd_am, h->gd_fgds);
  HIPOK(hipGetLastError());
  h->gas_drag_ready = true;
  return MFIXDEM_OK;
}

int32_t mfixdem_des_time_init(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (!h->S.p.des_explicitly_coupled) return MFIXDEM_OK;
  if (!h->inited) return MFIXDEM_ERR_NOT_READY;
  if (h->n > 0)
    hipLaunchKernelGGL(k_drag_fc, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->P, step_ptrs(h));
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

int32_t mfixdem_get_gas_drag(mfixdem_handle* h, double* f_gds, double* drag_am,
                             double* drag_bm) {
  if (!h) return MFIXDEM_ERR_ARG;
  if (!h->gas_drag_ready) return MFIXDEM_ERR_NOT_READY;
  const size_t nc = (size_t)h->S.ncell;
  if (f_gds)
    HIPOK(hipMemcpyAsync(f_gds, h->gd_fgds, sizeof(double) * nc, hipMemcpyDeviceToHost,
                         h->stream));
  if (drag_am)
    HIPOK(hipMemcpyAsync(drag_am, h->gd_am, sizeof(double) * nc, hipMemcpyDeviceToHost,
                         h->stream));
  if (drag_bm)
    HIPOK(hipMemcpyAsync(drag_bm, h->gd_bm, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost,
                         h->stream));
  HIPOK(sync_stream(h));
  return MFIXDEM_OK;
}

// mfixdem_drag_gs_gas without allocation or host-device copies: results go
// to the caller's device buffers; the F_GP records and DES_VOL_NODE live in
// the device's `scratch` (shared by every call on that device).
int32_t mfixdem_drag_gs_gas_noalloc(mfixdem_handle* h, double* f_gds, double* drag_am,
                                    double* drag_bm) {
  if (!h || !f_gds || !drag_am || !drag_bm) return MFIXDEM_ERR_ARG;
  const mfixdem_params& p = h->S.p;
  if (h->nsys > 1 || !p.coupled) return MFIXDEM_ERR_UNSUPPORTED;
  const int nc = h->S.ncell;
  const bool garg = p.interp_scheme == MFIXDEM_INTERP_GARG_2012;
  if (garg && (int64_t)kGasRec * h->n + nc > kScratchDoubles) return MFIXDEM_ERR_SCRATCH;
  if (!h->inited) return MFIXDEM_ERR_NOT_READY;
  if (!h->scratch_ptr) return MFIXDEM_ERR_HIP;
  if (!garg) {  // DRAG_GS_GAS1: no DRAG_AM, no records
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nc)), dim3(kBS), 0, h->stream, drag_am, nc, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nc)), dim3(kBS), 0, h->stream, f_gds, nc, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(3 * (int64_t)nc)), dim3(kBS), 0, h->stream, drag_bm,
                       3 * nc, 0.0);
    if (h->n > 0)
      hipLaunchKernelGGL(k_gas_filter, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->P,
                         step_ptrs(h), h->vol, nc, f_gds, drag_bm);
    HIPOK(hipGetLastError());
    return MFIXDEM_OK;
  }
  double* rec = h->scratch_ptr;                          // kGasRec * n, all written
  double* vol_node = h->scratch_ptr + (int64_t)kGasRec * h->n;  // nc, all written
  hipLaunchKernelGGL(k_vol_node, dim3(nblk(nc)), dim3(kBS), 0, h->stream, p.imax, p.jmax, p.kmax,
                     h->S.ni, h->S.nj, h->S.nk, h->fluid, h->csz, vol_node);
  CHECKRC(do_pic_lists(h));
  if (h->n > 0)
    hipLaunchKernelGGL(k_gas_fgp, dim3(nblk(h->n)), dim3(kBS), 0, h->stream, h->P, step_ptrs(h),
                       rec);
  hipLaunchKernelGGL(k_gas_nodes, dim3(nblk(nc)), dim3(kBS), 0, h->stream, p.imax, p.jmax, p.kmax,
                     h->S.ni, h->S.nj, h->S.nk, h->fluid, h->pic_off, h->pic, rec, vol_node,
                     drag_am, drag_bm);
  hipLaunchKernelGGL(k_gas_fgds, dim3(nblk(nc)), dim3(kBS), 0, h->stream, h->S.ni, h->S.nj,
                     h->S.nk, h->fluid, drag_am, f_gds);
  HIPOK(hipGetLastError());
  return MFIXDEM_OK;
}

int32_t mfixdem_set_debug(mfixdem_handle* h, int32_t flags) {
  if (!h) return MFIXDEM_ERR_ARG;
  HIPOK(sync_stream(h));  // queued graph launches finish under the old flags
  h->debug = flags;
  invalidate_graphs(h);
  return MFIXDEM_OK;
}

int32_t mfixdem_get_forces(mfixdem_handle* h, double* fc, double* tow, int64_t ld) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || ld < h->n) return MFIXDEM_ERR_ARG;
  const int64_t n = h->n, S = h->stride;
  std::vector<double> tmp;
  if (fc) {
    CHECKRC(download(h, tmp, h->fc_id, 3 * S));
    for (int c = 0; c < 3; c++) std::memcpy(fc + c * ld, tmp.data() + c * S, sizeof(double) * n);
  }
  if (tow) {
    CHECKRC(download(h, tmp, h->tow_id, 3 * S));
    for (int c = 0; c < 3; c++) std::memcpy(tow + c * ld, tmp.data() + c * S, sizeof(double) * n);
  }
  return MFIXDEM_OK;
}

int32_t mfixdem_get_cells(mfixdem_handle* h, int32_t* dg_pijk, int32_t* pijk, int64_t ld) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h || ld < h->n) return MFIXDEM_ERR_ARG;
  const int n = h->n;
  std::vector<int> g, d, f;
  CHECKRC(download(h, g, h->gid[h->tw], n));
  CHECKRC(download(h, d, h->dgc[h->tw], n));
  CHECKRC(download(h, f, h->fcell[h->tw], n));
  const int nj = h->S.nj, ni = h->S.ni;
  for (int i = 0; i < n; i++) {
    const int q = g[i];
    if (dg_pijk) dg_pijk[q] = d[i] + 1;
    if (pijk) {
      const int c = f[i];  // packed I | J << 10 | K << 20
      if (c >= 0) {
        const int I = c & 1023, J = (c >> 10) & 1023, K = c >> 20;
        pijk[q] = I;
        pijk[ld + q] = J;
        pijk[2 * ld + q] = K;
        pijk[3 * ld + q] = J + (I - 1) * nj + (K - 1) * nj * ni;
      } else {
        pijk[q] = pijk[ld + q] = pijk[2 * ld + q] = pijk[3 * ld + q] = 0;
      }
    }
  }
  return MFIXDEM_OK;
}

int32_t mfixdem_get_pairs(mfixdem_handle* h, int32_t* neighbor_index, int32_t* neighbors,
                          double* pft, int64_t capacity, int64_t* npairs) {
  if (h && h->nsys > 1) return MFIXDEM_ERR_ARG;  // single-system call
  if (!h) return MFIXDEM_ERR_ARG;
  if (npairs) *npairs = h->n_pairs;
  if (capacity == 0) return MFIXDEM_OK;
  if (capacity < h->n_pairs) return MFIXDEM_ERR_CAPACITY;
  const int n = h->n, l = h->lt;
  std::vector<int> g, st, sp, nb, pid;
  std::vector<double> pf;
  CHECKRC(download(h, g, h->gid[h->tw], n));
  CHECKRC(download(h, st, h->nbr_start[l], n + 1));
  CHECKRC(download(h, sp, h->nbr_split[l], n));
  CHECKRC(download(h, nb, h->nbr[l], st[n]));
  CHECKRC(download(h, pid, h->nbr_pid[l], st[n]));
  CHECKRC(download(h, pf, h->pft[h->pcur], 3 * h->n_pairs));
  std::vector<uint8_t> pfl;
  CHECKRC(download(h, pfl, h->pfl[h->pcur], h->n_pairs));
  std::vector<int> slot_of(n);
  for (int i = 0; i < n; i++) slot_of[g[i]] = i;
  int64_t w = 0;
  for (int q = 0; q < n; q++) {
    const int i = slot_of[q];
    for (int e = sp[i]; e < st[i + 1]; e++) {
      if (neighbors) neighbors[w] = g[nb[e]] + 1;
      if (pft)
        for (int d = 0; d < 3; d++)
          pft[3 * w + d] = pfl[pid[e]] ? pf[3 * (size_t)pid[e] + d] : 0.0;
      w++;
    }
    if (neighbor_index) neighbor_index[q] = (int32_t)(w + 1);
  }
  return MFIXDEM_OK;
}

// Restart of the pair history (read_res1_des_mod.f:108 reads NEIGHBORS,
// NEIGHBOR_INDEX and PFT_NEIGHBOR back; neighbour.f:77-117 carries them):
// the rows become the previous neighbour list, own entries only, in the
// slot order that mfixdem_set_particles leaves (global-id order), with the
// histories as the previous history buffer, so that the first neighbour
// build (mfixdem_init) carries each surviving pair's history over.
int32_t mfixdem_set_pairs(mfixdem_handle* h, const int32_t* neighbor_index,
                          const int32_t* neighbors, const double* pft, int64_t npairs) {
  if (!h || h->nsys > 1 || h->inited || npairs < 0 || npairs > INT_MAX / 2)
    return MFIXDEM_ERR_ARG;
  const int n = h->n;
  if (n > 0 && !neighbor_index) return MFIXDEM_ERR_ARG;
  if (npairs > 0 && (!neighbors || !pft)) return MFIXDEM_ERR_ARG;
  std::vector<int> st((size_t)n + 1), nb((size_t)npairs), pid((size_t)npairs);
  std::vector<uint8_t> fl((size_t)npairs);
  int64_t prev = 0;
  for (int q = 0; q < n; q++) {
    const int64_t end = (int64_t)neighbor_index[q] - 1;  // 1-based exclusive end
    if (end < prev || end > npairs) return MFIXDEM_ERR_ARG;
    st[q] = (int)prev;
    for (int64_t e = prev; e < end; e++) {
      const int j = neighbors[e] - 1;
      if (j <= q || j >= n) return MFIXDEM_ERR_ARG;  // a half list: partner id above the row's
      nb[e] = j;
      pid[e] = (int)e;
      fl[e] = (pft[3 * e] != 0.0 || pft[3 * e + 1] != 0.0 || pft[3 * e + 2] != 0.0);
    }
    prev = end;
  }
  if (prev != npairs) return MFIXDEM_ERR_ARG;
  st[n] = (int)npairs;
  HIPOK(sync_stream(h));
  const int l = h->lt, pc = h->pcur;
  CHECKRC(ensure_list_capacity(h, l, std::max<int64_t>(npairs, 1)));
  CHECKRC(ensure_pft_capacity(h, pc, std::max<int64_t>(npairs, 1)));
  HIPOK(hipMemcpy(h->nbr_start[l], st.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(h->pair_base[l], st.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  if (n > 0)
    HIPOK(hipMemcpy(h->nbr_split[l], st.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  if (npairs > 0) {
    HIPOK(hipMemcpy(h->nbr[l], nb.data(), sizeof(int) * npairs, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->nbr_pid[l], pid.data(), sizeof(int) * npairs, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->pft[pc], pft, sizeof(double) * 3 * npairs, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->pfl[pc], fl.data(), npairs, hipMemcpyHostToDevice));
  }
  h->n_pairs = npairs;
  return MFIXDEM_OK;
}

int32_t mfixdem_get_status(mfixdem_handle* h, mfixdem_status* out) {
  if (!h || !out) return MFIXDEM_ERR_ARG;
  int fl[kFlagCount];
  HIPOK(hipMemcpyAsync(fl, h->flags, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
  std::vector<uint8_t> s;
  CHECKRC(download(h, s, h->state[h->tw], h->n));
  *out = h->st;
  int64_t np = 0;
  for (uint8_t v : s) np += (v != 0);
  out->n_particles = np;
  out->n_pairs = h->n_pairs;
  out->n_contacts = fl[kFlagContacts];
  out->epg_errors = fl[kFlagEpg];
  out->zero_distance = fl[kFlagZeroDist];
  out->out_of_domain = fl[kFlagOutOfDomain];
  return MFIXDEM_OK;
}

int32_t mfixdem_get_stream(mfixdem_handle* h, void** stream) {
  if (!h || !stream) return MFIXDEM_ERR_ARG;
  *stream = (void*)h->stream;
  return MFIXDEM_OK;
}

int32_t mfixdem_get_kernel_times(mfixdem_handle* h, double* ms, int64_t* calls, int32_t n) {
  if (!h || n < 0) return MFIXDEM_ERR_ARG;
  t_flush(h);
  for (int i = 0; i < n && i < kTimeSlots; i++) {
    if (ms) ms[i] = h->t_ms[i];
    if (calls) calls[i] = h->t_calls[i];
  }
  return MFIXDEM_OK;
}

int32_t mfixdem_reset_kernel_times(mfixdem_handle* h) {
  if (!h) return MFIXDEM_ERR_ARG;
  t_flush(h);
  for (int i = 0; i < kTimeSlots; i++) {
    h->t_ms[i] = 0.0;
    h->t_calls[i] = 0;
  }
  return MFIXDEM_OK;
}

// ================================================================ ensembles
// One handle for nsys independent pure-DEM systems: every launch (the step
// kernel, each rebuild stage) covers all systems; see BatchDev.

int32_t mfixdem_ensemble_create(int32_t nsys, const mfixdem_params* p, mfixdem_handle** out) {
  if (!p || !out || nsys < 1 || nsys > 65535) return MFIXDEM_ERR_ARG;
  *out = nullptr;
  mfixdem_handle* h = new (std::nothrow) mfixdem_handle();
  if (!h) return MFIXDEM_ERR_OOM;
  h->nsys = nsys;
  h->sys_setup.resize(nsys);
  h->sys_off.assign(nsys + 1, 0);
  h->sys_n.assign(nsys, 0);
  for (int s = 0; s < nsys; s++) {
    const mfixdem_params& q = p[s];
    int32_t rc = build_setup(q, h->sys_setup[s]);
    if (rc == MFIXDEM_OK &&
        (q.coupled || q.des_neighbor_search != MFIXDEM_NSEARCH_GRID ||
         q.precision != p[0].precision || q.neighbor_search_n != p[0].neighbor_search_n ||
         q.remove_rogue != p[0].remove_rogue))
      rc = MFIXDEM_ERR_UNSUPPORTED;  // per-launch choices must agree; no gas coupling
    if (rc != MFIXDEM_OK) {
      delete h;
      return rc;
    }
    h->sys_off[s + 1] = h->sys_off[s] + q.max_pip;
  }
  if ((int64_t)h->sys_off[nsys] > INT_MAX / 4) {
    delete h;
    return MFIXDEM_ERR_ARG;
  }
  h->S = h->sys_setup[0];
  h->n = h->sys_off[nsys];
  h->stride = (h->n + 63) / 64 * 64;
  fill_devparams(h);
  std::vector<DevParams> ptab(nsys);
  std::vector<GridDev> gtab(nsys);
  std::vector<int> cell_off(nsys + 1, 0), dg_off(nsys + 1, 0);
  int mbits = 3;
  for (int s = 0; s < nsys; s++) {
    const Setup& S = h->sys_setup[s];
    fill_params(S, h->n, h->stride, ptab[s], gtab[s]);
    int nfine = 0;
    if (search_grid(S.p, S.p.max_pip, gtab[s], nfine) != MFIXDEM_OK) {
      delete h;
      return MFIXDEM_ERR_UNSUPPORTED;
    }
    mbits = std::max(mbits, gtab[s].mbits);
    cell_off[s + 1] = cell_off[s] + nfine;
    dg_off[s + 1] = dg_off[s] + S.dg_ncell;
  }
  int sys_bits = 0;
  while ((1 << sys_bits) < nsys) sys_bits++;
  if (mbits + sys_bits > 64) {
    delete h;
    return MFIXDEM_ERR_UNSUPPORTED;
  }
  for (auto& g : gtab) g.mbits = mbits;
  h->G = gtab[0];
  // a handle-level flag that rebuild stages apply to every slot: carry
  // Adams-Bashforth history if any member uses it
  for (int s = 0; s < nsys; s++) h->P.ab2 = std::max(h->P.ab2, ptab[s].ab2);
  h->n_fine = cell_off[nsys];
  h->key_bits = mbits + sys_bits;
  int32_t rc = alloc_buffers(h);
  std::vector<uint16_t> sys(h->stride, 0);
  for (int s = 0; s < nsys; s++)
    for (int q = h->sys_off[s]; q < h->sys_off[s + 1]; q++) sys[q] = (uint16_t)s;
  h->rogue_init.resize(2 * nsys);
  for (int s = 0; s < nsys; s++) {
    h->rogue_init[2 * s] = INT_MAX;
    h->rogue_init[2 * s + 1] = 0;
  }
  if (rc == MFIXDEM_OK &&
      ((rc = dalloc(&h->d_sys, h->stride)) || (rc = dalloc(&h->d_ptab, nsys)) ||
       (rc = dalloc(&h->d_gtab, nsys)) || (rc = dalloc(&h->d_cell_off, nsys + 1)) ||
       (rc = dalloc(&h->d_dg_off, nsys + 1)) || (rc = dalloc(&h->d_rogue, 2 * nsys))))
    rc = MFIXDEM_ERR_OOM;
  if (rc == MFIXDEM_OK &&
      (hipMemcpy(h->d_sys, sys.data(), sizeof(uint16_t) * h->stride, hipMemcpyHostToDevice) ||
       hipMemcpy(h->d_ptab, ptab.data(), sizeof(DevParams) * nsys, hipMemcpyHostToDevice) ||
       hipMemcpy(h->d_gtab, gtab.data(), sizeof(GridDev) * nsys, hipMemcpyHostToDevice) ||
       hipMemcpy(h->d_cell_off, cell_off.data(), sizeof(int) * (nsys + 1),
                 hipMemcpyHostToDevice) ||
       hipMemcpy(h->d_dg_off, dg_off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice)))
    rc = MFIXDEM_ERR_HIP;
  if (rc == MFIXDEM_OK) {
    std::vector<const Setup*> sp;
    for (int s = 0; s < nsys; s++) sp.push_back(&h->sys_setup[s]);
    rc = upload_facets(h, sp);
  }
  if (rc != MFIXDEM_OK) {
    mfixdem_destroy(h);
    return rc;
  }
  h->B.sys = h->d_sys;
  h->B.ptab = h->d_ptab;
  h->B.gtab = h->d_gtab;
  h->B.cell_off = h->d_cell_off;
  h->B.dg_off = h->d_dg_off;
  h->B.rogue = h->d_rogue;
  h->B.nsys = nsys;
  h->B.sys_shift = mbits;
  h->sys_grid = gtab;
  const size_t S = h->stride;
  h->st_pos.assign(3 * S, 0.0);
  h->st_vel.assign(3 * S, 0.0);
  h->st_om.assign(3 * S, 0.0);
  h->st_r.assign(S, 0.0);
  h->st_rho.assign(S, 0.0);
  h->st_phase.assign(S, 1);
  *out = h;
  return MFIXDEM_OK;
}

int32_t mfixdem_ensemble_set_particles(mfixdem_handle* h, int32_t s, int64_t n, const double* pos,
                                       const double* vel, const double* omega,
                                       const double* radius, const double* ro_sol,
                                       const int32_t* phase, int64_t ld) {
  if (!h || h->nsys < 2 || s < 0 || s >= h->nsys || !pos || !radius || !ro_sol) return MFIXDEM_ERR_ARG;
  const int off = h->sys_off[s], cap = h->sys_off[s + 1] - off;
  if (n < 0 || n > cap || ld < n) return MFIXDEM_ERR_ARG;
  const size_t S = h->stride;
  for (int64_t q = 0; q < n; q++) {
    const int64_t t = off + q;
    for (int c = 0; c < 3; c++) {
      h->st_pos[c * S + t] = pos[c * ld + q];
      h->st_vel[c * S + t] = vel ? vel[c * ld + q] : 0.0;
      h->st_om[c * S + t] = omega ? omega[c * ld + q] : 0.0;
    }
    h->st_r[t] = radius[q];
    h->st_rho[t] = ro_sol[q];
    const int ph = phase ? phase[q] : 1;
    if (ph < 1 || ph > h->sys_setup[s].p.nphases) return MFIXDEM_ERR_ARG;
    h->st_phase[t] = ph;
  }
  h->sys_n[s] = (int)n;
  h->inited = false;
  return MFIXDEM_OK;
}

int32_t mfixdem_ensemble_get_particles(mfixdem_handle* h, int32_t s, double* pos, double* vel,
                                       double* omega, int8_t* state, int64_t ld) {
  if (!h || h->nsys < 2 || s < 0 || s >= h->nsys) return MFIXDEM_ERR_ARG;
  const int off = h->sys_off[s], cap = h->sys_off[s + 1] - off;
  if (ld < cap) return MFIXDEM_ERR_ARG;
  const int64_t S = h->stride;
  std::vector<int> g(cap);
  std::vector<int8_t> st(cap);
  std::vector<double> tmp(cap);
  HIPOK(hipMemcpyAsync(g.data(), h->gid[h->tw] + off, sizeof(int) * cap, hipMemcpyDeviceToHost,
                       h->stream));
  HIPOK(hipMemcpyAsync(st.data(), h->state[h->tw] + off, cap, hipMemcpyDeviceToHost, h->stream));
  HIPOK(sync_stream(h));
  const double* src[3] = {h->pos[h->cur], h->vel[h->cur], h->omega[h->cur]};
  double* dst[3] = {pos, vel, omega};
  for (int a = 0; a < 3; a++) {
    if (!dst[a]) continue;
    for (int c = 0; c < 3; c++) {
      HIPOK(hipMemcpy(tmp.data(), src[a] + c * S + off, sizeof(double) * cap,
                      hipMemcpyDeviceToHost));
      for (int q = 0; q < cap; q++) dst[a][c * ld + g[q]] = tmp[q];
    }
  }
  if (state)
    for (int q = 0; q < cap; q++) state[g[q]] = st[q];
  return MFIXDEM_OK;
}

int32_t mfixdem_set_tuning(mfixdem_handle* h, double cell_scale, int32_t row_kernel) {
  if (!h || !(cell_scale >= 1.0) || cell_scale > 4.0 || row_kernel < MFIXDEM_ROWS_THREAD ||
      row_kernel > MFIXDEM_ROWS_WAVE256)
    return MFIXDEM_ERR_ARG;
  HIPOK(sync_stream(h));
  int n_fine = 0;
  if (h->nsys == 1) {
    GridDev G = h->G;
    if (search_grid(h->S.p, h->n, G, n_fine, cell_scale) != MFIXDEM_OK)
      return MFIXDEM_ERR_UNSUPPORTED;
    G.mbits = h->G.mbits;  // the key layout stays (a coarser grid needs fewer bits)
    h->G = G;
  } else {
    std::vector<GridDev> gt = h->sys_grid;
    std::vector<int> off(h->nsys + 1, 0);
    for (int s = 0; s < h->nsys; s++) {
      int nf = 0;
      const int gb = gt[s].mbits;
      if (search_grid(h->sys_setup[s].p, h->sys_setup[s].p.max_pip, gt[s], nf, cell_scale) !=
          MFIXDEM_OK)
        return MFIXDEM_ERR_UNSUPPORTED;
      gt[s].mbits = gb;
      off[s + 1] = off[s] + nf;
    }
    n_fine = off[h->nsys];
    HIPOK(hipMemcpy(h->d_gtab, gt.data(), sizeof(GridDev) * h->nsys, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(h->d_cell_off, off.data(), sizeof(int) * (h->nsys + 1),
                    hipMemcpyHostToDevice));
    h->sys_grid = gt;
    h->G = gt[0];
  }
  dfree(h->fstart);
  dfree(h->fend);
  CHECKRC(dalloc(&h->fstart, n_fine));
  CHECKRC(dalloc(&h->fend, n_fine));
  h->n_fine = n_fine;
  CHECKRC(setup_fine_cells(h));
  h->cell_scale = cell_scale;
  h->row_kernel = row_kernel;
  h->binned = false;
  return MFIXDEM_OK;
}

int32_t mfixdem_get_device_bytes(mfixdem_handle* h, int64_t* bytes) {
  if (!h || !bytes) return MFIXDEM_ERR_ARG;
  int64_t total = 0;
  int32_t rc = MFIXDEM_OK;
  for_each_buffer(h, [&](auto*& p) {
    if (!p) return;
    size_t sz = 0;
    if (hipMemPtrGetInfo((void*)p, &sz) != hipSuccess) rc = MFIXDEM_ERR_HIP;
    total += (int64_t)sz;
  });
  *bytes = total;
  return rc;
}

}  // extern "C"

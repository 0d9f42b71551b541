// sv_internal.h -- device data layout and kernel entry points shared by the
// translation units of libsunset_amd.so.
//
// Layout in HBM (per context):
//   node arrays  x, y, h, hs, s, node_type     [np]      (real nodes first, then
//                                                        ghosts, reference order)
//   irel                                       [np]      ghost -> parent (0-based),
//                                                        identity for real nodes
//   CSR          row_ptr [npfb+1], col_raw [nnz]         sorted by (r^2, j);
//                                                        col_raw keeps ghost ids
//                wx, wy, wl, wh [nnz]                     weights in CSR order
//   SELL-64      slices of 64 consecutive real nodes; entry (lane l, link k) of
//                slice s at sell_off[s] + k*64 + l; width = max links in slice.
//                scol_raw (ghost ids), scol (ghosts folded to parents),
//                swx, swy, swl, swh.  Padding: col = own node, weight 0.
//   sums         gsx, gsy, lsum, hsum, vol      [npfb]
//   state        records on real nodes only (ghosts are resolved through scol
//                at gather time), see below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/sunset_amd.h"

#define SV_WAVE 64
#define SV_PI 3.141592653589793238462643383279502884197

// RK3(2)4[2R+]C, common_parameter.F90:75-94
#define SV_RKA1 (11847461282814.0 / 36547543011857.0)
#define SV_RKA2 (3943225443063.0 / 7078155732230.0)
#define SV_RKA3 (-346793006927.0 / 4029903576067.0)
#define SV_RKB1 (1017324711453.0 / 9774461848756.0)
#define SV_RKB2 (8237718856693.0 / 13685301971492.0)
#define SV_RKB3 (57731312506979.0 / 19404895981398.0)
#define SV_RKB4 (-101169746363290.0 / 37734290219643.0)
#define SV_RKBH1 (15763415370699.0 / 46270243929542.0)
#define SV_RKBH2 (514528521746.0 / 5659431552419.0)
#define SV_RKBH3 (27030193851939.0 / 9429696342944.0)
#define SV_RKBH4 (-69544964788955.0 / 30262026368149.0)
// PID, common_parameter.F90:97-101
#define SV_PID_TOL 1.0e-5
#define SV_PID_A (0.49 / 2.0)
#define SV_PID_B (0.34 / 2.0)
#define SV_PID_C (0.1 / 2.0)
#define SV_PID_K 0.9

// Device-resident control block: everything the time loop decides lives here so
// a captured step replays without host involvement.
struct sv_ctrl {
  double dt, time, time0, dt_cfl, dt_par;
  double emax_np1, emax_n, emax_nm1;
  double RKa[3], RKb[4], RKbmbh[4], RKc[4];
  double dro;
  int itime, dt_mode;
  int flags;   // bit 0: non-finite value met in a reduction
  // ramped inflow velocity (update_u_inflow, characteristic_boundaries.F90:
  // 256-289): the value the last apply_time_dependent_bounds left, once one
  // has run at time <= ramp_time (uin_set); before that each inflow node
  // keeps its initial velocity
  double uin_ramp;
  int uin_set;
};

// Physical constants passed by value to the stage kernels
struct sv_physk {
  double csq, cs, visc_total, coef_solvent, coef_polymeric, lambda, eps, Mdiff;
  double bfx, bfy, kol_amp, kol_k, rho_char, U_char, L_char;
  double err_norm_rou;   // 1/U_char (setup_flow.F90:156)
  // boundaries: characteristic conditions (characteristic_boundaries.F90),
  // inflow_type 0 soft, 1 hard, 2 inflow-outflow; u_control 1 ramps the
  // inflow velocity u_start -> u_end over ramp_time (load_data.F90:152)
  int inflow_type, u_control;
  double u_start, u_end, ramp_time;
  double Lx;             // L_domain_x (setup_domain.F90:98)
  double p_ref;          // p_inflow = p_outflow = rho_char*csq (load_data.F90:147-148)
};

// Boundary nodes (node types 0 wall, 1 inflow, 2 outflow), each followed in
// node order by its four FD rows (types -1..-4) along the unit normal
// (setup_domain.F90:179-193).  Per boundary node jb the time loop reads one
// record of SV_BREC doubles: kind (0 wall, 1 inflow, 2 outflow), nx, ny, the
// initial velocity u of the node (u_inflow_local, setup_flow.F90:130-137), the
// loop slots of its four FD rows, its spacing s, the wall curvature 1/R and
// L_char (what the direct form's wall diffusion needs, rhs_di.F90:528-545), and
// a pad.
#define SV_BREC 12

// State records (real nodes only; ghosts are resolved through the folded
// stencil columns).  Gather record of node j, what a neighbour sweep reads:
//   [ro, u, v, cxx, cxy, cyy, psixx, psixy, psiyy]  (LC: 10 doubles =  80 B)
//   [ro, u, v, cxx, cxy, cyy]                        (DI:  6 doubles =  48 B)
// packed (16-byte aligned, no padding to a cache line): a sweep's gathers of
// Morton-neighbouring records then read whole lines of useful fields.
// State / register / error records: [ro, rou, rov, f1, f2, f3], 6 doubles
// (48 B, packed), f = psi (LC) or c (DI).
#define SV_GS_LC 10
#define SV_GS_DI 6
#define SV_US 6
struct sv_adj;   // adjoint work space (k_adjoint.hip)
struct sv_ctx;
void sv_adj_free(sv_ctx *c);
#define SV_G_RO 0
#define SV_G_U 1
#define SV_G_V 2
#define SV_G_CXX 3
#define SV_G_CXY 4
#define SV_G_CYY 5
#define SV_G_PXX 6
#define SV_G_PXY 7
#define SV_G_PYY 8

// A segment is one simulation inside a context.  A single context has one; an
// ensemble context concatenates its members, each starting at a multiple of
// SV_BLOCK nodes, so every block of the node kernels lies in one member and
// per-block partial reductions are per-member partial reductions.
#define SV_BLOCK 256
// LDS budget of a block's staged halo records (bytes); above it the sweeps
// gather from global memory instead
#define SV_HALO_LDS_MAX (60 * 1024)
struct sv_segk {
  int base;      // first node (concatenated numbering), a multiple of SV_BLOCK
  int npfb;      // real nodes
  int blk0;      // first logical block
  int nblk;      // blocks
  double volsum; // sum of node volumes (the mass-correction denominator)
  sv_physk pk;
};

// Time-loop layout: the arrays the RK sweeps read, with the real nodes in
// "loop order" -- a Morton (Z-order) sort of the positions, so a block of 256
// consecutive loop slots is a compact patch of the domain and the neighbour
// records it gathers are shared across its nodes in the caches.  perm[k] is
// the caller's node in loop slot k, iperm its inverse.  SELL-64 over loop
// slots; columns are loop slots of the (ghost-folded) neighbours.  Per-node
// arrays are in loop order; the drop-in operators keep the caller's order.
struct sv_loopl {
  bool valid = false;
  int nslices = 0;
  int64_t nnz = 0;
  int64_t *off = nullptr;
  int *col = nullptr;
  double *wx = nullptr, *wy = nullptr, *wl = nullptr, *wh = nullptr;
  int *cnt = nullptr, *cnt_small = nullptr;
  double *gsx = nullptr, *gsy = nullptr, *lsum = nullptr, *hsum = nullptr, *vol = nullptr;
  double *s = nullptr;
  int *perm = nullptr, *iperm = nullptr;
  std::vector<int> hperm;
  // Per-block neighbour halos: the distinct neighbour slots of the SV_BLOCK
  // slots of block b are hidx[hoff[b] .. hoff[b+1]) (ascending); lcol[e] is
  // the position of col[e] in its block's halo.  The sweeps stage a block's
  // halo records in LDS once and read every link from there.
  bool halo = false;
  int hmax = 0;             // largest halo of a block
  int64_t *hoff = nullptr;  // [nblocks + 1]
  int *hidx = nullptr;
  uint16_t *lcol = nullptr; // [nnz]
  // boundary nodes: per loop slot the boundary record index or -1, and the
  // records [nb][SV_BREC]; nullptr when the node set has no boundary nodes
  int *bidx = nullptr;
  double *brec = nullptr;
  // mirror ghosts: per link the sign-flip code of its column (SV_FLIP_*),
  // only when the node set has symmetric / no-slip mirror ghosts
  uint8_t *lflip = nullptr;
};

// Mirror ghosts (mirror_boundaries.F90:263-352): a link to a ghost of vrelation
// 2 (x mirror) reads its parent's record with u and the conformation's xy
// component negated, 3 (y mirror) with v and xy negated, 4 (no-slip, or a
// corner of two mirrors) with u and v negated.  In the halo index of a block
// the code sits above the loop slot: key = slot | code << SV_FLIP_SHIFT.
#define SV_FLIP_SHIFT 29
#define SV_FLIP_MASK ((1 << SV_FLIP_SHIFT) - 1)

struct sv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sv_config cfg{};
  std::string err;
  // domain
  bool have_domain = false;
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1, smax = 0, h0 = 0, h2 = 0, L_char = 1;
  int bc[4] = {1, 1, 1, 1};
  // nodes
  int npfb = 0, np = 0, cap = 0;
  double *x = nullptr, *y = nullptr, *h = nullptr, *hs = nullptr, *s = nullptr;
  int *node_type = nullptr, *irel = nullptr, *vrel = nullptr;
  double hmax = 0, hmin = 0;
  // boundary nodes (types 0..2) in node order, node types on the host, unit
  // normals [cap] (FD rows and ghosts carry their parent's)
  int nb = 0;
  std::vector<int> hblist, hnt;
  int *blist = nullptr;
  double *rnx = nullptr, *rny = nullptr;
  double *ooR = nullptr;     // [nb] wall curvature 1/R (calc_boundary_weights)
  // [nb][wtt_ld] tangential second-derivative weights of the boundary nodes
  // over their (pruned) rows (ij_wb_grad2(2,:,jj)) and [nb] their sums,
  // both divided by L_char^2 once the sums are computed
  double *wtt = nullptr, *wtt_sum = nullptr;
  int wtt_ld = 0;
  // staging area of the host-array entry points on the device ([12][npfb],
  // kept between calls): sv_get_state, sv_set_state, sv_calc_all_rhs
  double *fields_buf = nullptr;
  bool have_normals = false, have_bweights = false;
  // inflow settings (sv_set_inflow)
  int inflow_type = 0, u_control = 0;
  double u_start = 0.0, u_end = 1.0, ramp_time = 1.0;
  // neighbours
  bool have_lists = false;
  int nplink = 0;
  int64_t nnz = 0;
  int *cnt = nullptr, *cnt_small = nullptr;
  int64_t *row_ptr = nullptr;
  int *col_raw = nullptr;
  int maxK = 0;
  // weights (CSR) and SELL
  bool have_weights = false, have_sums = false, have_filter = false;
  double *wx = nullptr, *wy = nullptr, *wl = nullptr, *wh = nullptr;
  double *gsx = nullptr, *gsy = nullptr, *lsum = nullptr, *hsum = nullptr, *vol = nullptr;
  int nslices = 0;
  int64_t sell_nnz = 0;
  int64_t *sell_off = nullptr;
  int *sell_w = nullptr;
  int *scol = nullptr, *scol_raw = nullptr;
  double *swx = nullptr, *swy = nullptr, *swl = nullptr, *swh = nullptr;
  bool sell_dirty = true;
  int *singular = nullptr;  // count of zero pivots in the last solve
  int n_singular = 0, n_systems = 0;
  int weights_fixed = 0;    // launch variant when not autotuning (see launch_order)
  // weight-solve launch autotuning per order (4, 6, 8): chosen variant (0 one
  // block per system, 1 persistent grid), -1 until tuned; sample times
  bool weights_autotune = true;
  int wtune[3] = {-1, -1, -1};
  float wtune_ms[3][4] = {};
  int wvariant[3] = {-1, -1, -1};   // variant of the last launch per order
  // several orders: 0 per-order launches, 1/2 one mixed-order persistent launch
  // with 16/32-neighbour tiles; chosen by the autotuner (-1 untuned) or fixed
  int wmixed = -1, weights_mixed = 0, wmixed_used = 0;
  float wmixed_ms[3] = {};
  // physics/state
  bool have_phys = false, have_state = false;
  sv_physk pk{};
  sv_phys phys{};
  double *U = nullptr, *Uf = nullptr, *reg = nullptr, *eacc = nullptr;  // [npfb][SV_US]
  double *GA = nullptr, *GB = nullptr;                                 // [npfb][gs]
  int gs = SV_GS_LC;
  double *fx = nullptr;    // per-node Kolmogorov forcing
  double *partial = nullptr; int npartial = 0;
  double volsum = 0.0;     // sum of node volumes (the mass-correction denominator)
  sv_ctrl *ctrl = nullptr;  // [nseg]
  // segments: ntot = padded node count of the state arrays, nblk = blocks
  int nseg = 1, ntot = 0, nblk = 0;
  sv_segk *segk = nullptr;  // [nseg]
  int *blk_seg = nullptr;   // [nblk] segment of each logical block
  bool is_ensemble = false;
  sv_adj *adj = nullptr;
  sv_loopl L;
  int loop_order = 1;       // 1: Morton order for the time loop, 0: the caller's order
  int halo_sweeps = 1;      // 1: stage neighbour halos in LDS when they fit
  int adj_lanes = 0;        // lanes per node of the adjoint's transposed gathers
                            // (1, 4; 0: by node count), read by sv_adj_setup
  // ensemble: loop-slot -> member-node permutation of each member
  std::vector<std::vector<int>> member_perm;
  // ensemble launch groups: the blocks of the members of one LABFM order,
  // launched together with their own LDS halo budget (blist: logical blocks,
  // device).  lgroup_mode: -1 untuned, 0 one launch over all blocks, 1 one
  // launch per group (chosen by timing both at sv_ensemble_create)
  struct lgroup { int nblk = 0, hmax = 0; int *blist = nullptr; };
  std::vector<lgroup> lgroups;
  int lgroup_mode = 0;
  float lgroup_ms[2] = {0.0f, 0.0f};
  bool any_forcing = false, any_mdiff = false;  // over segments
  bool has_mirror = false;  // symmetric / no-slip mirror ghosts exist
  std::vector<sv_segk> hseg;
  // scratch for drop-in ops
  double *tmp_phi = nullptr, *tmp_out = nullptr; int tmp_cap = 0;
  // graph
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  // timings (ms)
  double t_neigh = 0, t_weights = 0, t_weights_tune = 0, t_step = 0;
  // device buffer registry for cleanup
  std::vector<void *> allocs;
};

// ---- helpers (sv_api.cpp) ----
// Sets the context's message (or, for c == NULL, the calling thread's
// context-free message that sv_last_error(NULL) returns) and returns code.
int sv_fail(sv_ctx *c, int code, const char *fmt, ...);
// Per-device address of the scratch array of sv_diagnostics_noalloc
// (k_diag.hip), looked up once per device with that device current.
void sv_diag_scratch_init(int device);
// LABFM orders the weight kernels are instantiated for
static inline bool sv_order_supported(int m) { return m == 4 || m == 6 || m == 8 || m == 10; }
#define SV_HIP(c, call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return sv_fail((c), SV_E_HIP, "%s failed: %s (%s:%d)", #call,            \
                     hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)
int sv_alloc(sv_ctx *c, void **p, size_t bytes);

// A temporary device buffer owned by the enclosing scope: freed on every exit
// path, error returns included (hipFree waits for work that still uses it).
template <class T>
struct DevTmp {
  T *p = nullptr;
  DevTmp() = default;
  DevTmp(const DevTmp &) = delete;
  DevTmp &operator=(const DevTmp &) = delete;
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) { return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1)); }
  operator T *() const { return p; }
};

// A HIP event owned by the enclosing scope.
struct EventTmp {
  hipEvent_t e = nullptr;
  EventTmp() = default;
  EventTmp(const EventTmp &) = delete;
  EventTmp &operator=(const EventTmp &) = delete;
  ~EventTmp() {
    if (e) (void)hipEventDestroy(e);
  }
  hipError_t create() { return hipEventCreate(&e); }
  operator hipEvent_t() const { return e; }
};
int sv_ensure_sell(sv_ctx *c);
void sv_free(sv_ctx *c, void *p);

// ---- kernels: nodes/neighbours (k_nodes.hip) ----
int sv_dev_create_ghosts(sv_ctx *c);
int sv_dev_find_neighbours(sv_ctx *c);
// adapt_stencils on the current lists: h, hs, cnt, cnt_small, col_raw change
int sv_dev_adapt(sv_ctx *c, int *reductions_host);
int sv_dev_build_sell(sv_ctx *c);
// time-loop layout from the final weights (after sums and filter coefficients)
int sv_dev_build_loop(sv_ctx *c);
int sv_ensure_loop(sv_ctx *c);
// ---- kernels: weights (k_weights.hip) ----
int sv_dev_weights(sv_ctx *c, const int *order_map_host);
int sv_dev_sums(sv_ctx *c);
int sv_dev_filter_coeffs(sv_ctx *c);
// ---- kernels: boundary nodes (k_bound.hip) ----
int sv_dev_boundary_normals(sv_ctx *c, const double *rnx_host, const double *rny_host);
int sv_dev_boundary_weights(sv_ctx *c);
// the boundary nodes' tangential second-derivative weights: sums and the
// 1/L_char^2 scaling (calc_labf_sums, labf.F90:428-443)
int sv_dev_boundary_tt_sums(sv_ctx *c);
// out[i] of every boundary node i = its tangential second derivative of phi
// (calc_laplacian_transverse_only_on_bound, derivatives.F90:327-337); phi
// sized np with raw (ghost) columns, else npfb with the ghosts' parents
int sv_dev_boundary_tt_apply(sv_ctx *c, const double *phi, double *out, bool raw_cols,
                             hipStream_t st);
// ---- kernels: stencil operators (k_stencil.hip) ----
int sv_dev_apply(sv_ctx *c, int op, const double *phi, double *out, bool raw_cols,
                 hipStream_t st);
// ro (npfb, caller order) -= the volume-weighted mean deviation from 1
int sv_dev_mass_correct(sv_ctx *c, double *ro, hipStream_t st);
// caller-order device arrays (six of npfb, null entries read as 0 / skipped)
// <-> loop-order records on the context's stream
int sv_dev_gather_records(sv_ctx *c, double *dst, const double *const f[6]);
int sv_dev_scatter_records(sv_ctx *c, const double *src, double *const f[6]);
// the eleven fields of sv_get_state in caller order into out[11][npfb]
// (single context; the psi rows are left untouched in DI)
int sv_dev_get_fields(sv_ctx *c, double *out);
// nf loop-order fields (in[f][npfb]) to the caller's order (out[f][npfb])
int sv_dev_to_caller_order(sv_ctx *c, int nf, const double *in, double *out);
// 0 if p is device memory of the context's device in this library's HIP
// runtime, else an SV_E_ARG failure naming `what`
int sv_check_device_ptr(sv_ctx *c, const void *p, const char *what);
// ---- kernels: time stepping (k_step.hip) ----
int sv_dev_init_state(sv_ctx *c);
// one step; ev (nullable) gets SV_NKERNELS+1 events recorded between launches
#define SV_NKERNELS 8
int sv_dev_step(sv_ctx *c, hipEvent_t *ev = nullptr);
int sv_dev_rhs(sv_ctx *c, int iRKstep, double *out6, double *tapeG = nullptr,
               double *tapeD = nullptr);
// DI step that also writes, per stage k, the own gather record to tape[2k]
// ([ntot][SV_US]) and the derivative values to tape[2k+1] ([ntot][SV_TAPE_D])
#define SV_TAPE_D 16
int sv_dev_step_tape(sv_ctx *c, double *const tape[8]);
// per-block partial minima of set_tstep for the current state (no correction)
int sv_dev_init_partials(sv_ctx *c);
// ensembles: time one launch against per-group launches, keep the faster
int sv_dev_tune_groups(sv_ctx *c, int reps);
// dst[k] += add for k < n (index rebasing when members are concatenated)
int sv_dev_offset_int(sv_ctx *c, int *dst, int64_t n, int add);
int sv_dev_offset_i64(sv_ctx *c, int64_t *dst, int64_t n, int64_t add);

// Logical block of this workgroup: the XCD-aware index, through a launch
// group's block list when the launch covers a subset of the blocks.
__device__ __forceinline__ int sv_xcd_block();
__device__ __forceinline__ int sv_logical_block(const int *blist) {
  const int b = sv_xcd_block();
  return blist ? blist[b] : b;
}

// XCD-aware block index: consecutive logical blocks land on the same XCD (one
// L2), so neighbouring slices share cached field lines.  Blocks are dispatched
// round-robin over the 8 XCDs of an MI355X.
__device__ __forceinline__ int sv_xcd_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, xcd = b & 7, idx = b >> 3;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + idx;
}

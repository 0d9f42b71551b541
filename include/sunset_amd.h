/*
 * sunset_amd.h -- C ABI of the MI355X (gfx950) LABFM hot-path library
 * (libsunset_amd.so).
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative SV_E_* code;
 *    sv_last_error(ctx) gives the message.  Host pointers are never retained
 *    past a call.  One host thread per context; calls are not reentrant.
 *  - Host arrays use the reference's layouts: node arrays are indexed by node
 *    (0-based here), neighbour lists are ij_link(nplink, npfb) column-major
 *    (node-contiguous), 1-based, ghost ids > npfb (common_vars.F90:87-89), and
 *    weights are ij_w_grad(2, nplink, npfb), ij_w_lap / ij_w_hyp(nplink, npfb)
 *    (common_vars.F90:92-95).  Any output pointer may be NULL (not copied).
 *  - Entry points ending in _dev take device pointers and a hipStream_t passed
 *    as void* (NULL = the context's stream); they neither allocate nor copy.
 *  - All arithmetic is FP64.
 */
#ifndef SUNSET_AMD_H
#define SUNSET_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sv_ctx sv_ctx;

enum sv_status {
  SV_OK = 0,
  SV_E_ARG = -1,        /* invalid argument or call order */
  SV_E_HIP = -2,        /* HIP runtime error */
  SV_E_NOMEM = -3,      /* device allocation failed */
  SV_E_OVERFLOW = -4,   /* neighbour list longer than nplink */
  SV_E_SINGULAR = -5,   /* a moment matrix had a zero pivot */
  SV_E_STATE = -6       /* operation needs a step that has not been done */
};

enum sv_form { SV_FORM_DI = 0, SV_FORM_LC = 1 };      /* -Ddi / -Dlc (Makefile:73-83) */
/* PID control (set_tstep_PID, step.F90:624-682); a fixed dt; or the CFL-limited
 * dt = min(dt_parabolic, dt_cfl) of every step (set_tstep's non-PID line,
 * step.F90:610-611) */
enum sv_dt_mode { SV_DT_PID = 0, SV_DT_FIXED = 1, SV_DT_CFL = 2 };
enum sv_op {                                          /* derivatives.F90 routines */
  SV_OP_GRAD = 0,        /* calc_gradient          :26-71   */
  SV_OP_LAP = 1,         /* calc_laplacian         :73-116  */
  SV_OP_GRAD_X = 2,      /* calc_gradient_only(1)  :416-464 */
  SV_OP_GRAD_Y = 3,      /* calc_gradient_only(2)  :416-464 */
  SV_OP_FILTER = 4,      /* calc_filtered_var      :360-414 */
  SV_OP_FILTER_TERM = 5, /* calc_filter_term       :466-511 */
  SV_OP_LAP_TRANSVERSE = 6 /* calc_laplacian_transverse_only_on_bound :306-358 */
};

typedef struct {
  int32_t morder;      /* derivative order (4,6,8,10): -Dmorder, common_parameter.F90:57-66 */
  int32_t forder;      /* filter order: -Dforder, common_parameter.F90:45-56 */
  int32_t form;        /* SV_FORM_DI or SV_FORM_LC */
  int32_t device;      /* HIP device ordinal */
} sv_config;

typedef struct {       /* control.in quantities (load_data.F90:21-159) */
  double L_char, U_char, Re, Wi, beta, epsPTT, Ma, Mdiff;
  double grav[3], driving_force[3];
  /* Kolmogorov forcing f_x = kol_amp*sin(kol_k*y) added to rhs_rou as the
   * forcing line of rhs.F90:275 (the LC/DI forms have it commented out,
   * rhs_lc.F90:263, rhs_di.F90:264).  0 disables it. */
  double kol_amp, kol_k;
} sv_phys;

/* ---- context --------------------------------------------------------- */
/* Refuses forms other than DI/LC and orders other than 4, 6, 8, 10 (the
 * orders the device weight kernels are built for; order 10 without boundary
 * nodes and without stencil adaptation). */
int  sv_create(const sv_config *cfg, sv_ctx **out);
void sv_destroy(sv_ctx *ctx);
/* The context's last message; with ctx == NULL, the calling thread's last
 * sv_create / sv_ensemble_create failure (no context exists to hold it). */
const char *sv_last_error(const sv_ctx *ctx);
const char *sv_version(void);

/* ---- domain and nodes: replaces build_domain's IPART read
 *      (setup_domain.F90:93-202) and create_mirror_particles
 *      (mirror_boundaries.F90:34-261) ------------------------------------ */
/* IPART header: box, bc flags (xL xU yL yU; 0 none 1 periodic 2 symmetric
 * 3 no-slip) and smax (sets h0 = hovs*smax, h2 = h0^2: setup_domain.F90:105) */
int sv_set_domain(sv_ctx *ctx, double xmin, double xmax, double ymin, double ymax,
                  const int32_t bc[4], double smax, double L_char);
/* Real nodes.  Ghost (mirror) nodes are created on the device, numbered and
 * positioned exactly as create_mirror_particles does. */
int sv_set_nodes(sv_ctx *ctx, int32_t npfb, const double *x, const double *y,
                 const double *s, const double *h, const double *h_small,
                 const int32_t *node_type);
/* Boundary nodes.  node_type 0 (wall), 1 (inflow) or 2 (outflow) marks a
 * boundary node; it must be followed in node order by its four FD rows
 * (types -1..-4) at s, 2s, 3s, 4s along its normal, as the IPART reader
 * generates them (setup_domain.F90:179-193); sv_set_nodes checks this layout.
 * sv_set_boundaries gives the unit normals into the fluid (rnorm,
 * setup_domain.F90:177), sized npfb; entries of FD rows are taken from their
 * boundary node, other non-boundary entries are ignored.  Required before
 * sv_calc_boundary_weights when the node set has boundary nodes.  In the
 * direct form with polymer diffusion (Mdiff != 0) the Laplacian of c on walls
 * is its tangential part (derivatives.F90:306-358) plus the no-flux term of
 * rhs_di.F90:528-545 over the FD rows; the adjoint takes node sets without
 * boundary nodes. */
int sv_set_boundaries(sv_ctx *ctx, const double *rnx, const double *rny);
/* The inflow line of control.in (load_data.F90:152): inflow_type 0 soft
 * (partially non-reflecting), 1 hard, 2 inflow-outflow
 * (characteristic_boundaries.F90:50-175); velocity_control 0 keeps each
 * inflow node's initial u, 1 ramps it from u_start to u_end over ramp_time
 * (update_u_inflow, :256-289).  Default 0 0 0 1 1. */
int sv_set_inflow(sv_ctx *ctx, int32_t inflow_type, int32_t velocity_control, double u_start,
                  double u_end, double ramp_time);
/* np = npfb + ghosts; x,y sized np; irelation/vrelation sized np-npfb
 * (1-based parents, as common_vars.F90:101) */
int sv_get_ghosts(sv_ctx *ctx, int32_t *np, double *x, double *y, int32_t *irelation,
                  int32_t *vrelation, int32_t cap);

/* ---- neighbours: find_neighbours (neighbours.F90:32-341) ---------------- */
/* Lists sorted by (r^2, j) (the reference's quicksort orders equal r^2
 * arbitrarily).  Host copies optional; the device lists stay resident. */
int sv_find_neighbours(sv_ctx *ctx, int32_t nplink, int32_t *ij_count,
                       int32_t *ij_count_small, int32_t *ij_link);

/* adapt_stencils (labf.F90:1079-1416) + adapt_stencils_main (:1418-1721),
 * called between sv_find_neighbours and the weights as sunset.F90:40-52 does
 * without a restart.  Each interior node (node_type > 2, not 998) reduces h
 * by 2% (the single-precision 0.98) at a time while the Laplacian moment
 * system's SVD residual stays below res_tol/h (res_tol = c n^4 eps / m,
 * c = 1e-3, 5e-3, 3e-2 for orders 4, 6, 8) and the d/dx, d/dy, Laplacian
 * operators amplify none of the wavenumbers i pi/(32 s), i = 1..33, by more
 * than 1.0001 (nor the Laplacian below 0), at most 26 times; pass 1 with the
 * filter order sets h and h_small, pass 2 with the derivative order sets
 * h_small.  The lists are then rebuilt from the new h, h_small (weights, sums
 * and filter coefficients must be recomputed).  Needs morder <= forder.
 * h, h_small (npfb, nullable): the adapted values; reductions (2*npfb,
 * nullable): accepted reductions per node, pass 1 then pass 2.  Refused after
 * sv_set_initial; SV_E_SINGULAR if an SVD does not converge (the reference
 * stops, svdlib.F90:218-221). */
int sv_adapt_stencils(sv_ctx *ctx, double *h, double *h_small, int32_t *reductions);

/* ---- weights: labf.F90 -------------------------------------------------- */
/* calc_labf_weights (labf.F90:34-231) incl. calc_filter_weights (:233-394)
 * and calc_node_volumes (:453-503).  order_map (nullable, per real node):
 * LABFM order of that node's derivative and filter systems, overriding
 * morder/forder (variable-order node sets). */
int sv_calc_labf_weights(sv_ctx *ctx, const int32_t *order_map);
/* calc_boundary_weights (labf.F90:505-1039), after sv_calc_labf_weights and
 * before sv_calc_labf_sums, as sunset.F90:54-55 (a no-op without boundary
 * nodes): boundary nodes keep their same-type neighbours and FD rows; their
 * operators are 5-point one-sided differences along the normal and a 6th-order
 * 1D LABFM along the tangent (walls: in the normal-tangent frame, with the
 * curvature term in the Laplacian); rows 1 and 2 get normal differences plus a
 * 1D LABFM along a tangent turned by the tangential gradient of s, mapped to
 * x-y.  The Laplacian weights of inflow/outflow nodes are zero (the reference
 * never writes them).  sv_calc_labf_weights skips boundary nodes and the
 * derivative systems of rows 1 and 2, which this call replaces. */
int sv_calc_boundary_weights(sv_ctx *ctx);
int sv_calc_labf_sums(sv_ctx *ctx);           /* labf.F90:396-451 */
int sv_filter_coefficients(sv_ctx *ctx);      /* labf.F90:1723-1811 */
/* Host copies in the reference layouts (see top); any pointer may be NULL */
int sv_get_weights(sv_ctx *ctx, int32_t nplink, double *ij_w_grad, double *ij_w_lap,
                   double *ij_w_hyp, double *ij_w_grad_sum, double *ij_w_lap_sum,
                   double *ij_w_hyp_sum, double *vol);
/* Device time of the last call, milliseconds: "neighbours", "weights" (the
 * weight-solve launches), "weights_tune" (autotuning sample launches of that
 * call, if any), "step" (the last sv_run);
 * "systems" gives the number of moment systems solved by the last
 * sv_calc_labf_weights; "weights_variant_m4|6|8" the launch variant used for
 * that order (bit 0: persistent grid, else one block per system; bit 1:
 * 32-neighbour assembly tiles, else 16; -1 none);
 * "weights_tune_m<order>_<variant>" the autotuning sample time of a variant. */
int sv_get_timing(sv_ctx *ctx, const char *what, double *ms);
/* Tuning switches:
 *   "weights_autotune" (default 1): for an order with >= 16384 systems, the
 *       first sv_calc_labf_weights times the four launch variants (one block
 *       per system or a persistent grid fed from a work queue, times assembly
 *       tiles of 16 or 32 neighbours) on a strided 4096-system sample and uses
 *       the fastest; the choice is kept for later calls (setting the option
 *       again re-tunes).
 *   "weights_persistent" (0/1): fixes the launch to one-block-per-system (0) or
 *       the persistent grid (1), with 16-neighbour tiles, and turns the
 *       autotuning off.
 *   "weights_variant" (0-3): fixes the launch variant (bits as in
 *       "weights_variant_m<order>") and turns the autotuning off.
 *   "weights_mixed" (0-2): with several orders (an order map), 0 launches each
 *       order separately, 1 and 2 run every system in one persistent launch
 *       (queue: order 8 first, longest stencils first; 16- or 32-neighbour
 *       tiles); turns the autotuning off.  Autotuned (the default), the
 *       single launch is timed against the per-order launches on a strided
 *       4096-system sample of the whole queue.  sv_get_timing "weights_mixed"
 *       reports the mode of the last call, "weights_mixed_tune_<mode>" the
 *       sample times.
 *   "loop_order" (default 1, before sv_set_initial): 1 runs the time loop in
 *       Morton order of the nodes, 0 in the caller's order.
 *   "halo_sweeps" (default 1, before sv_set_initial): the RK sweeps copy each
 *       256-node block's neighbour records into LDS once and read the links
 *       from there (when the largest block halo fits 60 KB); 0 gathers every
 *       link from global memory.
 *   "ensemble_groups" (ensemble contexts): with members of several LABFM
 *       orders, the stage and filter sweeps run either as one launch over all
 *       members (0) or as one launch per order group (1), each group with the
 *       LDS halo budget of its own largest halo.  sv_ensemble_create times the
 *       right-hand-side sweep both ways and keeps the faster; -1 re-tunes.
 *       sv_get_timing "ensemble_groups" reports the mode,
 *       "ensemble_groups_tune_<mode>" the sample times.
 *   "adj_lanes" (0, 1, 4; read by sv_adj_setup): lanes per node of the
 *       adjoint's transposed gathers; 4 splits each node's links over a quad
 *       of lanes (node sets too small to fill the chip at one lane per node),
 *       0 (default) picks 4 below 8 waves per CU. */
int sv_set_option(sv_ctx *ctx, const char *key, double value);

/* ---- derivative operators: derivatives.F90 ------------------------------ */
/* phi sized np (ghost values used as given, as the reference does);
 * out sized npfb, or 2*npfb (x then y) for SV_OP_GRAD; FILTER writes phi in place
 * (first npfb entries). */
int sv_apply(sv_ctx *ctx, int32_t op, double *phi, double *out);
int sv_calc_gradient(sv_ctx *ctx, const double *phi, double *gradphi);   /* :26-71 */
int sv_calc_laplacian(sv_ctx *ctx, const double *phi, double *lapphi);   /* :73-116 */
int sv_calc_gradient_only(sv_ctx *ctx, int32_t dim, const double *phi,
                          double *gradphi);                              /* :416-464 */
int sv_calc_filtered_var(sv_ctx *ctx, double *phi);                      /* :360-414 */
int sv_calc_filter_term(sv_ctx *ctx, const double *phi, double *filtphi);/* :466-511 */
/* The Laplacian, except on boundary nodes (walls, inflow, outflow): there the
 * tangential second derivative only, with the 1D LABFM weights of
 * calc_boundary_weights (labf.F90:755) scaled by 1/L_char^2 in
 * sv_calc_labf_sums (labf.F90:428-443).  With boundary nodes it needs
 * sv_calc_boundary_weights before the sums. */
int sv_calc_laplacian_transverse_only_on_bound(sv_ctx *ctx, const double *phi,
                                               double *lapphi);          /* :306-358 */
/* Device variant over fields stored on real nodes only (ghosts resolved to
 * their periodic parents): nfields fields, each npfb doubles, device pointers.
 * Asynchronous: the launches are queued on `stream` as given (NULL: the null
 * stream), in order with the caller's work on it; the outputs are complete
 * when that stream's work is. */
int sv_apply_dev(sv_ctx *ctx, int32_t op, int32_t nfields, const double *const *phi,
                 double *const *out, void *stream);

/* ---- physics and time stepping: rhs_lc.F90 / rhs_di.F90 / step.F90 ----- */
int sv_set_phys(sv_ctx *ctx, const sv_phys *phys);
/* initial_solution (setup_flow.F90:24-173) with the fields given on real
 * nodes: rou = ro*u, psi = log(c) (log_conf_psi_from_c), dt = 1e-6,
 * emax = pid_tol unless set by sv_set_control. */
int sv_set_initial(sv_ctx *ctx, const double *ro, const double *u, const double *v,
                   const double *cxx, const double *cxy, const double *cyy);
/* The evolved variables of the real nodes exactly as the reference's step
 * registers hold them (step.F90:218-512): ro, rou, rov and psi (LC) or c (DI),
 * replacing the state sv_set_initial derived from primitives (which must come
 * first: it builds the boundary records and inflow velocities). */
int sv_set_state(sv_ctx *ctx, const double *ro, const double *rou, const double *rov,
                 const double *xx, const double *xy, const double *yy);
/* Device-resident forms of sv_set_state / the registers of sv_get_state:
 * U[6] = (ro, rou, rov, xx, xy, yy) as device pointers (npfb each, the
 * caller's node order) on the context's GPU in this library's HIP runtime
 * (from a Python process, import torch first so that the library binds to
 * torch's runtime and its tensors' data pointers qualify); the caller's
 * stream is synchronised first (NULL: the null stream, a framework's default
 * stream), the copy is complete on return.  A pointer of another runtime or
 * device is refused with SV_E_ARG.  No counterpart in the reference: these
 * serve device-tensor bindings (no host copy). */
int sv_set_state_device(sv_ctx *ctx, const double *const U[6], void *stream);
int sv_get_state_device(sv_ctx *ctx, double *const U[6], void *stream);
/* filter_variables (rhs_lc.F90:818-893; rhs_di.F90 alike) on fields the
 * caller holds: each of the six arrays (np entries, ghosts included as the
 * reference keeps them; psi in LC, c in DI) filtered as calc_filtered_var
 * does, then the mass correction of ro (-Dmcorr): ro -= sum((ro - 1) vol) /
 * sum(vol) over the real nodes.  The real-node entries are overwritten. */
int sv_filter_variables(sv_ctx *ctx, double *ro, double *rou, double *rov, double *xx,
                        double *xy, double *yy);
/* restart-style control state (setup_flow.F90:495-545) */
int sv_set_control(sv_ctx *ctx, double dt, double time, int32_t itime, double emax_np1,
                   double emax_n, double emax_nm1, int32_t dt_mode);
int sv_get_control(sv_ctx *ctx, double *dt, double *time, int32_t *itime, double *emax3,
                   double *dt_cfl, double *dt_parabolic);
/* calc_all_rhs (rhs_lc.F90:47-117 / rhs_di.F90:47-117) on the current state;
 * iRKstep as the reference's stage counter (the LC first-stage rule,
 * rhs_lc.F90:503-506, applies at itime==1, iRKstep==1).  Outputs sized npfb. */
int sv_calc_all_rhs(sv_ctx *ctx, int32_t iRKstep, double *rhs_ro, double *rhs_rou,
                    double *rhs_rov, double *rhs_xx, double *rhs_xy, double *rhs_yy);
/* nsteps of the main loop (sunset.F90:75-106): set_tstep, set_tstep_PID
 * (PID mode), step_rk3_4S_2R_EE (step.F90:218-512), itime += 1.  State stays
 * on the device; dt lives on the device.  use_graph != 0 replays one captured
 * step as a hipGraph. */
int sv_run(sv_ctx *ctx, int32_t nsteps, int32_t use_graph);
/* As sv_run (eager launches) with HIP events recorded on the context's stream
 * between the launches of every step.  kernel_ms[8]: mean device time per step
 * of each launch, in launch order: dt control, stage 1, stage 2, stage 3,
 * stage 4, filter, mass-correction reduce, state apply.  total_ms: whole run. */
int sv_run_profiled(sv_ctx *ctx, int32_t nsteps, double *kernel_ms, double *total_ms);
/* stats[6]: mean ij_count, mean ij_count_small, max ij_count, SELL entries
 * (with padding), number of ghost nodes, neighbour-halo records per node
 * staged in LDS by the time-loop sweeps (0 when they gather from memory) */
int sv_get_stencil_stats(sv_ctx *ctx, double *stats);
/* Fields on real nodes (any pointer may be NULL): psi* only in the LC form. */
int sv_get_state(sv_ctx *ctx, double *ro, double *rou, double *rov, double *u, double *v,
                 double *psixx, double *psixy, double *psiyy, double *cxx, double *cxy,
                 double *cyy);
/* Synchronise the context's stream */
int sv_sync(sv_ctx *ctx);

/* ---- Ensembles: many independent simulations advanced together ----------
 * No single counterpart in the reference, which runs one simulation per
 * process (sunset.F90:36-106); an ensemble is N such runs.  Each member is a
 * context taken through the whole setup and sv_set_initial (its own nodes,
 * orders, stencils, physics and control state); all members must have the
 * same conformation form and device.  Members may have boundary nodes (walls,
 * inflows with their own inflow settings and ramp, outflows) and symmetric /
 * no-slip mirror ghosts; an ensemble with mirror members holds at most 2^29
 * nodes.  The ensemble context holds copies of the
 * members' device data in one concatenated layout, so every launch of a step,
 * and one captured hipGraph (sv_run with use_graph), advances all members;
 * each keeps its own dt / PID history / itime (set_tstep and set_tstep_PID per
 * member, step.F90:569-682) and its own mass correction (rhs_lc.F90:871-890).
 * On an ensemble context: sv_run, sv_run_profiled, sv_sync, sv_destroy and the
 * two accessors below; the setup entry points return SV_E_ARG. */
int sv_ensemble_create(sv_ctx *const *members, int32_t nmembers, sv_ctx **ensemble);
int32_t sv_ensemble_size(const sv_ctx *ensemble);
/* sv_get_state / sv_get_control of one member (sv_get_state on an ensemble
 * context is member 0) */
int sv_ensemble_get_state(sv_ctx *ensemble, int32_t member, double *ro, double *rou,
                          double *rov, double *u, double *v, double *psixx, double *psixy,
                          double *psiyy, double *cxx, double *cxy, double *cyy);
int sv_ensemble_get_control(sv_ctx *ensemble, int32_t member, double *dt, double *time,
                            int32_t *itime, double *emax3, double *dt_cfl,
                            double *dt_parabolic);

/* ---- Diagnostics of the current state ------------------------------------
 * The reductions of statistics.F90 on the device, returned instead of written
 * to the reference's statistics files.  member = 0 for a single simulation,
 * the member index on an ensemble context (MASS_ENERGY and CONF only).
 *   SV_DIAG_MASS_ENERGY (mass_and_energy_check, statistics.F90:305-355):
 *     out[0] = sum (ro - rho_char) vol / sum vol, out[1] = sum vol (times
 *     L_char^2), out[2] = sum ro |u|^2 / 2 vol / sum vol
 *   SV_DIAG_CONF (conf_check, :357-442): volume means of cxx, cxy, cyy, czz,
 *     rms of cxy, mean trace (czz = 1 in 2D)
 *   SV_DIAG_KOLMOGOROV (kolmogorov_l2norm, :552-612): volume L2 errors of u,
 *     v, cxx, cxy, cyy against u = cos 2y, cxx = 1 + 8 lambda^2 sin^2 2y,
 *     cxy = -2 lambda sin 2y, cyy = 1
 *   SV_DIAG_TG (error_TG, :614-654): relative L2 error of |u| against the
 *     decaying Taylor-Green vortex at the context's time */
enum { SV_DIAG_MASS_ENERGY = 0, SV_DIAG_CONF = 1, SV_DIAG_KOLMOGOROV = 2, SV_DIAG_TG = 3 };
int sv_diagnostics(sv_ctx *ctx, int32_t kind, int32_t member, double *out);
/* sv_diagnostics without allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  The same reductions, enqueued on `stream` (a
 * hipStream_t passed as void*; NULL = the context's stream) with no host
 * synchronisation, the result written to out_dev (device memory, 6 doubles
 * suffice for every kind).  Block partials and member sums live in a fixed
 * 4 MiB array per device held until the process ends, so a context may have
 * at most 7*(blocks + members) <= 2^19 (about 19M nodes); larger ones get
 * SV_E_ARG.  Every call on one GPU uses the same array: calls that overlap on
 * a GPU (several streams or host threads) race.  SV_DIAG_TG reads the time
 * from the device control block, so its exp may differ from sv_diagnostics'
 * host exp in the last bit. */
int sv_diagnostics_noalloc(sv_ctx *ctx, int32_t kind, int32_t member, double *out_dev,
                           void *stream);

/* ---- Reverse mode (DI form, fixed time step) ----------------------------
 * No counterpart in the reference (it has no adjoint); these are the
 * vector-Jacobian products of the reference's own operators: calc_all_rhs
 * (rhs_di.F90:47-476), the RK3(2)4[2R+]C stages and registers
 * (step.F90:281-409), the hyperviscous filter (derivatives.F90:360-414) and the
 * mass correction (rhs_di.F90:667-720), with stencil transposes built once
 * (no atomics, deterministic).  State vectors are (ro, rou, rov, cxx, cxy, cyy)
 * on real nodes, six arrays of npfb.
 * sv_adj_setup: after sv_set_initial and sv_set_control(..., SV_DT_FIXED);
 * checkpoint_every = steps between stored states in a rollout (the backward
 * sweep recomputes one interval at a time). */
int sv_adj_setup(sv_ctx *ctx, int32_t checkpoint_every);
/* ubar = (d calc_all_rhs / d state)^T rbar at the current state */
int sv_rhs_vjp(sv_ctx *ctx, const double *const rbar[6], double *const ubar[6]);
/* nsteps from the current state, storing checkpoints; loss = (1/nsteps)
 * sum_{t=1..nsteps} mean_i (cxx_i + cyy_i - 2) of the state after each step */
int sv_rollout_forward(sv_ctx *ctx, int32_t nsteps, double *loss);
/* gradient of  loss_bar*loss + <ubar_final, state at the end>  with respect to
 * the state at the start of the last sv_rollout_forward (ubar_final nullable);
 * the context's state is left at the end of the rollout */
int sv_rollout_vjp(sv_ctx *ctx, double loss_bar, const double *const ubar_final[6],
                   double *const grad0[6]);
/* As sv_rollout_vjp with ubar_final (nullable, or null fields) and grad0 as
 * device pointers (see sv_set_state_device); the caller's stream (NULL: the
 * null stream) is synchronised first, the result is complete on return. */
int sv_rollout_vjp_device(sv_ctx *ctx, double loss_bar, const double *const ubar_final[6],
                          double *const grad0[6], void *stream);
/* gradient of the same quantity with respect to the parameters, from the last
 * sv_rollout_vjp (or, for <rbar, rhs>, the last sv_rhs_vjp):
 * grad[0] = d/dWi, grad[1] = d/dbeta, grad[2] = d/d(epsPTT) */
int sv_adj_param_grad(sv_ctx *ctx, double grad[3]);
/* device time (ms) of the last sv_rollout_forward and sv_rollout_vjp */
int sv_adj_timing(sv_ctx *ctx, double *forward_ms, double *vjp_ms);

#ifdef __cplusplus
}
#endif
#endif

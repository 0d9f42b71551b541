// k_step.hip -- device-resident RK3(2)4[2R+]C step (step_rk3_4S_2R_EE,
// step.F90:218-512) with the right-hand sides of rhs_lc.F90 / rhs_di.F90 and
// filter_variables, set_tstep and set_tstep_PID.
//
// One step is 8 launches, all reading their scalars from the device control
// block so the step can be captured once and replayed as a hipGraph:
//   k_dt          (1 block)  set_tstep + set_tstep_PID, emax history shift,
//                            dt-scaled RK coefficients
//   k_stage<1..4>            ONE gather sweep per stage: every gradient and
//                            Laplacian the RHS needs (the reference makes 11
//                            separate sweeps, rhs_lc.F90:64-402), then the
//                            pointwise RHS and the low-storage RK update
//   k_filter                 hyperviscous filter of the evolved fields
//                            (calc_filtered_var, derivatives.F90:360-414) and
//                            the conformation tensor of the filtered psi
//   k_reduce      (1 block)  mass correction (rhs_lc.F90:871-890), emax,
//                            itime += 1
//   k_apply_state            rho -= drho, u = rou/rho, dt partial minima
// Ghost nodes do not exist at run time: periodic links point at the parent
// (mirror copies are exact copies, mirror_boundaries.F90:263-352).
// Reductions are a fixed-order two-level tree, so results are bitwise
// reproducible run to run.
//
// Memory layout of the state (records, see sv_internal.h): what a neighbour
// gather needs of node j sits in one packed 80-byte record (LC) or 48-byte
// record (DI), so a link costs at most two cache lines instead of one line
// per field, and records of Morton-adjacent nodes share lines.
#include "sv_internal.h"

namespace {

constexpr int BS = SV_BLOCK;   // threads per block of the node kernels

// Gather-loop tuning (compile time): links processed per iteration, and
// whether the streamed stencil arrays (columns, weights; read once per sweep)
// are loaded non-temporally so they do not displace the gathered records in
// the caches.
#ifndef SV_GATHER_UNROLL
#define SV_GATHER_UNROLL 2
#endif
#ifndef SV_STREAM_NT
#define SV_STREAM_NT 1
#endif
// links per batch in the halo-staged sweeps (global streams prefetched one
// batch ahead)
#ifndef SV_HALO_UNROLL
#define SV_HALO_UNROLL 2
#endif

template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
#if SV_STREAM_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

// The node handled by this thread: segment of the block, concatenated index i,
// a safe index ii (= i, or the segment's first node for an idle lane).
struct NodeRef {
  int sg, i, ii;
  bool act;
};

__device__ __forceinline__ NodeRef node_ref(const int *blk_seg, const sv_segk *seg, int bid) {
  NodeRef r;
  r.sg = blk_seg[bid];
  const sv_segk &S = seg[r.sg];
  const int li = (bid - S.blk0) * BS + (int)threadIdx.x;
  r.act = li < S.npfb;
  r.i = S.base + li;
  r.ii = r.act ? r.i : S.base;
  return r;
}

// ---- conformation algebra (conf_transforms.F90) -------------------------
// 2D eigens (:399-444); R column-major R[0]=R11 R[1]=R21 R[2]=R12 R[3]=R22
__device__ __forceinline__ void eigens2(double a, double b, double d, double &l1, double &l2,
                                        double R[4]) {
  if (fabs(b) <= 1.0e-16) {
    l1 = a; l2 = d;
    R[0] = 1.0; R[1] = 0.0; R[2] = 0.0; R[3] = 1.0;
    return;
  }
  const double tr = a + d;
  const double det = a * d - b * b;
  double t = 0.25 * tr * tr - det;
  if (t <= 1.0e-16) t = 0.0;
  t = sqrt(t);
  l1 = 0.5 * tr + t;
  l2 = 0.5 * tr - t;
  double r11 = l1 - d, r21 = b;
  const double nrm = sqrt(r11 * r11 + r21 * r21);
  r11 = r11 / nrm;
  r21 = r21 / nrm;
  R[0] = r11; R[1] = r21; R[2] = -r21; R[3] = r11;
}

// log_conf_c_from_psi (:133-169): C = R diag(exp L) R^T
__device__ __forceinline__ void c_from_psi(double pxx, double pxy, double pyy, double &cxx,
                                           double &cxy, double &cyy) {
  double l1, l2, R[4];
  eigens2(pxx, pxy, pyy, l1, l2, R);
  const double e1 = exp(l1), e2 = exp(l2);
  const double i11 = e1 * R[0], i12 = e1 * R[1], i21 = e2 * R[2], i22 = e2 * R[3];
  cxx = R[0] * i11 + R[2] * i21;
  cxy = R[1] * i11 + R[3] * i21;
  cyy = R[1] * i12 + R[3] * i22;
}

// calc_rhs_logconf pointwise part (rhs_lc.F90:416-557)
__device__ __forceinline__ void rhs_logconf(double u, double v, double ux, double uy, double vx,
                                            double vy, double pxx, double pxy, double pyy,
                                            double gpxx_x, double gpxx_y, double gpxy_x,
                                            double gpxy_y, double gpyy_x, double gpyy_y,
                                            bool first, double eps, double lam, double &rxx,
                                            double &rxy, double &ryy) {
  const double adxx = -(u * gpxx_x + v * gpxx_y + 0.0);
  const double adxy = -(u * gpxy_x + v * gpxy_y + 0.0);
  const double adyy = -(u * gpyy_x + v * gpyy_y + 0.0);
  double L1, L2, R[4];
  eigens2(pxx, pxy, pyy, L1, L2, R);
  L1 = exp(L1);
  L2 = exp(L2);
  const double tr_c = L1 + L2;
  const double G11 = ux, G12 = uy, G21 = vx, G22 = vy;
  const double GR11 = G11 * R[0] + G12 * R[1], GR12 = G11 * R[2] + G12 * R[3];
  const double GR21 = G21 * R[0] + G22 * R[1], GR22 = G21 * R[2] + G22 * R[3];
  const double M11 = R[0] * GR11 + R[1] * GR21, M12 = R[0] * GR12 + R[1] * GR22;
  const double M21 = R[2] * GR11 + R[3] * GR21, M22 = R[2] * GR12 + R[3] * GR22;
  double om = L2 * M12 + L1 * M21;
  if (fabs(L2 - L1) <= 1.0e-30) om = 0.0;
  else om = om / (L2 - L1);
  double B11, B12, B22, O11, O12, O21, O22;
  {
    const double i11 = M11 * R[0], i12 = M11 * R[1], i21 = M22 * R[2], i22 = M22 * R[3];
    B11 = R[0] * i11 + R[2] * i21;
    B12 = R[0] * i12 + R[2] * i22;
    B22 = R[1] * i12 + R[3] * i22;
  }
  {
    const double i11 = om * R[2], i12 = om * R[3], i21 = -om * R[0], i22 = -om * R[1];
    O11 = R[0] * i11 + R[2] * i21;
    O12 = R[0] * i12 + R[2] * i22;
    O21 = R[1] * i11 + R[3] * i21;
    O22 = R[1] * i12 + R[3] * i22;
  }
  if (first) {   // rhs_lc.F90:503-506
    O11 = O12 = O21 = O22 = 0.0;
    B11 = 0.5 * (G11 + G11);
    B12 = 0.5 * (G12 + G21);
    B22 = 0.5 * (G22 + G22);
  }
  const double UC11 = (O11 * pxx + O12 * pxy) - (pxx * O11 + pxy * O21) + 2.0 * B11;
  const double UC12 = (O11 * pxy + O12 * pyy) - (pxx * O12 + pxy * O22) + 2.0 * B12;
  const double UC22 = (O21 * pxy + O22 * pyy) - (pxy * O12 + pyy * O22) + 2.0 * B22;
  const double fR = 1.0 - 2.0 * eps + eps * tr_c;
  const double C1 = fR / L1 - fR, C2 = fR / L2 - fR;
  const double i11 = C1 * R[0], i12 = C1 * R[1], i21 = C2 * R[2], i22 = C2 * R[3];
  rxx = adxx + UC11 + (R[0] * i11 + R[2] * i21) / lam;
  rxy = adxy + UC12 + (R[0] * i12 + R[2] * i22) / lam;
  ryy = adyy + UC22 + (R[1] * i12 + R[3] * i22) / lam;
}

// ---- record access ----------------------------------------------------
struct Rec9 {   // gathered fields of one node
  double ro, u, v, cxx, cxy, cyy, pxx, pxy, pyy;
};

template <int FORM>
__device__ __forceinline__ Rec9 load_g(const double *G, int j) {
  constexpr int S = (FORM == SV_FORM_LC) ? SV_GS_LC : SV_GS_DI;
  const double2 *p = reinterpret_cast<const double2 *>(G + (size_t)j * S);
  Rec9 r;
  const double2 a = p[0], b = p[1], c = p[2];
  r.ro = a.x; r.u = a.y; r.v = b.x; r.cxx = b.y; r.cxy = c.x; r.cyy = c.y;
  if (FORM == SV_FORM_LC) {
    const double2 d = p[3];
    r.pxx = d.x; r.pxy = d.y;
    r.pyy = G[(size_t)j * S + SV_G_PYY];
  } else {
    r.pxx = r.pxy = r.pyy = 0.0;
  }
  return r;
}

template <int FORM>
__device__ __forceinline__ void store_g(double *G, int i, const Rec9 &r) {
  constexpr int S = (FORM == SV_FORM_LC) ? SV_GS_LC : SV_GS_DI;
  double2 *p = reinterpret_cast<double2 *>(G + (size_t)i * S);
  p[0] = double2{r.ro, r.u};
  p[1] = double2{r.v, r.cxx};
  p[2] = double2{r.cxy, r.cyy};
  if (FORM == SV_FORM_LC) {
    p[3] = double2{r.pxx, r.pxy};
    G[(size_t)i * S + SV_G_PYY] = r.pyy;
  }
}

__device__ __forceinline__ void load_u(const double *U, int i, double f[6]) {
  const double2 *p = reinterpret_cast<const double2 *>(U + (size_t)i * SV_US);
  const double2 a = p[0], b = p[1], c = p[2];
  f[0] = a.x; f[1] = a.y; f[2] = b.x; f[3] = b.y; f[4] = c.x; f[5] = c.y;
}

__device__ __forceinline__ void store_u(double *U, int i, const double f[6]) {
  double2 *p = reinterpret_cast<double2 *>(U + (size_t)i * SV_US);
  p[0] = double2{f[0], f[1]};
  p[1] = double2{f[2], f[3]};
  p[2] = double2{f[4], f[5]};
}

struct StageArgs {
  int rhs_stride;       // STAGE 0: length of each rhs_out array
  const int *blk_seg;
  const sv_segk *seg;
  const int64_t *off;
  const int *kcnt;
  const int *col;
  const double *wx, *wy, *wl;
  const double *gsx, *gsy, *lsum;
  const double *G;      // gathered records (stage input)
  const double *reg_in; // register 1 input (stage 1: the state)
  double *reg_out;      // register 1 output
  double *e;            // error accumulators
  double *Uout;         // stage 4 result
  double *Gout;         // stages 1-3: next gather records
  double *rhs_out;      // STAGE 0: six arrays of npfb
  const double *fx;
  const sv_ctrl *ctrl;  // [nseg]
  double *partial;      // stage 4: per-block max error
  int rk1;              // STAGE 0: reference stage counter is 1
  // DI only, nullable: the stage's own gather record [ntot][SV_US] and its
  // derivative values [ntot][16] (rx ry ux uy lu vx vy lv, cxx_x cxx_y, cxy_x
  // cxy_y, cyy_x cyy_y, 0 0), the tape the adjoint sweep reads
  double *tapeG, *tapeD;
  // per-block halos (k_stage_h)
  const int64_t *hoff;
  const int *hidx;
  const uint16_t *lcol;
  int hcap;             // records the dynamic LDS holds
  // boundary nodes (nullable): per loop slot the record index or -1, records
  const int *bidx;
  const double *brec;
  const int *blist;     // launch group's logical blocks (nullable: all blocks)
};

// u_inflow_local of an inflow node as the last apply_time_dependent_bounds
// before stage s of the current step left it (update_u_inflow,
// characteristic_boundaries.F90:256-289): while time <= ramp_time every apply
// sets the ramp value; the applies of this step ran at time0 + RKc[q-1],
// q < s.  s = 1: the value at the start of the step.
__device__ __forceinline__ double uin_now(const sv_physk &P, const sv_ctrl *C, double uin0,
                                          int s) {
  if (P.u_control != 1) return uin0;
  bool set = C->uin_set != 0;
  double v = C->uin_ramp;
  for (int q = 1; q < s; q++) {
    const double t = C->time0 + C->RKc[q - 1];
    if (t <= P.ramp_time) {
      set = true;
      v = P.u_start + (P.u_end - P.u_start) * t / P.ramp_time;
    }
  }
  return set ? v : uin0;
}

// apply_time_dependent_bounds on one state record (characteristic_boundaries.F90:
// 176-254): walls hold zero velocity, hard inflows the inflow velocity
__device__ __forceinline__ void bound_apply(const sv_physk &P, int kind, double uin, double f[6]) {
  if (kind == 0) {
    f[1] = 0.0;
    f[2] = 0.0;
  } else if (kind == 1 && P.inflow_type == 1) {
    f[1] = uin * f[0];
    f[2] = 0.0;
  }
}

// Right-hand side of a boundary node: the boundary branches of calc_rhs_ro
// (rhs_lc.F90:144-160), calc_rhs_rovel (:279-367) and, for the conformation,
// calc_rhs_logconf (:564-745) or calc_rhs_c (rhs_di.F90:477-594: on walls the
// upper-convected and relaxation terms with the rotated velocity gradient, the
// polymer diffusion added by the caller; on inflows/outflows zero), then the
// characteristic waves
// (characteristic_boundaries.F90:27-175) applied as calc_rhs_nscbc
// (rhs_lc.F90:752-813).  Wall derivatives are in the normal-tangent frame
// (x component = d/dn, y = d/dt), inflow/outflow ones in x-y.  In the
// log-conformation form the interior conformation right-hand side (r_xx..r_yy
// on entry) is kept on inflow and outflow nodes.  f1..f3: psi (LC) or c (DI).
template <int FORM>
__device__ __forceinline__ void bound_rhs(const sv_physk &P, int kind, double nx, double ny,
                                          double uin, bool first, double ro, double u, double v,
                                          double rx, double ry, double ux, double uy, double vx,
                                          double vy, double lu, double lv, double dcxx_x,
                                          double dcxy_y, double dcxy_x, double dcyy_y,
                                          double f1, double f2, double f3, double &r_ro,
                                          double &r_rou, double &r_rov, double &r_xx,
                                          double &r_xy, double &r_yy) {
  const double csq = P.csq, c = P.cs;
  const double nscbc = 2.87e-1;                   // common_parameter.F90:71
  const double pr = csq * ro, dpdn = csq * rx;
  double L1, L2, L3, L5;
  if (kind == 0) {
    const double dutdt = -ny * uy + nx * vy;
    r_ro = -ro * dutdt - ro * 0.0;
    const double un = u * nx + v * ny;
    const double dundn = nx * ux + ny * vx;
    const double dutdn = -ny * ux + nx * vx;
    L1 = 0.5 * (un - c) * (dpdn - ro * c * dundn);
    L2 = un * (rx - csq * rx / c / c);
    L3 = un * dutdn;
    L5 = 0.5 * (un + c) * (dpdn + ro * c * dundn);
    r_rou = 0.0;
    r_rov = 0.0;
    // velocity gradient rotated to x-y; no advection on a wall
    const double G11 = nx * ux - ny * uy, G12 = ny * ux + nx * uy;
    const double G21 = nx * vx - ny * vy, G22 = ny * vx + nx * vy;
    if (FORM == SV_FORM_LC) {
      rhs_logconf(0.0, 0.0, G11, G12, G21, G22, f1, f2, f3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                  first, P.eps, P.lambda, r_xx, r_xy, r_yy);
    } else {
      const double cxx = f1, cxy = f2, cyy = f3;
      const double ucxx = 2.0 * G11 * cxx + 2.0 * G12 * cxy + 0.0;
      const double ucxy = G11 * cxy + G12 * cyy + 0.0 + G21 * cxx + G22 * cxy + 0.0;
      const double ucyy = 2.0 * G21 * cxy + 2.0 * G22 * cyy + 0.0;
      const double fr = -(1.0 - P.eps * 2.0 + P.eps * (cxx + cyy + 1.0)) / P.lambda;
      r_xx = ucxx + fr * (cxx - 1.0) + 0.0;
      r_xy = ucxy + fr * cxy + 0.0;
      r_yy = ucyy + fr * (cyy - 1.0) + 0.0;
    }
    L5 = L1;                                      // isothermal wall
  } else {
    r_ro = -v * ry - ro * vy - 0.0 * 0.0 - ro * 0.0;
    L1 = 0.5 * (u - c) * (dpdn - ro * c * ux);
    L2 = u * (rx - csq * rx / c / c);
    L3 = u * vx;
    L5 = 0.5 * (u + c) * (dpdn + ro * c * ux);
    double fvu = P.coef_solvent * lu;
    double fvv = P.coef_solvent * lv;
    fvu = fvu + P.coef_polymeric * (dcxx_x + dcxy_y + 0.0);
    fvv = fvv + P.coef_polymeric * (dcxy_x + dcyy_y + 0.0);
    r_rou = -v * uy - 0.0 * 0.0 + fvu + ro * P.bfx;
    r_rov = -v * vy - 0.0 * 0.0 - csq * ry + fvv + ro * P.bfy;
    if (FORM == SV_FORM_DI) { r_xx = 0.0; r_xy = 0.0; r_yy = 0.0; }   // rhs_di.F90:579-590
    const double Lx = P.Lx;
    if (kind == 1) {
      if (P.inflow_type == 1) {                   // hard inflow
        L2 = 0.0; L3 = 0.0; L5 = L1;
      } else if (P.inflow_type == 2) {            // inflow-outflow
        if (u > 0.0) {
          L2 = 0.0;
          L3 = v * nscbc * c / Lx;
          L5 = (u - uin) * nscbc * (1.0 - u / c) * c * c * 1.0 / Lx;
        } else {
          L2 = 0.0;
          L5 = (pr - P.p_ref) * nscbc * c * 1.0 / 2.0 / Lx;
        }
      } else {                                    // soft inflow
        const double gp2 = csq * ry;
        L2 = 0.0;
        L3 = v * nscbc * c / Lx - v * vy - 0.0 * 0.0 - gp2 / ro;
        L5 = (u - uin) * nscbc * (1.0 - u / c) * c * c * 1.0 / Lx -
             0.5 * (v * gp2 + pr * vy + ro * c * v * uy) -
             0.5 * (0.0 * 0.0 + pr * 0.0 + ro * c * 0.0 * 0.0);
      }
    } else {                                      // outflow
      if (u < c) L1 = (pr - P.p_ref) * nscbc * c * 1.0 / 2.0 / Lx;
      L2 = 0.0;
      if (u <= 0.0) { L2 = 0.0; L3 = 0.0; }
    }
  }
  const double ts = (c * c * L2 + L5 + L1) / c / c;
  r_ro = r_ro - ts;
  r_rou = r_rou - u * ts - (L5 - L1) / c;
  r_rov = r_rov - v * ts - ro * L3;
}

// The neighbour sums of one node in a stage: gradients and Laplacians before
// the own-node term (sum_k w_k f_j), as named in the right-hand side.
struct Sums {
  double rx, ry, ux, uy, vx, vy, lu, lv, axx, ayx, axy, ayy, bxy, byy, cxx_l, cxy_l, cyy_l, q1x, q1y, q2x, q2y, q3x, q3y;
};

// Pointwise part of a stage for one node: the right-hand side from the
// gathered sums and the own record, then the RK update.  Returns the node's
// error norm contribution (stage 4).
template <int FORM, int STAGE, bool MDIFF, bool BND>
__device__ __forceinline__ double stage_point(const StageArgs &a, const NodeRef &nr,
                                              const Sums &S) {
  const int i = nr.i;
  double rx = S.rx, ry = S.ry, ux = S.ux, uy = S.uy, vx = S.vx, vy = S.vy, lu = S.lu, lv = S.lv;
  double axx = S.axx, ayx = S.ayx, axy = S.axy, ayy = S.ayy, bxy = S.bxy, byy = S.byy, cxx_l = S.cxx_l, cxy_l = S.cxy_l;
  double cyy_l = S.cyy_l, q1x = S.q1x, q1y = S.q1y, q2x = S.q2x, q2y = S.q2y, q3x = S.q3x, q3y = S.q3y;
  (void)ayx; (void)bxy; (void)cxx_l; (void)cxy_l; (void)cyy_l;
  (void)q1x; (void)q1y; (void)q2x; (void)q2y; (void)q3x; (void)q3y;
  double emax_loc = 0.0;
  {
    const double sx = a.gsx[i], sy = a.gsy[i], sl = a.lsum[i];
    const Rec9 s = load_g<FORM>(a.G, i);
    const double ro = s.ro, u = s.u, v = s.v, cxx = s.cxx, cxy = s.cxy, cyy = s.cyy;
    rx = rx - ro * sx; ry = ry - ro * sy;
    ux = ux - u * sx; uy = uy - u * sy; lu = lu - u * sl;
    vx = vx - v * sx; vy = vy - v * sy; lv = lv - v * sl;
    const sv_physk &P = a.seg[nr.sg].pk;
    const sv_ctrl *C = a.ctrl + nr.sg;
    // calc_rhs_ro (rhs_lc.F90:127-141)
    const double r_ro = -ro * (ux + vy + 0.0) - (u * rx + v * ry + 0.0 * 0.0);
    int bkind = -1;        // boundary node kind (BND), -1 interior
    double buin0 = 0.0;
    // calc_rhs_rovel (rhs_lc.F90:217-272)
    const double tsu = ro * (u * ux + v * uy + 0.0 * 0.0) - u * r_ro;
    const double tsv = ro * (u * vx + v * vy + 0.0 * 0.0) - v * r_ro;
    const double dcxx_x = axx - cxx * sx;
    const double dcxy_x = axy - cxy * sx;
    const double dcxy_y = ayy - cxy * sy;
    const double dcyy_y = byy - cyy * sy;
    double fvu = P.coef_solvent * lu;
    double fvv = P.coef_solvent * lv;
    fvu = fvu + P.coef_polymeric * (dcxx_x + dcxy_y + 0.0);
    fvv = fvv + P.coef_polymeric * (dcxy_x + dcyy_y + 0.0);
    double bfu = ro * P.bfx;
    const double bfv = ro * P.bfy;
    if (a.fx) bfu = bfu + a.fx[i];
    double r_ro_ = r_ro;
    double r_rou = -tsu - P.csq * rx + bfu + fvu;
    double r_rov = -tsv - P.csq * ry + bfv + fvv;
    double r_xx, r_xy, r_yy;
    if (FORM == SV_FORM_LC) {
      const bool first = (STAGE == 1 || (STAGE == 0 && a.rk1)) && (C->itime == 1);
      rhs_logconf(u, v, ux, uy, vx, vy, s.pxx, s.pxy, s.pyy, q1x - s.pxx * sx,
                  q1y - s.pxx * sy, q2x - s.pxy * sx, q2y - s.pxy * sy, q3x - s.pyy * sx,
                  q3y - s.pyy * sy, first, P.eps, P.lambda, r_xx, r_xy, r_yy);
    } else {
      // calc_rhs_c (rhs_di.F90:417-475) with full gradients of cxx, cyy
      const double gcxx_y = ayx - cxx * sy, gcyy_x = bxy - cyy * sx;
      const double adxx = -(u * dcxx_x + v * gcxx_y + 0.0);
      const double adxy = -(u * dcxy_x + v * dcxy_y + 0.0);
      const double adyy = -(u * gcyy_x + v * dcyy_y + 0.0);
      const double ucxx = 2.0 * ux * cxx + 2.0 * uy * cxy + 0.0;
      const double ucxy = ux * cxy + uy * cyy + 0.0 + vx * cxx + vy * cxy + 0.0;
      const double ucyy = 2.0 * vx * cxy + 2.0 * vy * cyy + 0.0;
      const double fr = -(1.0 - P.eps * 2.0 + P.eps * (cxx + cyy + 1.0)) / P.lambda;
      double mxx = 0.0, mxy = 0.0, myy = 0.0;
      if (MDIFF) {
        mxx = P.Mdiff * (cxx_l - cxx * sl);
        mxy = P.Mdiff * (cxy_l - cxy * sl);
        myy = P.Mdiff * (cyy_l - cyy * sl);
      }
      r_xx = adxx + ucxx + fr * (cxx - 1.0) + mxx;
      r_xy = adxy + ucxy + fr * cxy + mxy;
      r_yy = adyy + ucyy + fr * (cyy - 1.0) + myy;
      if (a.tapeD) {
        double2 *T = reinterpret_cast<double2 *>(a.tapeD + (size_t)i * SV_TAPE_D);
        T[0] = double2{rx, ry}; T[1] = double2{ux, uy}; T[2] = double2{lu, vx};
        T[3] = double2{vy, lv}; T[4] = double2{dcxx_x, gcxx_y}; T[5] = double2{dcxy_x, dcxy_y};
        T[6] = double2{gcyy_x, dcyy_y}; T[7] = double2{0.0, 0.0};
        double2 *Q = reinterpret_cast<double2 *>(a.tapeG + (size_t)i * SV_US);
        Q[0] = double2{ro, u}; Q[1] = double2{v, cxx}; Q[2] = double2{cxy, cyy};
      }
    }
    if (BND) {
      const int jb = a.bidx[i];
      if (jb >= 0) {
        const double *br = a.brec + (size_t)SV_BREC * jb;
        bkind = (int)br[0];
        buin0 = br[3];
        const int s_rhs = STAGE == 0 ? 1 : STAGE;
        const bool first = (STAGE == 1 || (STAGE == 0 && a.rk1)) && (C->itime == 1);
        const double f1 = FORM == SV_FORM_LC ? s.pxx : cxx;
        const double f2 = FORM == SV_FORM_LC ? s.pxy : cxy;
        const double f3 = FORM == SV_FORM_LC ? s.pyy : cyy;
        bound_rhs<FORM>(P, bkind, br[1], br[2], uin_now(P, C, buin0, s_rhs), first, ro, u, v,
                        rx, ry, ux, uy, vx, vy, lu, lv, dcxx_x, dcxy_y, dcxy_x, dcyy_y, f1, f2,
                        f3, r_ro_, r_rou, r_rov, r_xx, r_xy, r_yy);
        if (FORM == SV_FORM_DI && MDIFF && bkind == 0) {
          // polymer diffusion on a wall (rhs_di.F90:396-413, :528-556): the
          // tangential second derivative of c only
          // (calc_laplacian_transverse_only_on_bound, derivatives.F90:329-339),
          // i.e. the wall Laplacian less its normal part ij_wb_grad2(1,:)
          // (labf.F90:594-598 plus the curvature term of :1011), plus the
          // one-sided term that removes the diffusive flux through the wall;
          // both over the four FD rows along the normal
          double fr[4][3];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const Rec9 q = load_g<FORM>(a.G, (int)br[4 + k]);
            fr[k][0] = q.cxx; fr[k][1] = q.cxy; fr[k][2] = q.cyy;
          }
          const double dx = br[8], ooR = br[9], Lc = br[10];
          const double dx2 = dx * dx;
          const double wn[4] = {(-104.0 / 12.0 / dx2 + 4.0 / dx * ooR) / Lc / Lc,
                                (114.0 / 12.0 / dx2 + -3.0 / dx * ooR) / Lc / Lc,
                                (-56.0 / 12.0 / dx2 + (4.0 / 3.0) / dx * ooR) / Lc / Lc,
                                (11.0 / 12.0 / dx2 + -0.25 / dx * ooR) / Lc / Lc};
          const double wsum = wn[0] + wn[1] + wn[2] + wn[3];
          const double den = 72.0 * dx * dx * Lc * Lc;
          const double c0[3] = {cxx, cxy, cyy};
          const double lap[3] = {cxx_l - cxx * sl, cxy_l - cxy * sl, cyy_l - cyy * sl};
          double m[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            const double n2 = wn[0] * fr[0][q] + wn[1] * fr[1][q] + wn[2] * fr[2][q] +
                              wn[3] * fr[3][q] - c0[q] * wsum;
            m[q] = (lap[q] - n2) + (-415.0 * c0[q] + 576.0 * fr[0][q] - 216.0 * fr[1][q] +
                                    64.0 * fr[2][q] - 9.0 * fr[3][q]) / den;
          }
          r_xx = r_xx + P.Mdiff * m[0];
          r_xy = r_xy + P.Mdiff * m[1];
          r_yy = r_yy + P.Mdiff * m[2];
        }
      }
    }
    // ---- RK update (step.F90:291-343, :375-409) ----
    const double rhs[6] = {r_ro_, r_rou, r_rov, r_xx, r_xy, r_yy};
    if (STAGE == 0) {
#pragma unroll
      for (int f = 0; f < 6; f++) a.rhs_out[(size_t)f * a.rhs_stride + i] = rhs[f];
    } else if (STAGE <= 3) {
      const double ra = C->RKa[STAGE - 1], rb = C->RKb[STAGE - 1], rm = C->RKbmbh[STAGE - 1];
      double r1[6], eo[6], Un[6], ro_[6];
      load_u(a.reg_in, i, r1);
      if (STAGE > 1) load_u(a.e, i, eo);
#pragma unroll
      for (int f = 0; f < 6; f++) {
        Un[f] = r1[f] + ra * rhs[f];
        ro_[f] = r1[f] + rb * rhs[f];
        eo[f] = (STAGE == 1 ? 0.0 : eo[f]) + rm * rhs[f];
      }
      store_u(a.reg_out, i, ro_);
      store_u(a.e, i, eo);
      if (BND && bkind >= 0) bound_apply(P, bkind, uin_now(P, C, buin0, STAGE + 1), Un);
      // derived fields for the next stage (step.F90:350-356)
      Rec9 g;
      const double ooro = 1.0 / Un[0];
      g.ro = Un[0];
      g.u = Un[1] * ooro;
      g.v = Un[2] * ooro;
      if (FORM == SV_FORM_LC) {
        g.pxx = Un[3]; g.pxy = Un[4]; g.pyy = Un[5];
        c_from_psi(Un[3], Un[4], Un[5], g.cxx, g.cxy, g.cyy);
      } else {
        g.cxx = Un[3]; g.cxy = Un[4]; g.cyy = Un[5];
        g.pxx = g.pxy = g.pyy = 0.0;
      }
      store_g<FORM>(a.Gout, i, g);
    } else {
      const double rb = C->RKb[3], rm = C->RKbmbh[3];
      double r1[6], eo[6], uo[6];
      load_u(a.reg_in, i, r1);
      load_u(a.e, i, eo);
      const double nrm[6] = {1.0, P.err_norm_rou, P.err_norm_rou, 1.0, 1.0, 1.0};
#pragma unroll
      for (int f = 0; f < 6; f++) {
        uo[f] = r1[f] + rb * rhs[f];
        const double ef = eo[f] + rm * rhs[f];
        emax_loc = fmax(emax_loc, fabs(ef) * nrm[f]);
      }
      if (BND && bkind >= 0) bound_apply(P, bkind, uin_now(P, C, buin0, 5), uo);
      store_u(a.Uout, i, uo);
    }
  }
  return emax_loc;
}

// STAGE 1..4: RK stages; STAGE 0: write the right-hand side into rhs_out.
template <int FORM, int STAGE, bool MDIFF, bool BND>
__global__ __launch_bounds__(BS) void k_stage(StageArgs a) {
  __shared__ double red[BS / SV_WAVE];
  const int bid = sv_logical_block(a.blist);
  const NodeRef nr = node_ref(a.blk_seg, a.seg, bid);
  const int ii = nr.ii;
  const bool act = nr.act;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  double rx = 0, ry = 0, ux = 0, uy = 0, vx = 0, vy = 0, lu = 0, lv = 0;
  double axx = 0, ayx = 0, axy = 0, ayy = 0, bxy = 0, byy = 0;
  double cxx_l = 0, cxy_l = 0, cyy_l = 0;
  double q1x = 0, q1y = 0, q2x = 0, q2y = 0, q3x = 0, q3y = 0;
  auto accum = [&](const Rec9 &r, double wxk, double wyk, double wlk) {
    rx += r.ro * wxk; ry += r.ro * wyk;
    ux += r.u * wxk; uy += r.u * wyk; lu += r.u * wlk;
    vx += r.v * wxk; vy += r.v * wyk; lv += r.v * wlk;
    if (FORM == SV_FORM_LC) {
      axx += r.cxx * wxk;                         // calc_gradient_only(1,cxx)
      axy += r.cxy * wxk; ayy += r.cxy * wyk;     // calc_gradient(cxy)
      byy += r.cyy * wyk;                         // calc_gradient_only(2,cyy)
      q1x += r.pxx * wxk; q1y += r.pxx * wyk;
      q2x += r.pxy * wxk; q2y += r.pxy * wyk;
      q3x += r.pyy * wxk; q3y += r.pyy * wyk;
    } else {
      axx += r.cxx * wxk; ayx += r.cxx * wyk;
      axy += r.cxy * wxk; ayy += r.cxy * wyk;
      bxy += r.cyy * wxk; byy += r.cyy * wyk;
      if (MDIFF) { cxx_l += r.cxx * wlk; cxy_l += r.cxy * wlk; cyy_l += r.cyy * wlk; }
    }
  };
  // Links in pairs, the column indices of the next pair loaded before the
  // current pair is consumed, so each lane keeps two record gathers and the
  // next index loads in flight.  Summation stays sequential in k.
  constexpr int U = SV_GATHER_UNROLL;
  int jn[U];
#pragma unroll
  for (int u = 0; u < U; u++) jn[u] = (u < K) ? ld_stream(a.col + b + (int64_t)u * SV_WAVE) : ii;
  for (int k = 0; k < K; k += U) {
    Rec9 r[U];
    double wxk[U], wyk[U], wlk[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k + u < K) {
        const int64_t e = b + (int64_t)(k + u) * SV_WAVE;
        r[u] = load_g<FORM>(a.G, jn[u]);
        wxk[u] = ld_stream(a.wx + e); wyk[u] = ld_stream(a.wy + e); wlk[u] = ld_stream(a.wl + e);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int kn = k + U + u;
      jn[u] = (kn < K) ? ld_stream(a.col + b + (int64_t)kn * SV_WAVE) : ii;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k + u < K) accum(r[u], wxk[u], wyk[u], wlk[u]);
  }
  double emax_loc = 0.0;
  if (act) {
    const Sums S{rx, ry, ux, uy, vx, vy, lu, lv, axx, ayx, axy, ayy, bxy, byy, cxx_l, cxy_l, cyy_l, q1x, q1y, q2x, q2y, q3x, q3y};
    emax_loc = stage_point<FORM, STAGE, MDIFF, BND>(a, nr, S);
  }
  if (STAGE == 4) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) emax_loc = fmax(emax_loc, __shfl_xor(emax_loc, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = emax_loc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int w = 1; w < BS / SV_WAVE; w++) m = fmax(m, red[w]);
      a.partial[bid] = m;
    }
  }
}

// ---- halo-staged sweeps ---------------------------------------------------
// A block first copies the gather records of its halo (the distinct neighbours
// of its 256 nodes, ~2.6 per node on a Morton-ordered lattice) into LDS, then
// every link reads its record from LDS by a 16-bit block-local index.  The
// L1 then serves only the streamed arrays (indices, weights) and the halo copy,
// instead of one scattered record per link.
// LDS record layout: 8 doubles (ro u v cxx cxy cyy pxx pxy) + pyy in a
// separate array for LC (72 bytes per record); 6 doubles for DI.
template <int FORM>
struct HaloL {
  static constexpr int RS = FORM == SV_FORM_LC ? 8 : 6;   // doubles in the main record
  static constexpr int BYTES = FORM == SV_FORM_LC ? 72 : 48;
};

template <int FORM>
__device__ __forceinline__ void halo_stage(const double *G, const int *hidx, int64_t h0, int nh,
                                           double *hl, double *hz) {
  constexpr int RS = HaloL<FORM>::RS;
  for (int h = threadIdx.x; h < nh; h += BS) {
    const int key = hidx[h0 + h];
    Rec9 r = load_g<FORM>(G, key & SV_FLIP_MASK);
    const int fl = key >> SV_FLIP_SHIFT;
    if (fl) {   // mirror ghost (reapply_mirror_bcs, mirror_boundaries.F90:285-340)
      if (fl != 2) r.u = -r.u;
      if (fl != 1) r.v = -r.v;
      if (fl != 3) { r.cxy = -r.cxy; r.pxy = -r.pxy; }
    }
    double2 *d = reinterpret_cast<double2 *>(hl + (size_t)h * RS);
    d[0] = double2{r.ro, r.u};
    d[1] = double2{r.v, r.cxx};
    d[2] = double2{r.cxy, r.cyy};
    if (FORM == SV_FORM_LC) {
      d[3] = double2{r.pxx, r.pxy};
      hz[h] = r.pyy;
    }
  }
}

template <int FORM>
__device__ __forceinline__ Rec9 halo_rec(const double *hl, const double *hz, int h) {
  constexpr int RS = HaloL<FORM>::RS;
  const double2 *d = reinterpret_cast<const double2 *>(hl + (size_t)h * RS);
  Rec9 r;
  const double2 a = d[0], b = d[1], c = d[2];
  r.ro = a.x; r.u = a.y; r.v = b.x; r.cxx = b.y; r.cxy = c.x; r.cyy = c.y;
  if (FORM == SV_FORM_LC) {
    const double2 e = d[3];
    r.pxx = e.x; r.pxy = e.y; r.pyy = hz[h];
  } else {
    r.pxx = r.pxy = r.pyy = 0.0;
  }
  return r;
}

template <int FORM, int STAGE, bool MDIFF, bool BND>
__global__ __launch_bounds__(BS) void k_stage_h(StageArgs a) {
  extern __shared__ double hl[];
  __shared__ double red[BS / SV_WAVE];
  double *hz = hl + (size_t)a.hcap * HaloL<FORM>::RS;
  const int bid = sv_logical_block(a.blist);
  const NodeRef nr = node_ref(a.blk_seg, a.seg, bid);
  const int ii = nr.ii;
  const bool act = nr.act;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  double rx = 0, ry = 0, ux = 0, uy = 0, vx = 0, vy = 0, lu = 0, lv = 0;
  double axx = 0, ayx = 0, axy = 0, ayy = 0, bxy = 0, byy = 0;
  double cxx_l = 0, cxy_l = 0, cyy_l = 0;
  double q1x = 0, q1y = 0, q2x = 0, q2y = 0, q3x = 0, q3y = 0;
  // The global streams (block-local column, three weights) of the next batch
  // of links are loaded before the current batch is consumed, so each lane
  // keeps SV_HALO_UNROLL links (26 bytes each) in flight; records come from LDS.
  // The first batch is requested before the halo is staged.
  constexpr int U = SV_HALO_UNROLL;
  int jn[U];
  double wxn[U], wyn[U], wln[U];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u < K) {
        const int64_t e = b + (int64_t)(k0 + u) * SV_WAVE;
        jn[u] = (int)ld_stream(a.lcol + e);
        wxn[u] = ld_stream(a.wx + e); wyn[u] = ld_stream(a.wy + e); wln[u] = ld_stream(a.wl + e);
      }
    }
  };
  fetch(0);
  const int64_t h0 = a.hoff[bid];
  halo_stage<FORM>(a.G, a.hidx, h0, (int)(a.hoff[bid + 1] - h0), hl, hz);
  __syncthreads();
  for (int k = 0; k < K; k += U) {
    int jc[U];
    double wxc[U], wyc[U], wlc[U];
#pragma unroll
    for (int u = 0; u < U; u++) { jc[u] = jn[u]; wxc[u] = wxn[u]; wyc[u] = wyn[u]; wlc[u] = wln[u]; }
    fetch(k + U);
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k + u < K) {
        const Rec9 q = halo_rec<FORM>(hl, hz, jc[u]);
        const double wx_ = wxc[u], wy_ = wyc[u], wl_ = wlc[u];
        rx += q.ro * wx_; ry += q.ro * wy_;
        ux += q.u * wx_; uy += q.u * wy_; lu += q.u * wl_;
        vx += q.v * wx_; vy += q.v * wy_; lv += q.v * wl_;
        if (FORM == SV_FORM_LC) {
          axx += q.cxx * wx_;
          axy += q.cxy * wx_; ayy += q.cxy * wy_;
          byy += q.cyy * wy_;
          q1x += q.pxx * wx_; q1y += q.pxx * wy_;
          q2x += q.pxy * wx_; q2y += q.pxy * wy_;
          q3x += q.pyy * wx_; q3y += q.pyy * wy_;
        } else {
          axx += q.cxx * wx_; ayx += q.cxx * wy_;
          axy += q.cxy * wx_; ayy += q.cxy * wy_;
          bxy += q.cyy * wx_; byy += q.cyy * wy_;
          if (MDIFF) { cxx_l += q.cxx * wl_; cxy_l += q.cxy * wl_; cyy_l += q.cyy * wl_; }
        }
      }
  }
  double emax_loc = 0.0;
  if (act) {
    const Sums S{rx, ry, ux, uy, vx, vy, lu, lv, axx, ayx, axy, ayy, bxy, byy, cxx_l, cxy_l,
                 cyy_l, q1x, q1y, q2x, q2y, q3x, q3y};
    emax_loc = stage_point<FORM, STAGE, MDIFF, BND>(a, nr, S);
  }
  if (STAGE == 4) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) emax_loc = fmax(emax_loc, __shfl_xor(emax_loc, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = emax_loc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int w = 1; w < BS / SV_WAVE; w++) m = fmax(m, red[w]);
      a.partial[bid] = m;
    }
  }
}

struct FilterArgs {
  const int *blk_seg;
  const sv_segk *seg;
  const int64_t *off;
  const int *kcnt;
  const int *col;
  const double *wh, *hsum, *vol;
  const double *U;   // gathered state records
  double *Uf;        // filtered state records
  double *Gout;      // psi (LC) and c of the filtered state
  double *partial;
  const int64_t *hoff;   // per-block halos (k_filter_h)
  const int *hidx;
  const uint16_t *lcol;   // per-block sum of (ro_f - rho_char)*vol
  const int *blist;       // launch group's logical blocks (nullable: all blocks)
};

// Pointwise part of the filter for one node: filtered record, its gather
// record (psi and c = exp(psi) for LC), and the mass-correction term.
template <int FORM>
__device__ __forceinline__ double filter_point(const FilterArgs &a, const NodeRef &nr,
                                               const double t[6]) {
  const int i = nr.i;
  const double hs = a.hsum[i];
  double f[6], g[6];
  load_u(a.U, i, f);
#pragma unroll
  for (int q = 0; q < 6; q++) g[q] = f[q] + (t[q] - f[q] * hs);
  store_u(a.Uf, i, g);
  constexpr int S = (FORM == SV_FORM_LC) ? SV_GS_LC : SV_GS_DI;
  double *G = a.Gout + (size_t)i * S;
  double c1 = g[3], c2 = g[4], c3 = g[5];
  if (FORM == SV_FORM_LC) {
    G[SV_G_PXX] = g[3]; G[SV_G_PXY] = g[4]; G[SV_G_PYY] = g[5];
    c_from_psi(g[3], g[4], g[5], c1, c2, c3);
  }
  G[SV_G_CXX] = c1; G[SV_G_CXY] = c2; G[SV_G_CYY] = c3;
  return (g[0] - a.seg[nr.sg].pk.rho_char) * a.vol[i];
}

template <int FORM>
__global__ __launch_bounds__(BS) void k_filter(FilterArgs a) {
  __shared__ double red[BS];
  const int bid = sv_logical_block(a.blist);
  const NodeRef nr = node_ref(a.blk_seg, a.seg, bid);
  const int ii = nr.ii;
  const bool act = nr.act;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  double t[6] = {0, 0, 0, 0, 0, 0};
  // pairs of links with the next pair's column indices prefetched, as k_stage
  constexpr int U = SV_GATHER_UNROLL;
  int jn[U];
#pragma unroll
  for (int u = 0; u < U; u++) jn[u] = (u < K) ? ld_stream(a.col + b + (int64_t)u * SV_WAVE) : ii;
  for (int k = 0; k < K; k += U) {
    double f[U][6], w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k + u < K) {
        load_u(a.U, jn[u], f[u]);
        w[u] = ld_stream(a.wh + b + (int64_t)(k + u) * SV_WAVE);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int kn = k + U + u;
      jn[u] = (kn < K) ? ld_stream(a.col + b + (int64_t)kn * SV_WAVE) : ii;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k + u < K) {
#pragma unroll
        for (int q = 0; q < 6; q++) t[q] += f[u][q] * w[u];
      }
  }
  const double m = act ? filter_point<FORM>(a, nr, t) : 0.0;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partial[bid] = red[0];
}

// Halo-staged form of k_filter (see k_stage_h): the block's neighbour state
// records (6 doubles) go to LDS once, links read them by block-local index.
template <int FORM>
__global__ __launch_bounds__(BS) void k_filter_h(FilterArgs a) {
  extern __shared__ double hl[];
  __shared__ double red[BS];
  const int bid = sv_logical_block(a.blist);
  const NodeRef nr = node_ref(a.blk_seg, a.seg, bid);
  const int ii = nr.ii;
  const bool act = nr.act;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  double t[6] = {0, 0, 0, 0, 0, 0};
  constexpr int U = SV_HALO_UNROLL;
  int jn[U];
  double wn[U];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u < K) {
        const int64_t e = b + (int64_t)(k0 + u) * SV_WAVE;
        jn[u] = (int)ld_stream(a.lcol + e);
        wn[u] = ld_stream(a.wh + e);
      }
  };
  fetch(0);   // the first batch is in flight while the halo is staged
  const int64_t h0 = a.hoff[bid];
  const int nh = (int)(a.hoff[bid + 1] - h0);
  for (int h = threadIdx.x; h < nh; h += BS) {
    double f[6];
    const int key = a.hidx[h0 + h];
    load_u(a.U, key & SV_FLIP_MASK, f);
    const int fl = key >> SV_FLIP_SHIFT;
    if (fl) {   // mirror ghost: rho u, rho v and the xy component
      if (fl != 2) f[1] = -f[1];
      if (fl != 1) f[2] = -f[2];
      if (fl != 3) f[4] = -f[4];
    }
    double2 *d = reinterpret_cast<double2 *>(hl + (size_t)h * 6);
    d[0] = double2{f[0], f[1]};
    d[1] = double2{f[2], f[3]};
    d[2] = double2{f[4], f[5]};
  }
  __syncthreads();
  for (int k = 0; k < K; k += U) {
    int jc[U];
    double wc[U];
#pragma unroll
    for (int u = 0; u < U; u++) { jc[u] = jn[u]; wc[u] = wn[u]; }
    fetch(k + U);
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k + u < K) {
        const double2 *d = reinterpret_cast<const double2 *>(hl + (size_t)jc[u] * 6);
        const double2 d0 = d[0], d1 = d[1], d2 = d[2];
        const double w = wc[u];
        t[0] += d0.x * w; t[1] += d0.y * w; t[2] += d1.x * w;
        t[3] += d1.y * w; t[4] += d2.x * w; t[5] += d2.y * w;
      }
  }
  const double m = act ? filter_point<FORM>(a, nr, t) : 0.0;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partial[bid] = red[0];
}

// mass correction value, emax and the end-of-step counters; one block per
// segment over that segment's per-block partials
__global__ __launch_bounds__(1024) void k_reduce(const sv_segk *seg, const double *pmass,
                                                 const double *perr, sv_ctrl *ctrl_all) {
  __shared__ double s1[1024], s2[1024];
  const sv_segk &S = seg[blockIdx.x];
  sv_ctrl *ctrl = ctrl_all + blockIdx.x;
  const double volsum = S.volsum;
  double a = 0.0, m = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += 1024) {
    a += pmass[S.blk0 + k];
    m = fmax(m, perr[S.blk0 + k]);
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = m;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctrl->dro = s1[0] / volsum;
    ctrl->emax_np1 = fmax(s2[0], 1.0e-16);       // step.F90:480-485
    if (!(s2[0] == s2[0]) || !(s1[0] == s1[0])) ctrl->flags |= 1;
    // the step's four applies of the boundary conditions ran at
    // time0 + RKc[0..3]; the ramped inflow velocity follows the last one at
    // time <= ramp_time (characteristic_boundaries.F90:268-283)
    if (S.pk.u_control == 1)
      for (int q = 0; q < 4; q++) {
        const double t = ctrl->time0 + ctrl->RKc[q];
        if (t <= S.pk.ramp_time) {
          ctrl->uin_set = 1;
          ctrl->uin_ramp = S.pk.u_start + (S.pk.u_end - S.pk.u_start) * t / S.pk.ramp_time;
        }
      }
    ctrl->time = ctrl->time0 + ctrl->RKc[3];
    ctrl->itime = ctrl->itime + 1;               // sunset.F90:99
  }
}

struct ApplyArgs {
  int gs;
  const int *blk_seg;
  const sv_segk *seg;
  const double *Uf;
  double *U;
  double *G;
  const double *s;
  const sv_ctrl *ctrl;   // ctrl[sg].dro is subtracted; nullptr for no correction
  double *pcfl, *pvisc;  // per-block minima for set_tstep
  const int *bidx;       // boundary nodes (nullable), as StageArgs
  const double *brec;
};

// rho -= drho (rhs_lc.F90:887-889); u = rou/rho (step.F90:514-529); set_tstep
// partial minima (step.F90:579-602)
__global__ __launch_bounds__(BS) void k_apply_state(ApplyArgs a) {
  __shared__ double r1[BS], r2[BS];
  const int bid = sv_xcd_block();
  const NodeRef nr = node_ref(a.blk_seg, a.seg, bid);
  const int i = nr.i;
  double mc = 1.0e10, mv = 1.0e10;
  if (nr.act) {
    const sv_physk &P = a.seg[nr.sg].pk;
    const double dro = a.ctrl ? a.ctrl[nr.sg].dro : 0.0;
    double f[6];
    load_u(a.Uf, i, f);
    f[0] = f[0] - dro;
    int bk = -1;
    if (a.bidx) {
      const int jb = a.bidx[i];
      if (jb >= 0) {
        const double *br = a.brec + (size_t)SV_BREC * jb;
        bk = (int)br[0];
        // the apply after filter_variables (step.F90:498)
        if (a.ctrl) bound_apply(P, bk, uin_now(P, a.ctrl + nr.sg, br[3], 1), f);
      }
    }
    store_u(a.U, i, f);
    const double ooro = 1.0 / f[0];
    const double u = f[1] * ooro, v = f[2] * ooro;
    double *G = a.G + (size_t)i * a.gs;
    reinterpret_cast<double2 *>(G)[0] = double2{f[0], u};
    G[SV_G_V] = v;
    const double um = sqrt(u * u + v * v + 0.0);
    const double si = a.s[i];
    // slightly smaller on outflows (step.F90:591-595)
    mc = bk == 2 ? 0.8 * si / (um + P.cs) : si / (um + P.cs);
    mv = si * si * f[0] / P.visc_total;
  }
  r1[threadIdx.x] = mc;
  r2[threadIdx.x] = mv;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r1[threadIdx.x] = fmin(r1[threadIdx.x], r1[threadIdx.x + s]);
      r2[threadIdx.x] = fmin(r2[threadIdx.x], r2[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { a.pcfl[bid] = r1[0]; a.pvisc[bid] = r2[0]; }
}

// set_tstep (step.F90:569-622) + set_tstep_PID (:624-682) + the start of
// step_rk3_4S_2R_EE (:239-246)
// one block per segment
__global__ __launch_bounds__(1024) void k_dt(const sv_segk *seg, const double *pcfl,
                                             const double *pvisc, sv_ctrl *ctrl_all) {
  __shared__ double s1[1024], s2[1024];
  const sv_segk &S = seg[blockIdx.x];
  sv_ctrl *c = ctrl_all + blockIdx.x;
  const double L_char = S.pk.L_char;
  double a = 1.0e10, m = 1.0e10;
  for (int k = threadIdx.x; k < S.nblk; k += 1024) {
    a = fmin(a, pcfl[S.blk0 + k]);
    m = fmin(m, pvisc[S.blk0 + k]);
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = m;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] = fmin(s1[threadIdx.x], s1[threadIdx.x + s]);
      s2[threadIdx.x] = fmin(s2[threadIdx.x], s2[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  c->dt_cfl = 1.0 * s1[0] * L_char;
  c->dt_par = (2.0 / 3.0) * s2[0] * L_char * L_char;
  if (c->dt_mode == SV_DT_PID) {
    const double fA = SV_PID_A * log(SV_PID_TOL / c->emax_np1);
    const double fB = SV_PID_B * log(c->emax_n / SV_PID_TOL);
    const double fC = SV_PID_C * log(SV_PID_TOL / c->emax_nm1);
    double f = SV_PID_K * exp(fA + fB + fC);
    f = fmax(0.5, f);
    f = fmin(1.02, f);
    double dt = c->dt * f;
    const double dmax = fmin(c->dt_cfl, c->dt_par);
    if (dt > dmax) dt = dmax;
    c->dt = dt;
  } else if (c->dt_mode == SV_DT_CFL) {
    c->dt = fmin(c->dt_par, c->dt_cfl);
  }
  c->emax_nm1 = c->emax_n;
  c->emax_n = c->emax_np1;
  const double dt = c->dt;
  c->RKa[0] = dt * SV_RKA1; c->RKa[1] = dt * SV_RKA2; c->RKa[2] = dt * SV_RKA3;
  const double b[4] = {SV_RKB1, SV_RKB2, SV_RKB3, SV_RKB4};
  const double bh[4] = {SV_RKBH1, SV_RKBH2, SV_RKBH3, SV_RKBH4};
  const double cc[4] = {0.0, SV_RKA1, SV_RKA2 + SV_RKB1, 1.0};
  for (int k = 0; k < 4; k++) {
    c->RKb[k] = dt * b[k];
    c->RKbmbh[k] = dt * (b[k] - bh[k]);
    c->RKc[k] = dt * cc[k];
  }
  c->time0 = c->time;
}

// initial gather records: psi (LC) and c of the state (ro, u, v by k_apply_state)
template <int FORM>
__global__ __launch_bounds__(BS) void k_init_gather(const int *blk_seg, const sv_segk *seg,
                                                     const double *U, double *G) {
  const NodeRef nr = node_ref(blk_seg, seg, (int)blockIdx.x);
  if (!nr.act) return;
  const int i = nr.i;
  double f[6];
  load_u(U, i, f);
  constexpr int S = (FORM == SV_FORM_LC) ? SV_GS_LC : SV_GS_DI;
  double *g = G + (size_t)i * S;
  double c1 = f[3], c2 = f[4], c3 = f[5];
  if (FORM == SV_FORM_LC) {
    g[SV_G_PXX] = f[3]; g[SV_G_PXY] = f[4]; g[SV_G_PYY] = f[5];
    c_from_psi(f[3], f[4], f[5], c1, c2, c3);
  }
  g[SV_G_CXX] = c1; g[SV_G_CXY] = c2; g[SV_G_CYY] = c3;
}

StageArgs stage_args(sv_ctx *c, int stage, const double *Gin, double *Gout) {
  StageArgs a{};
  a.rhs_stride = c->npfb;
  a.blk_seg = c->blk_seg;
  a.seg = c->segk;
  a.off = c->L.off;
  a.kcnt = c->L.cnt_small;
  a.col = c->L.col;
  a.wx = c->L.wx; a.wy = c->L.wy; a.wl = c->L.wl;
  a.gsx = c->L.gsx; a.gsy = c->L.gsy; a.lsum = c->L.lsum;
  a.G = Gin;
  a.reg_in = (stage == 1) ? c->U : c->reg;
  a.reg_out = c->reg;
  a.e = c->eacc;
  a.Uout = c->U;
  a.Gout = Gout;
  a.fx = c->any_forcing ? c->fx : nullptr;
  a.ctrl = c->ctrl;
  a.hoff = c->L.hoff;
  a.hidx = c->L.hidx;
  a.lcol = c->L.lcol;
  a.hcap = c->L.hmax;
  a.partial = c->partial;
  a.bidx = c->L.bidx;
  a.brec = c->L.brec;
  return a;
}

template <int FORM, int STAGE>
void launch_stage_grid(sv_ctx *c, const StageArgs &a, int nb, int hmax) {
  const bool md = FORM == SV_FORM_DI && c->any_mdiff;
  if constexpr (FORM == SV_FORM_LC) {
    const bool bnd = a.bidx != nullptr;
    if (c->L.halo) {
      const size_t shm = (size_t)hmax * HaloL<FORM>::BYTES;
      if (bnd) k_stage_h<FORM, STAGE, false, true><<<nb, BS, shm, c->stream>>>(a);
      else k_stage_h<FORM, STAGE, false, false><<<nb, BS, shm, c->stream>>>(a);
    } else {
      if (bnd) k_stage<FORM, STAGE, false, true><<<nb, BS, 0, c->stream>>>(a);
      else k_stage<FORM, STAGE, false, false><<<nb, BS, 0, c->stream>>>(a);
    }
  } else {
    const bool bnd = a.bidx != nullptr;
    if (c->L.halo) {
      const size_t shm = (size_t)hmax * HaloL<FORM>::BYTES;
      if (md && bnd) k_stage_h<FORM, STAGE, true, true><<<nb, BS, shm, c->stream>>>(a);
      else if (md) k_stage_h<FORM, STAGE, true, false><<<nb, BS, shm, c->stream>>>(a);
      else if (bnd) k_stage_h<FORM, STAGE, false, true><<<nb, BS, shm, c->stream>>>(a);
      else k_stage_h<FORM, STAGE, false, false><<<nb, BS, shm, c->stream>>>(a);
    } else {
      if (md && bnd) k_stage<FORM, STAGE, true, true><<<nb, BS, 0, c->stream>>>(a);
      else if (md) k_stage<FORM, STAGE, true, false><<<nb, BS, 0, c->stream>>>(a);
      else if (bnd) k_stage<FORM, STAGE, false, true><<<nb, BS, 0, c->stream>>>(a);
      else k_stage<FORM, STAGE, false, false><<<nb, BS, 0, c->stream>>>(a);
    }
  }
}

// One stage over every block: one launch, or one launch per ensemble launch
// group, each with its own halo budget (its blocks' largest halo).
template <int FORM, int STAGE>
void launch_stage(sv_ctx *c, const StageArgs &a, int nb) {
  if (c->lgroup_mode == 1 && !c->lgroups.empty()) {
    for (const sv_ctx::lgroup &g : c->lgroups) {
      StageArgs b = a;
      b.blist = g.blist;
      b.hcap = g.hmax;
      launch_stage_grid<FORM, STAGE>(c, b, g.nblk, g.hmax);
    }
    return;
  }
  launch_stage_grid<FORM, STAGE>(c, a, nb, c->L.hmax);
}

template <int FORM>
void launch_filter(sv_ctx *c, const FilterArgs &f, int nb) {
  if (!c->L.halo) {
    k_filter<FORM><<<nb, BS, 0, c->stream>>>(f);
    return;
  }
  if (c->lgroup_mode == 1 && !c->lgroups.empty()) {
    for (const sv_ctx::lgroup &g : c->lgroups) {
      FilterArgs b = f;
      b.blist = g.blist;
      k_filter_h<FORM><<<g.nblk, BS, sizeof(double) * 6 * (size_t)g.hmax, c->stream>>>(b);
    }
    return;
  }
  k_filter_h<FORM><<<nb, BS, sizeof(double) * 6 * (size_t)c->L.hmax, c->stream>>>(f);
}

// ev (nullable): SV_NKERNELS+1 events recorded between the launches of one step
template <int FORM>
int step_form(sv_ctx *c, hipEvent_t *ev, double *const *tape) {
  const int nb = c->nblk;
  double *pm = c->partial, *pe = c->partial + nb, *pc = c->partial + 2 * nb,
         *pv = c->partial + 3 * nb;
  int ke = 0;
  auto mark = [&]() {
    if (ev) (void)hipEventRecord(ev[ke++], c->stream);
  };
  mark();
  k_dt<<<c->nseg, 1024, 0, c->stream>>>(c->segk, pc, pv, c->ctrl);
  mark();
  StageArgs as[4] = {stage_args(c, 1, c->GA, c->GB), stage_args(c, 2, c->GB, c->GA),
                     stage_args(c, 3, c->GA, c->GB), stage_args(c, 4, c->GB, c->GA)};
  as[3].partial = pe;
  if (tape)
    for (int k = 0; k < 4; k++) { as[k].tapeG = tape[2 * k]; as[k].tapeD = tape[2 * k + 1]; }
  launch_stage<FORM, 1>(c, as[0], nb);
  mark();
  launch_stage<FORM, 2>(c, as[1], nb);
  mark();
  launch_stage<FORM, 3>(c, as[2], nb);
  mark();
  launch_stage<FORM, 4>(c, as[3], nb);
  mark();
  FilterArgs f{};
  f.blk_seg = c->blk_seg;
  f.seg = c->segk;
  f.off = c->L.off;
  f.kcnt = c->L.cnt;
  f.col = c->L.col;
  f.wh = c->L.wh;
  f.hsum = c->L.hsum;
  f.vol = c->L.vol;
  f.U = c->U;
  f.Uf = c->Uf;
  f.Gout = c->GA;
  f.partial = pm;
  f.hoff = c->L.hoff;
  f.hidx = c->L.hidx;
  f.lcol = c->L.lcol;
  launch_filter<FORM>(c, f, nb);
  mark();
  k_reduce<<<c->nseg, 1024, 0, c->stream>>>(c->segk, pm, pe, c->ctrl);
  mark();
  ApplyArgs ap{};
  ap.gs = c->gs;
  ap.blk_seg = c->blk_seg;
  ap.seg = c->segk;
  ap.Uf = c->Uf;
  ap.U = c->U;
  ap.G = c->GA;
  ap.s = c->L.s;
  ap.ctrl = c->ctrl;
  ap.pcfl = pc;
  ap.pvisc = pv;
  ap.bidx = c->L.bidx;
  ap.brec = c->L.brec;
  k_apply_state<<<nb, BS, 0, c->stream>>>(ap);
  mark();
  SV_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace

int sv_dev_init_partials(sv_ctx *c) {
  // u, v and the set_tstep partial minima of the current state (no correction)
  const int nb = c->nblk;
  ApplyArgs ap{};
  ap.gs = c->gs;
  ap.blk_seg = c->blk_seg;
  ap.seg = c->segk;
  ap.Uf = c->U;
  ap.U = c->U;
  ap.G = c->GA;
  ap.s = c->L.s;
  ap.ctrl = nullptr;
  ap.pcfl = c->partial + 2 * nb;
  ap.pvisc = c->partial + 3 * nb;
  ap.bidx = c->L.bidx;
  ap.brec = c->L.brec;
  k_apply_state<<<nb, BS, 0, c->stream>>>(ap);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_dev_init_state(sv_ctx *c) {
  const int nb = c->nblk;
  if (c->cfg.form == SV_FORM_LC)
    k_init_gather<SV_FORM_LC><<<nb, BS, 0, c->stream>>>(c->blk_seg, c->segk, c->U, c->GA);
  else
    k_init_gather<SV_FORM_DI><<<nb, BS, 0, c->stream>>>(c->blk_seg, c->segk, c->U, c->GA);
  SV_HIP(c, hipGetLastError());
  return sv_dev_init_partials(c);
}

int sv_dev_step(sv_ctx *c, hipEvent_t *ev) {
  return c->cfg.form == SV_FORM_LC ? step_form<SV_FORM_LC>(c, ev, nullptr)
                                   : step_form<SV_FORM_DI>(c, ev, nullptr);
}

int sv_dev_step_tape(sv_ctx *c, double *const tape[8]) {
  if (c->cfg.form != SV_FORM_DI) return sv_fail(c, SV_E_STATE, "the tape is DI-only");
  return step_form<SV_FORM_DI>(c, nullptr, tape);
}

// calc_all_rhs on the current state (interior nodes): the stage kernel in its
// right-hand-side-only form, written into the six arrays of out6 (npfb each).
int sv_dev_rhs(sv_ctx *c, int iRKstep, double *out6, double *tapeG, double *tapeD) {
  const int nb = c->nblk;
  StageArgs a = stage_args(c, 1, c->GA, c->GB);
  a.rk1 = (iRKstep == 1);
  a.rhs_out = out6;
  a.tapeG = tapeG;
  a.tapeD = tapeD;
  if (c->cfg.form == SV_FORM_LC) launch_stage<SV_FORM_LC, 0>(c, a, nb);
  else launch_stage<SV_FORM_DI, 0>(c, a, nb);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Launch-group autotuning of an ensemble: the right-hand-side-only form of the
// stage sweep (it leaves the state untouched) timed over every block, one
// launch against one launch per group, reps times each after a warm-up; the
// faster mode is kept.
int sv_dev_tune_groups(sv_ctx *c, int reps) {
  if (c->lgroups.size() < 2) { c->lgroup_mode = 0; return 0; }
  DevTmp<double> out;
  SV_HIP(c, out.alloc(6 * (size_t)c->ntot));
  EventTmp e0, e1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  float ms[2] = {0.0f, 0.0f};
  for (int mode = 0; mode < 2; mode++) {
    c->lgroup_mode = mode;
    StageArgs a = stage_args(c, 1, c->GA, c->GB);
    a.rhs_stride = c->ntot;
    a.rhs_out = out;
    for (int r = -1; r < reps; r++) {
      if (r == 0) SV_HIP(c, hipEventRecord(e0, c->stream));
      if (c->cfg.form == SV_FORM_LC) launch_stage<SV_FORM_LC, 0>(c, a, c->nblk);
      else launch_stage<SV_FORM_DI, 0>(c, a, c->nblk);
    }
    SV_HIP(c, hipEventRecord(e1, c->stream));
    SV_HIP(c, hipGetLastError());
    SV_HIP(c, hipEventSynchronize(e1));
    SV_HIP(c, hipEventElapsedTime(&ms[mode], e0, e1));
  }
  c->lgroup_ms[0] = ms[0] / reps;
  c->lgroup_ms[1] = ms[1] / reps;
  c->lgroup_mode = ms[1] < ms[0] ? 1 : 0;
  return 0;
}

namespace {
__global__ void k_offset_int(int *dst, int64_t n, int add) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] += add;
}
__global__ void k_offset_i64(int64_t *dst, int64_t n, int64_t add) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] += add;
}
}  // namespace

int sv_dev_offset_int(sv_ctx *c, int *dst, int64_t n, int add) {
  if (n <= 0) return 0;
  k_offset_int<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(dst, n, add);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_offset_i64(sv_ctx *c, int64_t *dst, int64_t n, int64_t add) {
  if (n <= 0) return 0;
  k_offset_i64<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(dst, n, add);
  SV_HIP(c, hipGetLastError());
  return 0;
}
